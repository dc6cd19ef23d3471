"""GPU parity of the batched column-major <-> panel-major conversions
(blasfeo_amd_dpack_batched / dpack_tran_batched / dunpack_batched) against the
oracle's restatement of blasfeo_pack_dmat, blasfeo_pack_tran_dmat and
blasfeo_unpack_dmat (auxiliary/d_aux_lib4.c:1607, :2086, :2238): bit-exact on
every byte of both buffers (sentinels outside the block included), the
use_dA flags, random sizes 1..200, offsets 0..9 and leading dimensions above
the row count, both panel sizes."""
import numpy as np
import pytest

import blasfeo_amd as bfa
import oracle
from blasfeo_amd.device import to_device
from pack_cases import SENTINEL, make_cases, records, descriptors

pytestmark = pytest.mark.gpu


def _run(c, name, cm_init, pm_init):
    import torch
    cm_d = torch.from_numpy(cm_init.copy()).to("cuda")
    pm_d = torch.from_numpy(pm_init.copy()).to("cuda")
    dd = descriptors(c, pm_d.data_ptr())
    dd["use_dA"] = 5
    ddev = to_device(dd)
    rec = to_device(records(c, cm_d.data_ptr(), ddev.data_ptr()))
    maxc = int(c["C"].max())
    print("%s ps=%d batch=%d" % (name, c["ps"], c["batch"]), flush=True)
    bfa.call(name, c["ps"], c["batch"], rec.data_ptr(), (0, maxc, 0))
    torch.cuda.synchronize()
    # the same on the host
    cm_h, pm_h = cm_init.copy(), pm_init.copy()
    dh = descriptors(c, pm_h.ctypes.data)
    dh["use_dA"] = 5
    rh = records(c, cm_h.ctypes.data, dh.ctypes.data)
    getattr(oracle.lib(), "or_%s_batch" % name)(c["ps"], c["batch"], rh.ctypes.data)
    gdesc = ddev.cpu().numpy().view(bfa.DMAT)
    return cm_d.cpu().numpy(), pm_d.cpu().numpy(), gdesc, cm_h, pm_h, dh


@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("name", ["dpack", "dpack_tran"])
def test_pack(gpu, ps, name):
    rng = np.random.default_rng(800 + ps + (name == "dpack_tran"))
    # instances 5..8 have empty blocks: pack leaves their use_dA (5) alone,
    # pack_tran clears it (auxiliary/d_aux_lib4.c:1609, :2090)
    c = make_cases(rng, 40, ps, tran=(name == "dpack_tran"), empties=(5, 6, 7, 8))
    cm = rng.standard_normal(c["cm_len"])
    pm = np.full(c["pm_bytes"] // 8, SENTINEL)
    gcm, gpm, gdesc, rcm, rpm, rdesc = _run(c, name, cm, pm)
    np.testing.assert_array_equal(gpm, rpm)
    np.testing.assert_array_equal(gcm, cm)
    np.testing.assert_array_equal(gdesc["use_dA"], rdesc["use_dA"])
    empty = (c["ms"] == 0) | (c["ns"] == 0)
    assert np.all(gdesc["use_dA"][~empty] == 0)
    assert np.all(gdesc["use_dA"][empty] == (0 if name == "dpack_tran" else 5))


@pytest.mark.parametrize("ps", [4, 8])
def test_unpack(gpu, ps):
    rng = np.random.default_rng(820 + ps)
    c = make_cases(rng, 40, ps)
    cm = np.full(c["cm_len"], SENTINEL)
    pm = rng.standard_normal(c["pm_bytes"] // 8)
    gcm, gpm, gdesc, rcm, rpm, rdesc = _run(c, "dunpack", cm, pm)
    np.testing.assert_array_equal(gcm, rcm)
    np.testing.assert_array_equal(gpm, pm)
    assert np.all(gdesc["use_dA"] == 5)


def test_pack_unpack_roundtrip_large(gpu):
    """A 1024 x 700 block at offset (3, 5): pack then unpack returns it."""
    import torch
    rng = np.random.default_rng(830)
    m, n, ai, aj, lda = 1024, 700, 3, 5, 1030
    A = torch.from_numpy(rng.standard_normal(lda * n)).to("cuda")
    B = torch.full((lda * n,), SENTINEL, dtype=torch.float64, device="cuda")
    d, _, nb = bfa.bulk_layout(np.float64, 4, [ai + m], [aj + n], 0)
    pm = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
    d = bfa.bulk_layout(np.float64, 4, [ai + m], [aj + n], pm.data_ptr())[0]
    dd = to_device(d)
    rec = np.zeros(1, dtype=bfa.DPACK_ARG)
    rec["m"], rec["n"], rec["lda"], rec["ai"], rec["aj"] = m, n, lda, ai, aj
    rec["A"], rec["sA"] = A.data_ptr(), dd.data_ptr()
    print("roundtrip 1024x700", flush=True)
    r1 = to_device(rec)
    bfa.call("dpack", 4, 1, r1.data_ptr(), (0, n, 0))
    rec["A"] = B.data_ptr()
    r2 = to_device(rec)
    bfa.call("dunpack", 4, 1, r2.data_ptr(), (0, n, 0))
    torch.cuda.synchronize()
    a = A.cpu().numpy().reshape(n, lda).T
    b = B.cpu().numpy().reshape(n, lda).T
    np.testing.assert_array_equal(b[:m], a[:m])
    assert np.all(b[m:] == SENTINEL)


def test_pm_dgecp(gpu):
    """blasfeo_amd_pm_dgecp_batched: one batch mixing all four (psA, psB)
    pairs, random sizes up to 200, offsets 0..9; every byte of the
    destination buffer and the use_dA flags against the oracle."""
    import torch
    from pack_cases import pm_layout
    rng = np.random.default_rng(850)
    batch = 48
    ms = rng.integers(1, 201, batch)
    ns = rng.integers(1, 201, batch)
    ais, ajs, bis, bjs = (rng.integers(0, 10, batch) for _ in range(4))
    pa = rng.choice([4, 8], batch)
    pb = rng.choice([4, 8], batch)
    pss = np.concatenate([pa, pb])
    rows = np.concatenate([ais + ms + 1, bis + ms + 2])
    cols = np.concatenate([ajs + ns, bjs + ns + 1])
    _, offs, nb = pm_layout(pss, rows, cols, 0)
    init = rng.standard_normal(nb // 8)
    g = torch.from_numpy(init.copy()).to("cuda")
    dg = pm_layout(pss, rows, cols, g.data_ptr())[0]
    dg["use_dA"] = 2
    ddev = to_device(dg)
    sz = bfa.PM_DMAT.itemsize

    def recs(dbase):
        r = np.zeros(batch, dtype=bfa.PM_DGECP_ARG)
        r["m"], r["n"], r["ai"], r["aj"], r["bi"], r["bj"] = ms, ns, ais, ajs, bis, bjs
        idx = np.arange(batch, dtype=np.uint64)
        r["sA"] = np.uint64(dbase) + idx * sz
        r["sB"] = np.uint64(dbase) + (idx + batch) * sz
        return r
    print("pm_dgecp batch=%d" % batch, flush=True)
    drec = to_device(recs(ddev.data_ptr()))        # kept alive until the launch has run
    bfa.pm_dgecp(batch, drec.data_ptr(), (0, int(ns.max()), 0))
    torch.cuda.synchronize()
    h = init.copy()
    dh = pm_layout(pss, rows, cols, h.ctypes.data)[0]
    dh["use_dA"] = 2
    hrec = recs(dh.ctypes.data)
    oracle.lib().or_pm_dgecp_batch(batch, hrec.ctypes.data)
    np.testing.assert_array_equal(g.cpu().numpy(), h)
    gd = ddev.cpu().numpy().view(bfa.PM_DMAT)
    np.testing.assert_array_equal(gd["use_dA"], dh["use_dA"])
