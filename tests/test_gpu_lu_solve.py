"""GPU parity of the LU-solve pieces against the oracle's restatements of
blasfeo_dtrsm_llnu, blasfeo_dtrsm_lunn (blasfeo_ref/x_blas3_ref.c:572, :920)
and blasfeo_drowpe (auxiliary/d_aux_lib4.c:2767), and their composition after
getrf_rp (examples/example_d_lu_factorization.c:163-171).  The triangles a
solve must not read (llnu: the diagonal and the upper part; lunn: the lower
part) hold large values, so reading them would fail the comparison."""
import numpy as np
import pytest

import blasfeo_amd as bfa
import oracle
from blasfeo_amd.device import DeviceBatch, HostBatch, to_device
from helpers import Batch, run_both, compare_buffers

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("thin", [False, True])
@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("op", ["trsm_llnu", "trsm_lunn"])
def test_trsm_unit_and_upper(gpu, ps, op, thin):
    """thin: at most 4 right-hand sides, m <= 256 (the one-wavefront kernel)."""
    rng = np.random.default_rng(900 + ps + (op == "trsm_lunn") + 5 * thin)
    b = Batch("d", ps, seed=ps + 21)
    recs = []
    for r in range(18):
        m = int(rng.integers(1, 257 if thin else 260)) if r else 256
        n = (int(rng.integers(1, 5)) if thin else int(rng.integers(1, 40))) if r else 4
        ai, bi, di = (int(x) for x in rng.integers(0, 2 * ps, 3))
        aj, bj, dj = (int(x) for x in rng.integers(0, 3, 3))
        if r % 3 == 0:
            ai = aj = 0
        Ad = np.full((ai + m, aj + m), 1e3)
        blk = rng.standard_normal((m, m)) * 0.1
        if op == "trsm_llnu":
            T = np.tril(blk, -1) + 1e3 * np.triu(np.ones((m, m)), 0)     # diagonal not read
        else:
            T = np.triu(blk, 1) + np.diag(1.0 + rng.uniform(0.5, 1.5, m)) + 1e3 * np.tril(np.ones((m, m)), -1)
        Ad[ai:, aj:] = T
        A = b.mat(ai + m, aj + m, "data", Ad)
        B = b.mat(bi + m, bj + n)
        if r % 5 == 4:
            D, di, dj = B, bi, bj
        else:
            D = b.mat(di + m, dj + n, "const", -1.0)
        recs.append(dict(m=m, n=n, ai=ai, aj=aj, bi=bi, bj=bj, di=di, dj=dj,
                         alpha=float(rng.choice([1.0, -0.5, 2.0])), sA=A, sB=B, sD=D))
    buf = b.host()
    f = oracle.fn("d", op)

    def ocall(h, r):
        f(ps, r["m"], r["n"], r["alpha"], h.desc_addr(r["sA"]), r["ai"], r["aj"],
          h.desc_addr(r["sB"]), r["bi"], r["bj"], h.desc_addr(r["sD"]), r["di"], r["dj"])

    print("%s ps=%d batch=%d" % (op, ps, len(recs)), flush=True)
    # thin: the one-wavefront kernel is forced (the built-in choice takes it only
    # for batches of at least 16 instances per CU)
    assert bfa.lib().blasfeo_amd_tune_set(b"trsm_thin", 0, 1 if thin else 0) == 0
    try:
        g, gd, rr, rd = run_both(b, buf, bfa.DTRSM_ARG, recs, "d" + op, ocall, (256, 4, 0) if thin else (260, 40, 0))
    finally:
        bfa.lib().blasfeo_amd_tune_set(b"trsm_thin", 0, 0)
    compare_buffers(b, g, rr, "d%s ps=%d" % (op, ps))
    assert (gd["use_dA"] == rd["use_dA"]).all()


def _pivots(rng, m):
    return np.array([int(rng.integers(i, m)) for i in range(m)], dtype=np.int32)


@pytest.mark.parametrize("ps", [4, 8])
def test_rowpe(gpu, ps):
    import torch
    rng = np.random.default_rng(940 + ps)
    b = Batch("d", ps, seed=ps)
    shapes = [(int(rng.integers(1, 300)), int(rng.integers(1, 90))) for _ in range(30)] + [(256, 4), (1, 1)]
    for m, n in shapes:
        b.mat(m, n)
    buf = b.host()
    piv = [_pivots(rng, m) for m, n in shapes]
    kmax = [m if i % 4 else max(1, m // 2) for i, (m, n) in enumerate(shapes)]
    flat = np.concatenate(piv)
    offs = np.concatenate([[0], np.cumsum([len(p) for p in piv])[:-1]])
    dev, host = DeviceBatch(b.layout, buf), HostBatch(b.layout, buf)
    dpiv = torch.from_numpy(flat).to("cuda")
    rec = np.zeros(len(shapes), dtype=bfa.DROWPE_ARG)
    rec["kmax"] = kmax
    rec["ipiv"] = np.uint64(dpiv.data_ptr()) + (4 * offs).astype(np.uint64)
    rec["sA"] = [dev.desc_addr(i) for i in range(len(shapes))]
    print("rowpe ps=%d batch=%d" % (ps, len(shapes)), flush=True)
    drec = to_device(rec)
    bfa.call("drowpe", ps, len(shapes), drec.data_ptr())
    torch.cuda.synchronize()
    for i in range(len(shapes)):
        oracle.fn("d", "rowpe")(ps, kmax[i], piv[i].ctypes.data, host.desc_addr(i))
    np.testing.assert_array_equal(dev.download(), host.buf)
    assert (dev.download_desc()["use_dA"] == 0).all()


@pytest.mark.parametrize("ps", [4, 8])
def test_lu_solve_pipeline(gpu, ps):
    """getrf_rp, rowpe, trsm_llnu, trsm_lunn on the GPU: pivots equal the
    oracle's, X has a normwise backward error <= 20 n eps, and X agrees with
    the oracle's pipeline to cond-level (1e-8 relative)."""
    import torch
    rng = np.random.default_rng(960 + ps)
    ns = [int(x) for x in rng.integers(1, 257, 24)] + [256, 1, 16]
    nrhs = 4
    b = Batch("d", ps, seed=ps + 5)
    As = [rng.uniform(-1, 1, (n, n)) for n in ns]
    Bs = [rng.standard_normal((n, nrhs)) for n in ns]
    for A, B in zip(As, Bs):
        b.mat(*A.shape, "data", A)
        b.mat(A.shape[0], A.shape[0], "const", 0.0)
        b.mat(*B.shape, "data", B)
    buf = b.host()
    dev, host = DeviceBatch(b.layout, buf), HostBatch(b.layout, buf)
    N = len(ns)
    offs = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.uint64)
    dpiv = torch.zeros(int(sum(ns)), dtype=torch.int32, device="cuda")
    f = np.zeros(N, dtype=bfa.DFACT_ARG)
    f["m"], f["n"] = ns, ns
    f["sC"] = [dev.desc_addr(3 * i) for i in range(N)]
    f["sD"] = [dev.desc_addr(3 * i + 1) for i in range(N)]
    f["ipiv"] = np.uint64(dpiv.data_ptr()) + 4 * offs
    p = np.zeros(N, dtype=bfa.DROWPE_ARG)
    p["kmax"] = ns
    p["ipiv"] = f["ipiv"]
    p["sA"] = [dev.desc_addr(3 * i + 2) for i in range(N)]
    t = np.zeros(N, dtype=bfa.DTRSM_ARG)
    t["m"], t["n"], t["alpha"] = ns, nrhs, 1.0
    t["sA"] = f["sD"]
    t["sB"] = p["sA"]
    t["sD"] = p["sA"]
    print("lu pipeline ps=%d batch=%d" % (ps, N), flush=True)
    df, dp = to_device(f), to_device(p)      # records kept alive until the launches ran
    bfa.call("dgetrf_rp", ps, N, df.data_ptr(), (256, 256, 0))
    bfa.call("drowpe", ps, N, dp.data_ptr())
    dt = to_device(t)
    bfa.call("dtrsm_llnu", ps, N, dt.data_ptr(), (256, nrhs, 0))
    bfa.call("dtrsm_lunn", ps, N, dt.data_ptr(), (256, nrhs, 0))
    torch.cuda.synchronize()
    g = dev.download()
    gp = dpiv.cpu().numpy()
    L = oracle.lib()
    eps = np.finfo(float).eps
    for i, n in enumerate(ns):
        hp = np.zeros(n, dtype=np.int32)
        a0, a1, a2 = (host.desc_addr(3 * i + q) for q in range(3))
        oracle.fn("d", "getrf_rp")(ps, n, n, a0, 0, 0, a1, 0, 0, hp.ctypes.data)
        oracle.fn("d", "rowpe")(ps, n, hp.ctypes.data, a2)
        oracle.fn("d", "trsm_llnu")(ps, n, nrhs, 1.0, a1, 0, 0, a2, 0, 0, a2, 0, 0)
        oracle.fn("d", "trsm_lunn")(ps, n, nrhs, 1.0, a1, 0, 0, a2, 0, 0, a2, 0, 0)
        np.testing.assert_array_equal(gp[int(offs[i]):int(offs[i]) + n], hp, err_msg="pivots %d" % i)
        X = b.layout.get(g, 3 * i + 2)
        Xr = b.layout.get(host.buf, 3 * i + 2)
        A, B = As[i], Bs[i]
        berr = np.abs(A @ X - B).max() / (np.abs(A).sum(1).max() * np.abs(X).max() + np.abs(B).max())
        assert berr <= 20 * n * eps, "instance %d n=%d berr %.3g" % (i, n, berr)
        np.testing.assert_allclose(X, Xr, rtol=1e-8, atol=1e-8 * np.abs(Xr).max(), err_msg="X %d" % i)


@pytest.mark.parametrize("ps", [4, 8])
def test_getrf_np(gpu, ps):
    """LU without pivoting on diagonally dominant blocks (single-wave and
    multi-wave panel paths, rectangular blocks, offsets), ipiv NULL: L, U,
    dA and use_dA against the oracle."""
    rng = np.random.default_rng(980 + ps)
    b = Batch("d", ps, seed=ps + 9)
    recs = []
    shapes = [(int(rng.integers(1, 300)),) * 2 for _ in range(16)] + [(40, 70), (150, 90), (256, 256), (1, 1), (64, 64), (65, 65)]
    for r, (m, n) in enumerate(shapes):
        ci, di = (int(x) for x in rng.integers(0, 2 * ps, 2))
        cj, dj = (int(x) for x in rng.integers(0, 3, 2))
        if r % 3 == 0:
            di = dj = 0
        A = rng.uniform(-1, 1, (ci + m, cj + n))
        for t in range(min(m, n)):
            A[ci + t, cj + t] += (n if t % 2 else -n)
        C = b.mat(ci + m, cj + n, "data", A)
        if r % 4 == 3:
            D, di, dj = C, ci, cj
        else:
            D = b.mat(di + m, dj + n, "const", -1.0)
        recs.append(dict(m=m, n=n, ci=ci, cj=cj, di=di, dj=dj, sC=C, sD=D, ipiv=0))
    buf = b.host()
    f = oracle.fn("d", "getrf_np")

    def ocall(h, r):
        f(ps, r["m"], r["n"], h.desc_addr(r["sC"]), r["ci"], r["cj"], h.desc_addr(r["sD"]), r["di"], r["dj"])

    print("getrf_np ps=%d batch=%d" % (ps, len(recs)), flush=True)
    g, gd, rr, rd = run_both(b, buf, bfa.DFACT_ARG, recs, "dgetrf_np", ocall, (300, 300, 0))
    compare_buffers(b, g, rr, "dgetrf_np ps=%d" % ps)
    assert (gd["use_dA"] == rd["use_dA"]).all()


@pytest.mark.parametrize("thin", [True, False])
@pytest.mark.parametrize("op", ["trsm_llnn", "trsm_lltn", "trsm_llnu", "trsm_lunn"])
def test_trsm_left_nonfinite(gpu, op, thin):
    """A zero diagonal entry (reciprocal Inf) or a NaN in B: every element,
    NaN and Inf included, equals the oracle's (rows solved before the
    non-finite one keep their values)."""
    import trsm_cases
    from chol_cases import to_batch
    ps = 4
    b = Batch("d", ps, seed=5)
    recs = to_batch(b, trsm_cases.nonfinite(ps, op, 970 + 3 * len(op) + thin))
    buf = b.host()
    f = oracle.fn("d", op)

    def ocall(h, r):
        f(ps, r["m"], r["n"], r["alpha"], h.desc_addr(r["sA"]), r["ai"], r["aj"],
          h.desc_addr(r["sB"]), r["bi"], r["bj"], h.desc_addr(r["sD"]), r["di"], r["dj"])

    print("%s nonfinite thin=%d" % (op, thin), flush=True)
    assert bfa.lib().blasfeo_amd_tune_set(b"trsm_thin", 0, 1 if thin else 2) == 0
    try:
        g, gd, rr, rd = run_both(b, buf, bfa.DTRSM_ARG, recs, "d" + op, ocall, (256, 4, 0))
    finally:
        bfa.lib().blasfeo_amd_tune_set(b"trsm_thin", 0, 0)
    lay = b.layout
    for i, r in enumerate(recs):
        gv, rv = lay.view(g, r["sD"]), lay.view(rr, r["sD"])
        with np.errstate(invalid="ignore"):
            same = (gv == rv) | (np.isnan(gv) & np.isnan(rv)) | (np.abs(gv - rv) <= 1e-11 * (np.abs(gv) + np.abs(rv)))
        assert same.all(), "%s instance %d: %d elements differ, first %s gpu=%r ref=%r" % (
            op, i, int((~same).sum()), np.argwhere(~same)[:1].tolist(), gv[~same][:1], rv[~same][:1])
    assert (gd["use_dA"] == rd["use_dA"]).all()
