"""The LU workload at full size: 16384 blasfeo_dgetrf_rp calls, n ~ U{16..256},
entries U[-1,1] (seeded), ps = 4, in place.  Every instance's pivot vector
must equal the oracle's bit for bit (FP64), and the factors of the first 2048
instances must be backward stable, ||P A - L U||_F <= 10 n eps ||A||_F, with dA
the reciprocals of U's diagonal.  For
the FP32 twin the test counts the instances whose pivots differ from the
oracle's (a different summation order flips near-ties in single precision)
and requires each of them to be a backward-stable factorisation,
||P A - L U||_F <= 64 n eps32 ||A||_F."""
import numpy as np
import pytest

import blasfeo_amd as bfa
import oracle
from blasfeo_amd.device import to_device

pytestmark = pytest.mark.gpu

BATCH = 16384


def _batch(prec, seed):
    import torch
    dt = np.float64 if prec == "d" else np.float32
    rng = np.random.default_rng(seed)
    nn = rng.integers(16, 257, BATCH)
    d0, off, nb = bfa.bulk_layout(dt, 4, nn, nn, 0)
    host = rng.uniform(-1.0, 1.0, nb // np.dtype(dt).itemsize).astype(dt)
    buf = torch.from_numpy(host.copy()).to("cuda")
    hd = bfa.bulk_layout(dt, 4, nn, nn, buf.data_ptr())[0]
    d = to_device(hd)
    piv_off = np.concatenate([[0], np.cumsum(nn)[:-1]]).astype(np.uint64)
    piv = torch.full((int(nn.sum()),), -1, dtype=torch.int32, device="cuda")
    rec = np.zeros(BATCH, dtype=bfa.DFACT_ARG)
    rec["m"], rec["n"] = nn, nn
    rec["sC"] = d.data_ptr() + np.arange(BATCH, dtype=np.uint64) * bfa.DMAT.itemsize
    rec["sD"] = rec["sC"]
    rec["ipiv"] = piv.data_ptr() + 4 * piv_off
    drec = to_device(rec)
    print("%sgetrf_rp batch=%d" % (prec, BATCH), flush=True)
    bfa.call(prec + "getrf_rp", 4, BATCH, drec.data_ptr(), (256, 256, 0))
    torch.cuda.synchronize()
    gout = buf.cpu().numpy()
    gpiv = piv.cpu().numpy()
    # the oracle on the same input bytes
    hbuf = host.copy()
    hdesc = bfa.bulk_layout(dt, 4, nn, nn, hbuf.ctypes.data)[0]
    hpiv = np.full(int(nn.sum()), -1, dtype=np.int32)
    hrec = rec.copy()
    hrec["sC"] = hdesc.ctypes.data + np.arange(BATCH, dtype=np.uint64) * bfa.DMAT.itemsize
    hrec["sD"] = hrec["sC"]
    hrec["ipiv"] = hpiv.ctypes.data + 4 * piv_off
    getattr(oracle.lib(), "or_%sgetrf_rp_batch" % prec)(4, BATCH, hrec.ctypes.data)
    return nn, off, host, gout, hbuf, gpiv, hpiv, piv_off.astype(np.int64)


def test_getrf_rp_full_batch_pivots_fp64(gpu):
    nn, off, host, gout, hbuf, gpiv, hpiv, piv_off = _batch("d", 5005)
    bad = [i for i in range(BATCH)
           if not np.array_equal(gpiv[piv_off[i]: piv_off[i] + nn[i]], hpiv[piv_off[i]: piv_off[i] + nn[i]])]
    print("fp64: %d of %d pivot vectors differ" % (len(bad), BATCH), flush=True)
    assert not bad, "instances with differing pivots: %s" % bad[:10]
    # the factors: backward error ||P A - L U||_F <= 10 n eps ||A||_F (the FP64
    # bar; elementwise agreement with the oracle degrades with n for LU, whose
    # later columns accumulate n rounding steps in a different order), and dA
    # the reciprocal of the computed U's diagonal
    eps = np.finfo(np.float64).eps
    worst = 0.0
    for i in range(2048):
        n = int(nn[i])
        cn = pm = -(-n // 4) * 4
        o = int(off[i]) // 8
        A = bfa.pm_unpack(host[o: o + pm * cn], 4, cn, n, n)
        F = bfa.pm_unpack(gout[o: o + pm * cn], 4, cn, n, n)
        gp = gpiv[piv_off[i]: piv_off[i] + n]
        PA = A.copy()
        for j in range(n):
            p = int(gp[j])
            PA[[j, p]] = PA[[p, j]]
        L = np.tril(F, -1) + np.eye(n)
        U = np.triu(F)
        r = np.linalg.norm(PA - L @ U) / (np.linalg.norm(A) * n * eps)
        worst = max(worst, r)
        assert r <= 10.0, "instance %d (n=%d): backward error %.2f n eps" % (i, n, r)
        dA = gout[o + pm * cn: o + pm * cn + n]
        np.testing.assert_array_equal(dA, 1.0 / np.diag(U), err_msg="instance %d dA" % i)
    print("fp64: worst backward error of the first 2048 factors %.3f n eps" % worst, flush=True)


def test_getrf_rp_full_batch_pivots_fp32(gpu):
    nn, off, host, gout, hbuf, gpiv, hpiv, piv_off = _batch("s", 5006)
    eps = np.finfo(np.float32).eps
    bad = []
    for i in range(BATCH):
        n = int(nn[i])
        gp = gpiv[piv_off[i]: piv_off[i] + n]
        if np.array_equal(gp, hpiv[piv_off[i]: piv_off[i] + n]):
            continue
        bad.append(i)
        cn = pm = -(-n // 4) * 4
        o = int(off[i]) // 4
        A = bfa.pm_unpack(host[o: o + pm * cn].astype(np.float64), 4, cn, n, n)
        F = bfa.pm_unpack(gout[o: o + pm * cn].astype(np.float64), 4, cn, n, n)
        L = np.tril(F, -1) + np.eye(n)
        U = np.triu(F)
        PA = A.copy()
        for j in range(n):
            p = int(gp[j])
            PA[[j, p]] = PA[[p, j]]
        r = np.linalg.norm(PA - L @ U) / np.linalg.norm(A)
        assert r <= 64 * n * eps, "instance %d (n=%d): backward error %.3g" % (i, n, r)
    print("fp32: %d of %d pivot vectors differ from the oracle's (all backward stable)" % (len(bad), BATCH),
          flush=True)
