"""GPU tests of the batched SPD solve blasfeo_amd_dposv_mixed_batched.

The mixed path returns X with normwise backward error <= tol (reported in
the per-instance info record and recomputed here in FP64 from the returned
X).  Instances the FP32 factor cannot serve take the FP64 path, which
computes what blasfeo_dpotrf_l + blasfeo_dtrsm_llnn + blasfeo_dtrsm_lltn
compute (blasfeo_hp_pm/d_lapack_lib4.c:4269; d_blas3_lib4.c:7236,3782); for
those the factor L and X are compared with the oracle's restatement of that
sequence at the reference's 1e-11 rule."""
import numpy as np
import pytest

import blasfeo_amd as bfa
import oracle
from helpers import Batch, assert_close

pytestmark = pytest.mark.gpu


def _spd_cond(rng, n, cond):
    """Q diag(lambda) Q^T, Q Haar (QR of N(0,1) with sign fix), lambda
    geometric from 1 down to 1/cond."""
    Q, R = np.linalg.qr(rng.standard_normal((n, n)))
    Q *= np.sign(np.diag(R))
    lam = cond ** (-np.arange(n) / max(n - 1, 1))
    A = (Q * lam) @ Q.T
    return (A + A.T) / 2


def _berr(A, X, B):
    r = B - A @ X
    den = np.abs(A).sum(1).max() * np.abs(X).max(0) + np.abs(B).max(0)
    return float((np.abs(r).max(0) / den).max())


def _run(ps, cases, tol, max_iter, nmax, seed):
    """cases: list of (n, nrhs, A) -> per-instance (X, L, info) from the GPU and
    the oracle's FP64 solve of the same bytes."""
    import torch
    from blasfeo_amd.device import DeviceBatch, HostBatch, to_device
    rng = np.random.default_rng(seed)
    b = Batch("d", ps, seed=seed)
    recs = []
    for r, (n, nrhs, A) in enumerate(cases):
        ai, aj = (0, 0) if r % 2 == 0 else (int(rng.integers(0, 2 * ps)), int(rng.integers(0, 3)))
        bi, bj = (0, 0) if r % 3 == 0 else (int(rng.integers(0, 2 * ps)), int(rng.integers(0, 3)))
        S = rng.standard_normal((ai + n, aj + n))  # upper part of A's block is never read
        S[ai:, aj:] = np.tril(A) + np.triu(rng.standard_normal((n, n)), 1)
        sA = b.mat(ai + n, aj + n, "data", S)
        sB = b.mat(bi + n, bj + nrhs, "data", rng.standard_normal((bi + n, bj + nrhs)))
        sX = b.mat(n, nrhs, "const", -1.0)
        sL = b.mat(n, n, "const", -1.0)
        recs.append(dict(n=n, nrhs=nrhs, ai=ai, aj=aj, bi=bi, bj=bj, xi=0, xj=0, li=0, lj=0,
                         sA=sA, sB=sB, sX=sX, sL=sL))
    buf = b.host()
    dev = DeviceBatch(b.layout, buf)
    host = HostBatch(b.layout, buf)
    info = torch.zeros(len(recs) * bfa.SOLVE_INFO.itemsize, dtype=torch.uint8, device="cuda")
    darg = np.zeros(len(recs), dtype=bfa.DPOSV_ARG)
    harg = np.zeros(len(recs), dtype=bfa.DPOSV_ARG)
    for i, r in enumerate(recs):
        for f, v in r.items():
            if f.startswith("s"):
                darg[i][f] = dev.desc_addr(v)
                harg[i][f] = host.desc_addr(v)
            else:
                darg[i][f] = harg[i][f] = v
        darg[i]["info"] = info.data_ptr() + i * bfa.SOLVE_INFO.itemsize
    d = to_device(darg)
    print("posv_mixed ps=%d batch=%d sizes=%s" % (ps, len(recs), [c[0] for c in cases]), flush=True)
    bfa.posv_mixed(ps, len(recs), d.data_ptr(), tol, max_iter,
                   (nmax, max(c[1] for c in cases), 0))
    torch.cuda.synchronize()
    oracle.lib().or_dposv_batch(ps, len(recs), harg.ctypes.data)
    g = dev.download()
    inf = info.cpu().numpy().view(bfa.SOLVE_INFO)
    lay = b.layout
    out = []
    for i, r in enumerate(recs):
        n, nrhs = r["n"], r["nrhs"]
        A = np.tril(lay.get(buf, r["sA"], r["ai"], r["aj"], n, n))
        A = A + np.tril(A, -1).T
        B = lay.get(buf, r["sB"], r["bi"], r["bj"], n, nrhs)
        out.append(dict(A=A, B=B, X=lay.get(g, r["sX"]), L=lay.get(g, r["sL"]),
                        Xo=lay.get(host.buf, r["sX"]), Lo=lay.get(host.buf, r["sL"]),
                        dL=lay.dview(g, r["sL"]), dLo=lay.dview(host.buf, r["sL"]), info=inf[i]))
    return out


@pytest.mark.parametrize("ps", [4, 8])
def test_posv_mixed_converges(gpu, ps):
    """cfg-4-like instances (Haar Q, geometric spectrum, cond log-uniform in
    [1e2, 1e8]) at uneven n and nrhs: every instance converges on the mixed
    path (conjugate gradients preconditioned by the FP32 factor, at most 30
    iterations) to berr <= 1e-14, and X agrees with the FP64 solve to the
    accuracy the condition number allows."""
    rng = np.random.default_rng(900 + ps)
    cases = []
    for n, nrhs in [(1, 1), (5, 2), (16, 4), (17, 3), (33, 4), (64, 5), (100, 4),
                    (129, 8), (200, 4), (256, 4), (256, 1), (250, 6)]:
        cond = 10 ** rng.uniform(2, 8)
        cases.append((n, nrhs, _spd_cond(rng, n, cond)))
    cases.append((256, 4, _spd_cond(rng, 256, 1e8)))  # the top of the range
    out = _run(ps, cases, 1e-14, 30, 256, 31 + ps)
    for k, (o, (n, nrhs, A)) in enumerate(zip(out, cases)):
        inf = o["info"]
        assert inf["path"] == 0, (k, n, inf)
        assert 0 <= inf["berr"] <= 1e-14, (k, n, inf)
        be = _berr(o["A"], o["X"], o["B"])
        assert be <= 1.5e-14, (k, n, be, inf)
        cond = np.linalg.cond(o["A"])
        fe = np.abs(o["X"] - o["Xo"]).max() / np.abs(o["Xo"]).max()
        assert fe <= max(cond, 10.0) * 1e-13, (k, n, fe, cond)


@pytest.mark.parametrize("ps", [4, 8])
def test_posv_fallback_matches_fp64_sequence(gpu, ps):
    """Instances the FP32 factor cannot serve take the FP64 path and then equal
    the reference sequence potrf_l + trsm_llnn + trsm_lltn bit for bit up to
    the 1e-11 rule: an indefinite matrix (the FP32 factor breaks down; the
    FP64 factor follows the zero-pivot convention), max_iter = 0 on a
    moderate matrix (the first FP32 solve is not accurate to 1e-14), and n
    beyond the LDS-resident size (n > 256).  At cond 1e13 the refinement
    stalls and the FP64 path's backward error is checked instead."""
    rng = np.random.default_rng(950 + ps)
    indef = _spd_cond(rng, 40, 1e3)
    indef[7, 7] = -5.0
    cases = [(40, 3, indef), (96, 4, _spd_cond(rng, 96, 1e13)), (70, 2, _spd_cond(rng, 70, 1e3))]
    out = _run(ps, cases, 1e-14, 10, 256, 41 + ps)
    assert out[0]["info"]["path"] == 1
    assert out[1]["info"]["path"] == 1
    assert out[2]["info"]["path"] == 0
    o = out[0]
    assert_close(o["L"], o["Lo"], 1e-11, "fallback L")
    assert_close(o["dL"], o["dLo"], 1e-11, "fallback dA")
    assert_close(o["X"], o["Xo"], 1e-11, "fallback X")
    # at cond 1e13 the last pivots cancel, so elementwise parity is not a
    # property of the problem; the FP64 solve is still backward stable
    o = out[1]
    assert o["info"]["berr"] == -1.0
    assert _berr(o["A"], o["X"], o["B"]) <= 1e-13
    # max_iter = 0 and a size bound below n send everything to the FP64 path;
    # cond 10 keeps X itself comparable elementwise at 1e-11
    cases = [(70, 2, _spd_cond(rng, 70, 10.0)), (300, 4, _spd_cond(rng, 300, 10.0))]
    out = _run(ps, cases, 1e-14, 0, 320, 51 + ps)
    for k, o in enumerate(out):
        assert o["info"]["path"] == 1, (k, o["info"])
        assert_close(o["L"], o["Lo"], 1e-11, "fallback L")
        assert_close(o["X"], o["Xo"], 1e-11, "fallback X")
