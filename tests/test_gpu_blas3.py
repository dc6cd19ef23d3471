"""GPU parity of the batched level-3 routines against the CPU oracle:
gemm_nt, gemm_nn, syrk_ln, syrk_ln_mn, trmm_rlnn, trsm_rltn (FP64 and FP32),
with random sizes crossing the 64x64 tile edges and random row/column offsets
that cross panel boundaries, as the reference's sweep does (tests/test_x.c:320-388)."""
import numpy as np
import pytest

import blasfeo_amd as bfa
import oracle
from helpers import Batch, run_both, compare_buffers

pytestmark = pytest.mark.gpu


def _gemm_records(b, rng, nrec, op, maxdim=140):
    recs = []
    ps = b.ps
    for r in range(nrec):
        m, n, k = (int(x) for x in rng.integers(1, maxdim + 1, 3))
        if r == 0:
            k = 0
        if op in ("syrk_ln",):
            n = m
        ai, bi, ci, di = (int(x) for x in rng.integers(0, 2 * ps, 4))
        aj, bj, cj, dj = (int(x) for x in rng.integers(0, 4, 4))
        if op == "gemm_nn":
            A = b.mat(ai + m, aj + k)
            B = b.mat(bi + k, bj + n)
        else:
            A = b.mat(ai + m, aj + k)
            B = b.mat(bi + n, bj + k)
        C = b.mat(ci + m, cj + n)
        D = b.mat(di + m, dj + n, "const", -1.0)
        alpha = float(rng.choice([1.0, -1.0, 0.5]))
        beta = float(rng.choice([1.0, 0.0, 1.5]))
        recs.append(dict(m=m, n=n, k=k, ai=ai, aj=aj, bi=bi, bj=bj, ci=ci, cj=cj, di=di, dj=dj,
                         alpha=alpha, beta=beta, sA=A, sB=B, sC=C, sD=D))
    return recs


@pytest.mark.parametrize("engine", ["wave", "lds"])
@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("op", ["gemm_nt", "gemm_nn", "syrk_ln", "syrk_ln_mn"])
def test_gemm_family(gpu, prec, ps, op, engine):
    """engine: 'wave' (size bounds <= 256: one wavefront per 32 x 32 tile, the
    built-in choice here) or 'lds' (the LDS-staged 64 x 64 engine, forced with
    the gemm_wave knob)."""
    if prec == "s" and op == "syrk_ln_mn":
        pytest.skip("no FP32 syrk_ln_mn entry")
    rng = np.random.default_rng(100 + ps)
    b = Batch(prec, ps, seed=ps)
    recs = _gemm_records(b, rng, 24, op)
    buf = b.host()
    f = oracle.fn(prec, op)

    def ocall(h, r):
        if op == "syrk_ln":
            f(ps, r["m"], r["k"], r["alpha"], h.desc_addr(r["sA"]), r["ai"], r["aj"],
              h.desc_addr(r["sB"]), r["bi"], r["bj"], r["beta"], h.desc_addr(r["sC"]), r["ci"], r["cj"],
              h.desc_addr(r["sD"]), r["di"], r["dj"])
        else:
            f(ps, r["m"], r["n"], r["k"], r["alpha"], h.desc_addr(r["sA"]), r["ai"], r["aj"],
              h.desc_addr(r["sB"]), r["bi"], r["bj"], r["beta"], h.desc_addr(r["sC"]), r["ci"], r["cj"],
              h.desc_addr(r["sD"]), r["di"], r["dj"])

    argdt = bfa.DGEMM_ARG if prec == "d" else bfa.SGEMM_ARG
    hint = (160, 160, 160)
    assert bfa.lib().blasfeo_amd_tune_set(b"gemm_wave", 0, 0 if engine == "wave" else 2) == 0
    try:
        g, gd, rr, rd = run_both(b, buf, argdt, recs, prec + op, ocall, hint)
    finally:
        bfa.lib().blasfeo_amd_tune_set(b"gemm_wave", 0, 0)
    compare_buffers(b, g, rr, "%s%s ps=%d %s" % (prec, op, ps, engine))
    assert (gd["use_dA"] == rd["use_dA"]).all()


@pytest.mark.parametrize("ps", [4, 8])
def test_gemm_inplace_and_empty(gpu, ps):
    """D aliasing C (accumulate in place) and m=0 / n=0 no-ops."""
    rng = np.random.default_rng(7)
    b = Batch("d", ps, seed=11)
    recs = []
    for r in range(8):
        m, n, k = (int(x) for x in rng.integers(1, 100, 3))
        A = b.mat(m + 3, k)
        B = b.mat(n, k + 2)
        C = b.mat(m + 5, n + 1)
        recs.append(dict(m=m, n=n, k=k, ai=3, aj=0, bi=0, bj=2, ci=5, cj=1, di=5, dj=1,
                         alpha=1.0, beta=1.0, sA=A, sB=B, sC=C, sD=C))
    E = b.mat(8, 8, "const", 3.0)
    recs.append(dict(m=0, n=5, k=3, ai=0, aj=0, bi=0, bj=0, ci=0, cj=0, di=0, dj=0,
                     alpha=1.0, beta=1.0, sA=E, sB=E, sC=E, sD=E))
    recs.append(dict(m=5, n=0, k=3, ai=0, aj=0, bi=0, bj=0, ci=0, cj=0, di=0, dj=0,
                     alpha=1.0, beta=1.0, sA=E, sB=E, sC=E, sD=E))
    buf = b.host()
    f = oracle.fn("d", "gemm_nt")

    def ocall(h, r):
        f(ps, r["m"], r["n"], r["k"], r["alpha"], h.desc_addr(r["sA"]), r["ai"], r["aj"],
          h.desc_addr(r["sB"]), r["bi"], r["bj"], r["beta"], h.desc_addr(r["sC"]), r["ci"], r["cj"],
          h.desc_addr(r["sD"]), r["di"], r["dj"])

    g, gd, rr, rd = run_both(b, buf, bfa.DGEMM_ARG, recs, "dgemm_nt", ocall, (100, 100, 100))
    compare_buffers(b, g, rr, "dgemm_nt in place ps=%d" % ps)


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("op", ["trsm_rltn", "trmm_rlnn"])
def test_trxm(gpu, prec, ps, op):
    rng = np.random.default_rng(200 + ps)
    b = Batch(prec, ps, seed=ps + 3)
    recs = []
    for r in range(20):
        m = int(rng.integers(1, 150))
        n = int(rng.integers(1, 90))
        ai, bi, di = (int(x) for x in rng.integers(0, 2 * ps, 3))
        aj, bj, dj = (int(x) for x in rng.integers(0, 3, 3))
        if r % 3 == 0:
            ai = aj = 0
        # well-conditioned lower-triangular A: unit-ish diagonal
        Ad = np.tril(rng.standard_normal((ai + n, aj + n)) * 0.1)
        for t in range(n):
            Ad[ai + t, aj + t] = 1.0 + rng.uniform(0.5, 1.5)
        A = b.mat(ai + n, aj + n, "data", Ad)
        B = b.mat(bi + m, bj + n)
        D = b.mat(di + m, dj + n, "const", -1.0)
        recs.append(dict(m=m, n=n, ai=ai, aj=aj, bi=bi, bj=bj, di=di, dj=dj,
                         alpha=float(rng.choice([1.0, -0.5, 2.0])), sA=A, sB=B, sD=D))
    buf = b.host()
    f = oracle.fn(prec, op)

    def ocall(h, r):
        f(ps, r["m"], r["n"], r["alpha"], h.desc_addr(r["sA"]), r["ai"], r["aj"],
          h.desc_addr(r["sB"]), r["bi"], r["bj"], h.desc_addr(r["sD"]), r["di"], r["dj"])

    argdt = bfa.DTRSM_ARG if prec == "d" else bfa.STRSM_ARG
    name = prec + op
    if prec == "s" and op == "trmm_rlnn":
        pytest.skip("no FP32 trmm entry")
    g, gd, rr, rd = run_both(b, buf, argdt, recs, name, ocall, (160, 96, 96))
    compare_buffers(b, g, rr, "%s ps=%d" % (name, ps))
    assert (gd["use_dA"] == rd["use_dA"]).all()


@pytest.mark.parametrize("thin", [False, True])
@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("op", ["trsm_llnn", "trsm_lltn"])
def test_trsm_left(gpu, ps, op, thin):
    """D = alpha A^{-1} B and D = alpha A^{-T} B (A lower m x m), with offsets,
    the reciprocal-diagonal cache of A (valid, stale and offset cases), the
    right-hand sides spread over several column blocks, and in-place D = B.
    thin: at most 4 right-hand sides and m <= 256 (hint (256, 4)), the
    one-wavefront kernel."""
    rng = np.random.default_rng(400 + ps + 7 * thin)
    b = Batch("d", ps, seed=ps + 11)
    recs = []
    for r in range(18):
        m = int(rng.integers(1, 257 if thin else 260)) if r else 256
        n = (int(rng.integers(1, 5)) if thin else int(rng.integers(1, 40))) if r else 4
        ai, bi, di = (int(x) for x in rng.integers(0, 2 * ps, 3))
        aj, bj, dj = (int(x) for x in rng.integers(0, 3, 3))
        if r % 3 == 0:
            ai = aj = 0
        Ad = np.tril(rng.standard_normal((ai + m, aj + m)) * 0.1)
        for t in range(m):
            Ad[ai + t, aj + t] = 1.0 + rng.uniform(0.5, 1.5)
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

    # thin: the one-wavefront kernel is forced (the built-in choice takes it only
    # for batches of at least 16 instances per CU)
    assert bfa.lib().blasfeo_amd_tune_set(b"trsm_thin", 0, 1 if thin else 0) == 0
    try:
        g, gd, rr, rd = run_both(b, buf, bfa.DTRSM_ARG, recs, "d" + op, ocall, (256, 4, 0) if thin else (260, 40, 0))
    finally:
        bfa.lib().blasfeo_amd_tune_set(b"trsm_thin", 0, 0)
    compare_buffers(b, g, rr, "d%s ps=%d" % (op, ps))
    assert (gd["use_dA"] == rd["use_dA"]).all()
