"""GPU parity of the batched factorisations against the CPU oracle:
potrf_l, potrf_l_mn, syrk_dpotrf_ln and getrf_rp (FP64 and FP32), plus the
reference's failure conventions (non-SPD pivot -> zero reciprocal, zero LU
pivot -> zero reciprocal, NaN never chosen as pivot) and exact pivot ties
(first index wins, as idamax)."""
import numpy as np
import pytest

import blasfeo_amd as bfa
import oracle
from helpers import Batch, run_both, compare_buffers, assert_close
import chol_cases

pytestmark = pytest.mark.gpu


def _spd(rng, n):
    B = rng.uniform(-1, 1, (n, n))
    return B @ B.T + n * np.eye(n)


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("nmax", [32, 128])
def test_potrf_lds(gpu, prec, ps, nmax):
    """potrf_l with the whole matrix in LDS (max size <= 128; one wave per
    instance when <= 32)."""
    rng = np.random.default_rng(350 + ps + nmax)
    b = Batch(prec, ps, seed=7)
    recs = []
    sizes = [1, 2, 5, 16, 17, 31, 32] + ([33, 48, 63, 64, 65, 100, 127, 128] if nmax == 128 else [])
    for r, n in enumerate(sizes):
        ci, cj = (0, 0) if r % 2 == 0 else (int(rng.integers(0, 2 * ps)), int(rng.integers(0, 3)))
        di, dj = (0, 0) if r % 3 == 0 else (int(rng.integers(0, 2 * ps)), int(rng.integers(0, 3)))
        S = np.zeros((ci + n, cj + n))
        S[ci:, cj:] = _spd(rng, n)
        C = b.mat(ci + n, cj + n, "data", S)
        D = b.mat(di + n, dj + n, "const", -1.0)
        recs.append(dict(m=n, n=n, ci=ci, cj=cj, di=di, dj=dj, sC=C, sD=D, ipiv=0))
    buf = b.host()
    f = oracle.fn(prec, "potrf_l")

    def ocall(h, r):
        f(ps, r["m"], h.desc_addr(r["sC"]), r["ci"], r["cj"], h.desc_addr(r["sD"]), r["di"], r["dj"])

    tol = 1e-11 if prec == "d" else 2e-3
    g, gd, rr, rd = run_both(b, buf, bfa.DFACT_ARG, recs, prec + "potrf_l", ocall, (nmax, nmax, 0))
    compare_buffers(b, g, rr, "%spotrf_l lds ps=%d" % (prec, ps), tol=tol)
    assert (gd["use_dA"] == rd["use_dA"]).all()


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("op", ["potrf_l", "potrf_l_mn"])
def test_potrf(gpu, prec, ps, op):
    if prec == "s" and op == "potrf_l_mn":
        pytest.skip("no FP32 potrf_l_mn entry")
    b = Batch(prec, ps, seed=1)
    recs = chol_cases.to_batch(b, chol_cases.potrf(ps, op, chol_cases.POTRF_SIZES, 300 + ps))
    buf = b.host()
    f = oracle.fn(prec, op)

    def ocall(h, r):
        if op == "potrf_l":
            f(ps, r["m"], h.desc_addr(r["sC"]), r["ci"], r["cj"], h.desc_addr(r["sD"]), r["di"], r["dj"])
        else:
            f(ps, r["m"], r["n"], h.desc_addr(r["sC"]), r["ci"], r["cj"], h.desc_addr(r["sD"]), r["di"], r["dj"])

    argdt = bfa.DFACT_ARG
    tol = 1e-11 if prec == "d" else 2e-3
    # potrf_l_mn above m = 128: potrf_l of the n x n top block, then the rows
    # below solved with its reciprocal diagonal (n bounded by 256)
    hint = (280, 280, 0) if op == "potrf_l" else (280, 256, 0)
    g, gd, rr, rd = run_both(b, buf, argdt, recs, prec + op, ocall, hint)
    compare_buffers(b, g, rr, "%s%s ps=%d" % (prec, op, ps), tol=tol)
    assert (gd["use_dA"] == rd["use_dA"]).all()


@pytest.mark.parametrize("bw", [4, 2, 1])
@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("ps", [4, 8])
def test_potrf_tiles(gpu, prec, ps, bw):
    """potrf_l with a size bound in (128, 256]: block panels of 16 bw columns
    (the tuning knob potrf_bw: 64, 32 or 16 columns, one, two or four
    instances per CU) factored in LDS as 16x16 tiles, earlier block columns
    applied from D;
    sizes around the tile and panel edges, offsets, in place, a non-SPD pivot."""
    from blasfeo_amd import autotune
    autotune.set_knob("potrf_bw", 0, bw)
    try:
        _potrf_tiles_case(prec, ps)
    finally:
        autotune.set_knob("potrf_bw", 0, 0)


def _potrf_tiles_case(prec, ps):
    rng = np.random.default_rng(370 + ps)
    b = Batch(prec, ps, seed=9)
    recs = []
    sizes = [1, 16, 17, 63, 64, 65, 127, 128, 129, 143, 192, 200, 255, 256]
    for r, n in enumerate(sizes):
        ci, cj = (0, 0) if r % 2 == 0 else (int(rng.integers(0, 2 * ps)), int(rng.integers(0, 3)))
        di, dj = (0, 0) if r % 3 == 0 else (int(rng.integers(0, 2 * ps)), int(rng.integers(0, 3)))
        S = np.zeros((ci + n, cj + n))
        M = _spd(rng, n)
        if n == 143:
            M[100, 100] = -1.0  # pivot 100 fails: zero column, dA = 0
        S[ci:, cj:] = M
        C = b.mat(ci + n, cj + n, "data", S)
        if n == 200:
            D, di, dj = C, ci, cj
        else:
            D = b.mat(di + n, dj + n, "const", -1.0)
        recs.append(dict(m=n, n=n, ci=ci, cj=cj, di=di, dj=dj, sC=C, sD=D, ipiv=0))
    buf = b.host()
    f = oracle.fn(prec, "potrf_l")

    def ocall(h, r):
        f(ps, r["m"], h.desc_addr(r["sC"]), r["ci"], r["cj"], h.desc_addr(r["sD"]), r["di"], r["dj"])

    tol = 1e-11 if prec == "d" else 2e-3
    argdt = bfa.DFACT_ARG
    g, gd, rr, rd = run_both(b, buf, argdt, recs, prec + "potrf_l", ocall, (256, 256, 0))
    compare_buffers(b, g, rr, "%spotrf_l tiles ps=%d" % (prec, ps), tol=tol)
    assert (gd["use_dA"] == rd["use_dA"]).all()


@pytest.mark.parametrize("ps", [4, 8])
def test_potrf_inplace_and_failures(gpu, ps):
    """In-place factorisation, non-SPD pivots (zero column + dA = 0) and NaN."""
    rng = np.random.default_rng(5)
    b = Batch("d", ps, seed=2)
    recs = []
    for n, kind in [(20, "ok"), (40, "neg"), (70, "neg"), (33, "nan"), (5, "zero")]:
        M = _spd(rng, n)
        if kind == "neg":
            M[n // 2, n // 2] = -1.0
        elif kind == "nan":
            M[3, 3] = np.nan
        elif kind == "zero":
            M[:] = 0.0
        C = b.mat(n, n, "data", M)
        recs.append(dict(m=n, n=n, ci=0, cj=0, di=0, dj=0, sC=C, sD=C, ipiv=0))
    buf = b.host()
    f = oracle.fn("d", "potrf_l")

    def ocall(h, r):
        f(ps, r["m"], h.desc_addr(r["sC"]), 0, 0, h.desc_addr(r["sD"]), 0, 0)

    g, gd, rr, rd = run_both(b, buf, bfa.DFACT_ARG, recs, "dpotrf_l", ocall, (80, 80, 0))
    compare_buffers(b, g, rr, "dpotrf_l failures ps=%d" % ps)


@pytest.mark.parametrize("ps", [4, 8])
def test_syrk_dpotrf_ln(gpu, ps):
    """dsyrk_dpotrf_ln(_mn) for m up to 299 (n <= 256): C + A B^T formed in D
    (syrk_ln_mn), then the trapezoid factored in place; n = m and n < m,
    k = 0 .. 69, row offsets on A, B and D."""
    b = Batch("d", ps, seed=3)
    recs = chol_cases.to_batch(b, chol_cases.syrk_potrf(ps, 400 + ps))
    buf = b.host()
    f = oracle.fn("d", "syrk_potrf_ln_mn")

    def ocall(h, r):
        f(ps, r["m"], r["n"], r["k"], h.desc_addr(r["sA"]), r["ai"], r["aj"],
          h.desc_addr(r["sB"]), r["bi"], r["bj"], h.desc_addr(r["sC"]), r["ci"], r["cj"],
          h.desc_addr(r["sD"]), r["di"], r["dj"])

    g, gd, rr, rd = run_both(b, buf, bfa.DSYRK_DPOTRF_ARG, recs, "dsyrk_dpotrf_ln", ocall, (300, 256, 80))
    compare_buffers(b, g, rr, "dsyrk_dpotrf_ln ps=%d" % ps)
    assert (gd["use_dA"] == rd["use_dA"]).all()


@pytest.mark.parametrize("ps", [4, 8])
def test_potrf_large(gpu, ps):
    """potrf_l for n > 256 (16-column panels at full height in LDS, the
    earlier panels read back from D), with offsets and in place."""
    b = Batch("d", ps, seed=19)
    recs = chol_cases.to_batch(b, chol_cases.potrf(ps, "potrf_l", chol_cases.POTRF_LARGE_SIZES, 760 + ps, inplace_at=5))
    buf = b.host()
    f = oracle.fn("d", "potrf_l")

    def ocall(h, r):
        f(ps, r["m"], h.desc_addr(r["sC"]), r["ci"], r["cj"], h.desc_addr(r["sD"]), r["di"], r["dj"])

    g, gd, rr, rd = run_both(b, buf, bfa.DFACT_ARG, recs, "dpotrf_l", ocall, (420, 420, 0))
    compare_buffers(b, g, rr, "dpotrf_l large ps=%d" % ps)
    assert (gd["use_dA"] == rd["use_dA"]).all()


@pytest.mark.parametrize("threads", [0, 512])
@pytest.mark.parametrize("ps", [4, 8])
def test_potrf_mn_lds(gpu, ps, threads):
    """dpotrf_l_mn with m <= 128: the m x n lower trapezoid factored in LDS
    (L11 = chol(C11), L21 = C21 L11^{-T}), offsets, n = m and n < m, in place."""
    rng = np.random.default_rng(720 + ps)
    b = Batch("d", ps, seed=13)
    recs = []
    shapes = [(1, 1), (5, 3), (16, 16), (17, 9), (31, 31), (40, 17), (64, 33), (65, 64),
              (100, 64), (127, 100), (128, 128), (128, 1), (90, 90)]
    for r, (m, n) in enumerate(shapes):
        ci, cj = (0, 0) if r % 2 == 0 else (int(rng.integers(0, 2 * ps)), int(rng.integers(0, 3)))
        di, dj = (0, 0) if r % 3 == 0 else (int(rng.integers(0, 2 * ps)), int(rng.integers(0, 3)))
        S = np.zeros((ci + m, cj + n))
        S[ci:, cj:] = _spd(rng, m)[:, :n]
        C = b.mat(ci + m, cj + n, "data", S)
        if r == 12:
            D, di, dj = C, ci, cj
        else:
            D = b.mat(di + m, dj + n, "const", -1.0)
        recs.append(dict(m=m, n=n, ci=ci, cj=cj, di=di, dj=dj, sC=C, sD=D, ipiv=0))
    buf = b.host()
    f = oracle.fn("d", "potrf_l_mn")

    def ocall(h, r):
        f(ps, r["m"], r["n"], h.desc_addr(r["sC"]), r["ci"], r["cj"], h.desc_addr(r["sD"]), r["di"], r["dj"])

    # threads: the lds_threads knob of the 16-row buckets above 32 rows (0 = built-in)
    for bk in range(2, 8):
        assert bfa.lib().blasfeo_amd_tune_set(b"lds_threads", bk, threads) == 0
    try:
        g, gd, rr, rd = run_both(b, buf, bfa.DFACT_ARG, recs, "dpotrf_l_mn", ocall, (128, 128, 0))
    finally:
        for bk in range(2, 8):
            bfa.lib().blasfeo_amd_tune_set(b"lds_threads", bk, 0)
    compare_buffers(b, g, rr, "dpotrf_l_mn ps=%d threads=%d" % (ps, threads))
    assert (gd["use_dA"] == rd["use_dA"]).all()


@pytest.mark.parametrize("threads", [0, 256, 512])
@pytest.mark.parametrize("ps", [4, 8])
def test_syrk_dpotrf_ln_lds(gpu, ps, threads):
    """dsyrk_dpotrf_ln(_mn) with m <= 128: D = chol(C + A B^T) on the m x n lower
    trapezoid, A B^T added on the matrix cores from global memory; k = 0,
    A != B, offsets on every operand.  The instances go to 16-row buckets,
    each a persistent launch; threads: the lds_threads knob for the buckets
    above 32 rows (0 = built-in: four waves)."""
    rng = np.random.default_rng(740 + ps)
    b = Batch("d", ps, seed=17)
    recs = []
    cases = [(1, 1, 1), (7, 7, 0), (16, 16, 4), (33, 20, 9), (48, 48, 33), (64, 64, 64),
             (80, 80, 64), (100, 37, 5), (128, 128, 70), (120, 120, 1), (25, 25, 100)]
    for r, (m, n, k) in enumerate(cases):
        ai, bi, ci = (int(x) for x in rng.integers(0, 2 * ps, 3))
        aj, bj, cj = (int(x) for x in rng.integers(0, 3, 3))
        A = b.mat(ai + m, aj + max(k, 1))
        same = r % 2 == 0
        B = A if same else b.mat(ai + m, aj + max(k, 1))
        if same:
            bi, bj = ai, aj
        else:
            bi = ai  # B's rows are read at the same row offset as A's in this test
        S = np.zeros((ci + m, cj + n))
        S[ci:, cj:] = (m + 2 * k) * np.eye(m)[:, :n] + 0.01 * rng.standard_normal((m, n))
        C = b.mat(ci + m, cj + n, "data", S)
        D = b.mat(m, n, "const", -1.0)
        recs.append(dict(m=m, n=n, k=k, ai=ai, aj=aj, bi=bi, bj=bj, ci=ci, cj=cj, di=0, dj=0,
                         sA=A, sB=B, sC=C, sD=D))
    buf = b.host()
    f = oracle.fn("d", "syrk_potrf_ln_mn")

    def ocall(h, r):
        f(ps, r["m"], r["n"], r["k"], h.desc_addr(r["sA"]), r["ai"], r["aj"],
          h.desc_addr(r["sB"]), r["bi"], r["bj"], h.desc_addr(r["sC"]), r["ci"], r["cj"],
          h.desc_addr(r["sD"]), r["di"], r["dj"])

    for bk in range(2, 8):
        assert bfa.lib().blasfeo_amd_tune_set(b"lds_threads", bk, threads) == 0
    try:
        g, gd, rr, rd = run_both(b, buf, bfa.DSYRK_DPOTRF_ARG, recs, "dsyrk_dpotrf_ln", ocall, (128, 128, 100))
    finally:
        for bk in range(2, 8):
            bfa.lib().blasfeo_amd_tune_set(b"lds_threads", bk, 0)
    compare_buffers(b, g, rr, "dsyrk_dpotrf_ln ps=%d threads=%d" % (ps, threads))
    assert (gd["use_dA"] == rd["use_dA"]).all()


def _getrf_case(b, rng, m, n, data, ci=0, cj=0, di=0, dj=0, inplace=False):
    S = np.zeros((ci + m, cj + n))
    S[ci:, cj:] = data
    C = b.mat(ci + m, cj + n, "data", S)
    D = C if inplace else b.mat(di + m, dj + n, "const", -1.0)
    if inplace:
        di, dj = ci, cj
    return dict(m=m, n=n, ci=ci, cj=cj, di=di, dj=dj, sC=C, sD=D)


def _run_getrf(b, prec, recs, hint):
    """Runs getrf_rp on the GPU and the oracle; pivots go to one device array
    and one host array at the same per-instance offsets."""
    import torch
    ps = b.ps
    p = [min(r["m"], r["n"]) for r in recs]
    tot = max(1, sum(p))
    gpiv = torch.full((tot,), -7, dtype=torch.int32, device="cuda")
    hpiv = np.full(tot, -7, dtype=np.int32)
    off = np.concatenate([[0], np.cumsum(p)])[:-1]
    for r, o in zip(recs, off):
        r["ipiv"] = gpiv.data_ptr() + 4 * int(o)
        r["_o"] = int(o)
    f = oracle.fn(prec, "getrf_rp")

    def ocall(h, r):
        f(ps, r["m"], r["n"], h.desc_addr(r["sC"]), r["ci"], r["cj"], h.desc_addr(r["sD"]),
          r["di"], r["dj"], hpiv.ctypes.data + 4 * r["_o"])

    g, gd, rr, rd = run_both(b, b.host_buf, bfa.DFACT_ARG, recs, prec + "getrf_rp", ocall, hint)
    return g, gd, rr, rd, gpiv.cpu().numpy(), hpiv


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("ps", [4, 8])
def test_getrf_random(gpu, prec, ps):
    rng = np.random.default_rng(500 + ps)
    b = Batch(prec, ps, seed=4)
    recs = []
    shapes = [(1, 1), (2, 2), (5, 5), (16, 16), (17, 17), (33, 33), (64, 64), (100, 100),
              (129, 129), (200, 200), (256, 256), (40, 25), (25, 40), (150, 70), (70, 150)]
    for r, (m, n) in enumerate(shapes):
        ci, cj = (0, 0) if r % 2 == 0 else (int(rng.integers(0, 2 * ps)), int(rng.integers(0, 3)))
        di, dj = (0, 0) if r % 3 else (int(rng.integers(0, 2 * ps)), int(rng.integers(0, 3)))
        recs.append(_getrf_case(b, rng, m, n, rng.uniform(-1, 1, (m, n)), ci, cj, di, dj, inplace=(r % 4 == 1)))
    b.host_buf = b.host()
    g, gd, rr, rd, gp, hp = _run_getrf(b, prec, recs, (280, 280, 0))
    tol = 1e-11 if prec == "d" else 2e-3
    if prec == "d":
        np.testing.assert_array_equal(gp, hp)
        compare_buffers(b, g, rr, "%sgetrf_rp ps=%d" % (prec, ps), tol=tol)
    else:
        _check_fp32_getrf(b, recs, g, rr, gp, hp, tol)
    assert (gd["use_dA"] == rd["use_dA"]).all()


def _check_fp32_getrf(b, recs, g, rr, gp, hp, tol):
    """FP32 pivots must equal the oracle's except where the oracle itself met a
    near-tie: at the first differing column j its own L has another
    |L(r,j)| >= 1 - 1e-4 (a rounding-level tie between two candidate rows).
    Past such a divergence the instance is judged by its backward error
    ||P A - L U|| / ||A|| <= 64 n eps_fp32 instead of elementwise."""
    lay = b.layout
    eps = np.finfo(np.float32).eps
    for r in recs:
        m, n, o = r["m"], r["n"], r["_o"]
        p = min(m, n)
        gpv, hpv = gp[o:o + p], hp[o:o + p]
        if (gpv == hpv).all():
            compare_buffers(b, g, rr, "sgetrf_rp", tol=tol, mats=[r["sD"]])
            continue
        j = int(np.argmax(gpv != hpv))
        Lr = lay.get(rr, r["sD"], r["di"], r["dj"], m, n).astype(np.float64)
        col = np.abs(Lr[j + 1:, j])
        assert col.size and col.max() >= 1 - 1e-4, "pivot %d differs without a near-tie" % j
        A = lay.get(b.host_buf, r["sC"], r["ci"], r["cj"], m, n).astype(np.float64)
        LU = lay.get(g, r["sD"], r["di"], r["dj"], m, n).astype(np.float64)
        PA = A.copy()
        for t in range(p):
            PA[[t, gpv[t]]] = PA[[gpv[t], t]]
        L = np.tril(LU, -1)[:, :p] + np.eye(m, p)
        U = np.triu(LU)[:p, :]
        assert np.linalg.norm(PA - L @ U) <= 64 * max(m, n) * eps * np.linalg.norm(A)


@pytest.mark.parametrize("ps", [4, 8])
def test_getrf_ties_and_singular(gpu, ps):
    """Exact ties pick the first row; permutation matrices give known pivots;
    an all-zero column keeps its diagonal and stores dA = 0; NaN is never a pivot."""
    rng = np.random.default_rng(9)
    b = Batch("d", ps, seed=5)
    recs = []
    # integer-valued ties: every column's candidates have equal magnitude
    T = np.array([[1., 2., 3.], [-1., 5., 1.], [1., -2., 7.], [-1., 4., 2.]])
    recs.append(_getrf_case(b, rng, 4, 3, T))
    perm = rng.permutation(37)
    P = np.eye(37)[perm]
    recs.append(_getrf_case(b, rng, 37, 37, P))
    Z = rng.uniform(-1, 1, (20, 20))
    Z[:, 4] = 0.0
    recs.append(_getrf_case(b, rng, 20, 20, Z))
    N = rng.uniform(-1, 1, (18, 18))
    N[7, 0] = np.nan
    recs.append(_getrf_case(b, rng, 18, 18, N))
    H = np.ones((70, 70)) + np.eye(70)  # ties in every column of the first panel
    recs.append(_getrf_case(b, rng, 70, 70, H))
    b.host_buf = b.host()
    g, gd, rr, rd, gp, hp = _run_getrf(b, "d", recs, (80, 80, 0))
    np.testing.assert_array_equal(gp, hp)
    # permutation matrix: pivot rows follow the permutation
    assert gp[0] == 0  # first tie column: row 0 wins
    compare_buffers(b, g, rr, "dgetrf ties ps=%d" % ps)
