"""Pins the CPU oracle (oracle/oracle_impl.h) before it is trusted as the
checker.  The reference holds no golden vectors (its tests compare HP with
REF at run time, tests/test_x.c), and its C build needs a header its own
build system generates, so it is not built here.  Every oracle routine is
pinned instead by checks a plausible mistake would fail:

* exact answers: the reference's own integer-valued test data
  (tests/test_x.c:144-169) through gemm_nt / gemm_nn / syrk_ln / trmm_rlnn,
  compared bit-for-bit with exact integer arithmetic (committed fixtures in
  tests/golden/, made by tests/golden/make_golden.py);
* closed forms: A*I^T = A (examples/getting_started.c), trsm with A = I,
  potrf of a diagonal matrix = sqrt, LU of a permutation matrix, Riccati
  with A = I, B = 0, Q = R = I;
* identities: L L^T = A within 10 n eps, P A = L U, D A^T = alpha B;
* conventions: lower-only writes (upper part keeps the -1 sentinel of
  tests/test_x.c:354), zero/NaN pivots, first-index tie break, dA caching.
"""
import os

import numpy as np
import pytest

import blasfeo_amd as bfa
import oracle
from blasfeo_amd.device import HostBatch

GOLD = os.path.join(os.path.dirname(__file__), "golden")
EPS = np.finfo(np.float64).eps


def _hb(mats, ps=4, dtype=np.float64):
    """Host batch from a list of dense arrays."""
    lay = bfa.Layout(dtype, ps)
    for A in mats:
        lay.add(*np.shape(A))
    buf = lay.host_buffer()
    for i, A in enumerate(mats):
        lay.put(buf, i, A)
    return lay, HostBatch(lay, buf)


def _get(lay, hb, i, **kw):
    return lay.get(hb.buf, i, **kw)


@pytest.mark.parametrize("ps", [4, 8])
def test_gemm_exact_on_reference_test_data(ps):
    g = np.load(os.path.join(GOLD, "test_x_data_n12.npz"))
    A, B, C = g["A"], g["B"], g["C"]
    n = A.shape[0]
    lay, hb = _hb([A, B, C, np.full((n, n), -1.0)], ps)
    f = oracle.fn("d", "gemm_nt")
    f(ps, n, n, n, 1.0, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, 1.5, hb.desc_addr(2), 0, 0, hb.desc_addr(3), 0, 0)
    np.testing.assert_array_equal(_get(lay, hb, 3), g["gemm_nt"])
    f = oracle.fn("d", "gemm_nn")
    f(ps, n, n, n, 1.0, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, 1.5, hb.desc_addr(2), 0, 0, hb.desc_addr(3), 0, 0)
    np.testing.assert_array_equal(_get(lay, hb, 3), g["gemm_nn"])
    lay, hb = _hb([A, B, C, np.full((n, n), -1.0)], ps)
    f = oracle.fn("d", "syrk_ln")
    f(ps, n, n, 1.0, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, 1.5, hb.desc_addr(2), 0, 0, hb.desc_addr(3), 0, 0)
    np.testing.assert_array_equal(_get(lay, hb, 3), g["syrk_ln"])
    lay, hb = _hb([A, B, np.full((n, n), -1.0)], ps)
    f = oracle.fn("d", "trmm_rlnn")
    f(ps, n, n, 1.0, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0)
    np.testing.assert_array_equal(_get(lay, hb, 2), g["trmm_rlnn"])


@pytest.mark.parametrize("ps", [4, 8])
def test_gemm_offsets_identity(ps):
    """A * I^T = A at arbitrary offsets (panel-crossing), D outside untouched."""
    rng = np.random.default_rng(1)
    A = rng.standard_normal((23, 9))
    I = np.zeros((14, 12)); I[5:14, 3:12] = np.eye(9)
    D = np.full((30, 20), -1.0)
    lay, hb = _hb([A, I, D, D], ps)
    oracle.fn("d", "gemm_nt")(ps, 20, 9, 9, 1.0, hb.desc_addr(0), 3, 0, hb.desc_addr(1), 5, 3,
                              0.0, hb.desc_addr(2), 0, 0, hb.desc_addr(3), 7, 5)
    out = _get(lay, hb, 3)
    np.testing.assert_array_equal(out[7:27, 5:14], A[3:23, :])
    mask = np.ones_like(out, bool); mask[7:27, 5:14] = False
    assert (out[mask] == -1.0).all()


def test_syrk_lower_only_and_ln_mn():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((11, 6))
    C = rng.standard_normal((11, 11))
    lay, hb = _hb([A, C, np.full((11, 11), -1.0)])
    oracle.fn("d", "syrk_ln_mn")(4, 11, 7, 6, 1.0, hb.desc_addr(0), 0, 0, hb.desc_addr(0), 0, 0,
                                 1.0, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0)
    out = _get(lay, hb, 2)
    full = C + A @ A.T
    for j in range(11):
        for i in range(11):
            if j < 7 and i >= j:
                assert abs(out[i, j] - full[i, j]) <= 1e-14 * (1 + abs(full[i, j]))
            else:
                assert out[i, j] == -1.0


def test_trsm_identity_and_residual():
    rng = np.random.default_rng(3)
    n, m = 13, 9
    I = np.eye(n)
    B = rng.standard_normal((m, n))
    lay, hb = _hb([I, B, np.zeros((m, n))])
    oracle.fn("d", "trsm_rltn")(4, m, n, 2.0, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0)
    np.testing.assert_array_equal(_get(lay, hb, 2), 2.0 * B)
    assert hb.desc[0]["use_dA"] == n  # cache filled (blasfeo_hp_pm/d_blas3_lib4.c:4162-4168)
    Lw = np.tril(rng.standard_normal((n, n))) + 3 * np.eye(n)
    lay, hb = _hb([Lw, B, np.zeros((m, n))])
    oracle.fn("d", "trsm_rltn")(4, m, n, -0.5, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0)
    X = _get(lay, hb, 2)
    np.testing.assert_allclose(X @ Lw.T, -0.5 * B, rtol=0, atol=1e-13)
    np.testing.assert_array_equal(lay.dview(hb.buf, 0), 1.0 / np.diag(Lw))


def test_trsm_uses_cached_dA():
    """With use_dA >= n the cached reciprocals are used even if stale
    (blasfeo_hp_pm/d_blas3_lib4.c:4164): a planted cache changes the result."""
    n = 4
    Lw = np.eye(n) * 2.0
    B = np.ones((3, n))
    lay, hb = _hb([Lw, B, np.zeros((3, n))])
    lay.dview(hb.buf, 0)[:] = 10.0
    hb.desc[0]["use_dA"] = n
    oracle.fn("d", "trsm_rltn")(4, 3, n, 1.0, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0)
    np.testing.assert_array_equal(_get(lay, hb, 2), 10.0 * B)


@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("op", ["trsm_llnn", "trsm_lltn"])
def test_trsm_left_exact_identity_residual(ps, op):
    """Left-side solves: A = I gives alpha*B exactly; a unit-lower integer A
    with B = A X (or A^T X) for integer X is solved exactly (every partial
    sum is an integer), which fails for a wrong index, a dropped term or a
    transposed operand; a random well-conditioned A meets the residual
    identity; offsets leave the cache invalid (use_dA = 0)."""
    rng = np.random.default_rng(11 + ps)
    f = oracle.fn("d", op)
    m, n = 19, 5
    B = rng.standard_normal((m, n))
    lay, hb = _hb([np.eye(m), B, np.zeros((m, n))], ps)
    f(ps, m, n, -2.0, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0)
    np.testing.assert_array_equal(_get(lay, hb, 2), -2.0 * B)
    assert hb.desc[0]["use_dA"] == m
    A = np.tril(rng.integers(-3, 4, (m, m)).astype(float), -1) + np.eye(m)
    X = rng.integers(-5, 6, (m, n)).astype(float)
    Bx = (A if op == "trsm_llnn" else A.T) @ X
    lay, hb = _hb([A, Bx, np.zeros((m, n))], ps)
    f(ps, m, n, 1.0, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0)
    np.testing.assert_array_equal(_get(lay, hb, 2), X)
    Lw = rng.standard_normal((m + 3, m + 2))  # only the lower triangle of the block is read
    Lw[3:, 2:] = np.tril(Lw[3:, 2:]) * 0.3 + 2 * np.eye(m)
    lay, hb = _hb([Lw, B, np.zeros((m, n))], ps)
    f(ps, m, n, 0.5, hb.desc_addr(0), 3, 2, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0)
    T = np.tril(Lw[3:, 2:])
    Y = _get(lay, hb, 2)
    np.testing.assert_allclose((T if op == "trsm_llnn" else T.T) @ Y, 0.5 * B, rtol=0, atol=1e-13)
    assert hb.desc[0]["use_dA"] == 0
    np.testing.assert_array_equal(lay.dview(hb.buf, 0)[:m], 1.0 / np.diag(T))


def test_trsm_lltn_summation_order():
    """The reference's trsm_lltn subtracts the in-pair term last for the upper
    row of each bottom-up pair (blasfeo_ref/x_blas3_ref.c:730-744).  With
    values chosen so that this order changes the rounded result, the oracle
    matches the reference's order and not plain ascending order."""
    # m = 4: pairs (2,3) then (0,1); row 0 subtracts k = 2, 3 first and k = 1 last
    m = 4
    A = np.eye(m)
    A[1, 0] = 1.0
    A[2, 0] = 1.0
    A[3, 0] = 1.0
    b = np.array([[1.0], [1e16], [-1e16], [1.0]])
    lay, hb = _hb([A, b, np.zeros((m, 1))])
    oracle.fn("d", "trsm_lltn")(4, m, 1, 1.0, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0)
    x = _get(lay, hb, 2)[:, 0]
    # reference order for row 0: ((1 - x2) - x3) - x1 = ((1 + 1e16) - 1) - 1e16
    ref = ((1.0 - x[2]) - x[3]) - x[1]
    asc = ((1.0 - x[1]) - x[2]) - x[3]
    assert ref != asc
    assert x[0] == ref


@pytest.mark.parametrize("ps", [4, 8])
def test_potrf_closed_forms_and_backward_error(ps):
    rng = np.random.default_rng(4)
    d = rng.uniform(1, 5, 10)
    lay, hb = _hb([np.diag(d), np.full((10, 10), -1.0)], ps)
    oracle.fn("d", "potrf_l")(ps, 10, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0)
    L = _get(lay, hb, 1)
    np.testing.assert_allclose(np.diag(L), np.sqrt(d), rtol=2 * EPS)
    assert (L[np.triu_indices(10, 1)] == -1.0).all()
    np.testing.assert_array_equal(lay.dview(hb.buf, 1), 1.0 / np.sqrt(d))
    assert hb.desc[1]["use_dA"] == 10
    for n in (1, 7, 32, 61):
        B = rng.uniform(-1, 1, (n, n))
        A = B @ B.T + n * np.eye(n)
        lay, hb = _hb([A, np.zeros((n, n))], ps)
        oracle.fn("d", "potrf_l")(ps, n, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0)
        L = np.tril(_get(lay, hb, 1))
        assert np.linalg.norm(A - L @ L.T) / np.linalg.norm(A) <= 10 * n * EPS


def test_potrf_nonspd_convention():
    A = np.eye(5) * 4.0
    A[2, 2] = -1.0
    lay, hb = _hb([A, np.zeros((5, 5))])
    oracle.fn("d", "potrf_l")(4, 5, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0)
    L = _get(lay, hb, 1)
    assert L[2, 2] == -0.0 and lay.dview(hb.buf, 1)[2] == 0.0
    assert L[3, 3] == 2.0


def test_potrf_l_mn_and_syrk_potrf():
    rng = np.random.default_rng(5)
    m, n, k = 12, 7, 5
    A = rng.standard_normal((m, k))
    C = np.eye(m) * 3
    lay, hb = _hb([A, C, np.zeros((m, m))])
    oracle.fn("d", "syrk_potrf_ln_mn")(4, m, n, k, hb.desc_addr(0), 0, 0, hb.desc_addr(0), 0, 0,
                                       hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0)
    L = _get(lay, hb, 2)[:, :n]
    M = C + A @ A.T
    Lref = np.linalg.cholesky(M[:n, :n])
    np.testing.assert_allclose(np.tril(L[:n]), Lref, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(L[n:], M[n:, :n] @ np.linalg.inv(Lref).T, rtol=1e-12, atol=1e-13)


def _lu(ps, M, m=None, n=None):
    m = M.shape[0] if m is None else m
    n = M.shape[1] if n is None else n
    lay, hb = _hb([M, np.zeros(M.shape)], ps)
    piv = np.zeros(min(m, n), dtype=np.int32)
    oracle.fn("d", "getrf_rp")(ps, m, n, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, piv.ctypes.data)
    return _get(lay, hb, 1), piv, lay.dview(hb.buf, 1)


@pytest.mark.parametrize("ps", [4, 8])
def test_getrf_identities(ps):
    rng = np.random.default_rng(6)
    for (m, n) in [(9, 9), (30, 30), (17, 8), (8, 17)]:
        M = rng.uniform(-1, 1, (m, n))
        LU, piv, dA = _lu(ps, M)
        p = min(m, n)
        PA = M.copy()
        for j in range(p):
            PA[[j, piv[j]]] = PA[[piv[j], j]]
        L = np.tril(LU, -1)[:, :p] + np.eye(m, p)
        U = np.triu(LU)[:p, :]
        assert np.abs(PA - L @ U).max() <= 50 * max(m, n) * EPS
        np.testing.assert_array_equal(dA[:p], 1.0 / np.diag(U)[:p])
        # partial pivoting: |L| <= 1
        assert np.abs(np.tril(LU, -1)).max() <= 1.0


def test_getrf_permutation_and_ties():
    rng = np.random.default_rng(7)
    perm = rng.permutation(10)
    P = np.eye(10)[perm]
    LU, piv, _ = _lu(4, P)
    np.testing.assert_array_equal(LU, np.eye(10))
    # the pivot of column j is the row holding its 1 after the earlier swaps
    cur = list(perm)
    for j in range(10):
        r = cur.index(j)
        assert piv[j] == r
        cur[j], cur[r] = cur[r], cur[j]
    T = np.array([[1., 2.], [-1., 5.], [1., -2.]])
    _, piv, _ = _lu(4, T)
    assert piv[0] == 0  # exact tie in |x|: first index
    T2 = np.array([[0.5, 1.], [-1., 5.], [1., -2.]])
    _, piv, _ = _lu(4, T2)
    assert piv[0] == 1  # first of the two largest
    Z = np.zeros((4, 4)); Z[:, 1:] = rng.uniform(-1, 1, (4, 3))
    LU, piv, dA = _lu(4, Z)
    assert piv[0] == 0 and dA[0] == 0.0
    N = rng.uniform(-1, 1, (5, 5)); N[0, 0] = np.nan; N[3, 0] = 9.0
    _, piv, _ = _lu(4, N)
    assert piv[0] == 3


def test_riccati_closed_form_oracle():
    from blasfeo_amd.riccati import RiccatiBatch, HostRiccati
    N = 6
    rb = RiccatiBatch([4, 9], [2, 3], N=N, ps=4)
    inp = rb.fill_custom(lambda i, nx, nu: (np.vstack([np.zeros((nu, nx)), np.eye(nx)]),
                                            np.eye(nu + nx), np.eye(nx)))
    h = HostRiccati(rb, inp)
    oracle.lib().or_dric_trf_batch(4, rb.batch, h.argv.ctypes.data)
    for i, (nx, nu) in enumerate([(4, 2), (9, 3)]):
        for k in range(N):
            expect = np.diag(np.r_[np.ones(nu), np.full(nx, np.sqrt(N - k + 1))])
            np.testing.assert_allclose(np.tril(h.L(i, k)), expect, rtol=1e-15)


def test_riccati_matches_dense_dare_recursion():
    """Riccati factor L_k reproduces P_k = Q + A^T P A - A^T P B (R + B^T P B)^-1 B^T P A."""
    from blasfeo_amd.riccati import RiccatiBatch, HostRiccati
    rng = np.random.default_rng(8)
    N, nx, nu = 5, 6, 3
    A = np.eye(nx) + 0.1 * rng.standard_normal((nx, nx))
    B = rng.standard_normal((nx, nu))
    rb = RiccatiBatch([nx], [nu], N=N, ps=4)
    inp = rb.fill_custom(lambda i, nx_, nu_: (np.vstack([B.T, A.T]), np.eye(nu + nx), np.eye(nx)))
    h = HostRiccati(rb, inp)
    oracle.lib().or_dric_trf_batch(4, 1, h.argv.ctypes.data)
    P = np.eye(nx)
    for k in range(N - 1, -1, -1):
        G = np.eye(nu) + B.T @ P @ B
        P = np.eye(nx) + A.T @ P @ A - A.T @ P @ B @ np.linalg.solve(G, B.T @ P @ A)
        L = np.tril(h.L(0, k))
        Lx = L[nu:, nu:]
        np.testing.assert_allclose(Lx @ Lx.T, P, rtol=1e-12, atol=1e-12)


def test_float_twins_match_double_on_exact_data():
    g = np.load(os.path.join(GOLD, "test_x_data_n12.npz"))
    n = 6
    A = g["A"][:n, :n] / 16.0
    B = g["B"][:n, :n] / 16.0
    lay, hb = _hb([A, B, np.zeros((n, n)), np.zeros((n, n))], 8, np.float32)
    oracle.fn("s", "gemm_nt")(8, n, n, n, 1.0, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, 0.0,
                              hb.desc_addr(2), 0, 0, hb.desc_addr(3), 0, 0)
    np.testing.assert_array_equal(_get(lay, hb, 3), (A @ B.T).astype(np.float32))
