"""Pins the oracle's level-2 routines and its Riccati solve (oracle_impl.h:
gemv_n/t, trmv_lnn/ltn, trsv_lnn_mn/ltn_mn restating
blasfeo_ref/x_blas2_ref.c, and ric_trs restating d_back_ric_trs_libstr,
examples/example_d_riccati_recursion.c:203-267).

* exact answers: integer matrices and vectors (dyadic diagonals for the
  solves) at offsets that cross panels, both panel sizes, odd and even
  sizes, compared bit for bit with exact numpy integer arithmetic;
* conventions: the reciprocal-diagonal cache of trsv (formed and flagged at
  offset 0, flagged invalid elsewhere, a planted cache used as is);
* the Riccati solve against a dense solve of the same equality-constrained
  quadratic program's KKT system: for stages z_k = [u_k; x_k]
      min sum_k 1/2 z_k^T H_k z_k + g_k^T z_k
      s.t. x_{k+1} = BAbt_k^T z_k + b_k,
  the solve returns ux_k = z_k and pi_k = the multiplier of stage k's
  dynamics, and Pb_k = P_{k+1} b_k with P_{k+1} the cost-to-go Hessian.
"""
import numpy as np
import pytest

import blasfeo_amd as bfa
import oracle
from blasfeo_amd.device import HostBatch


def _mat(A, ps):
    lay = bfa.Layout(np.float64, ps)
    lay.add(*np.shape(A))
    buf = lay.host_buffer()
    lay.put(buf, 0, A)
    return lay, HostBatch(lay, buf)


class _Vec:
    """A host vector with its blasfeo_dvec descriptor."""

    def __init__(self, x):
        self.x = np.array(x, dtype=np.float64)
        self.d = np.zeros(1, dtype=bfa.DVEC)
        self.d["mem"] = self.d["pa"] = self.x.ctypes.data
        self.d["m"] = self.d["pm"] = len(self.x)
        self.d["memsize"] = 8 * len(self.x)

    @property
    def p(self):
        return self.d.ctypes.data


def _ints(rng, shape, lo=-5, hi=6):
    return rng.integers(lo, hi, shape).astype(np.float64)


@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("m,n,ai,aj", [(7, 5, 3, 2), (8, 6, 5, 9), (1, 1, 0, 0), (12, 9, 0, 1)])
def test_gemv_exact(ps, m, n, ai, aj):
    rng = np.random.default_rng(m * 100 + n + ps)
    A = _ints(rng, (ai + m + 2, aj + n + 1))
    lay, hb = _mat(A, ps)
    Ab = A[ai:ai + m, aj:aj + n]
    xn, y, z = _Vec(_ints(rng, n + 3)), _Vec(_ints(rng, m + 2)), _Vec(np.full(m + 4, -1.0))
    oracle.fn("d", "gemv_n")(ps, m, n, 2.0, hb.desc_addr(0), ai, aj, xn.p, 1, -1.0, y.p, 2, z.p, 1)
    np.testing.assert_array_equal(z.x[1:1 + m], 2 * Ab @ xn.x[1:1 + n] - y.x[2:2 + m])
    assert z.x[0] == -1 and np.all(z.x[1 + m:] == -1)
    xt, y, z = _Vec(_ints(rng, m + 1)), _Vec(_ints(rng, n)), _Vec(np.full(n, -1.0))
    oracle.fn("d", "gemv_t")(ps, m, n, -1.0, hb.desc_addr(0), ai, aj, xt.p, 1, 3.0, y.p, 0, z.p, 0)
    np.testing.assert_array_equal(z.x, -Ab.T @ xt.x[1:] + 3 * y.x)
    # beta = 0 ignores y, even a NaN in it
    y.x[:] = np.nan
    oracle.fn("d", "gemv_t")(ps, m, n, 1.0, hb.desc_addr(0), ai, aj, xt.p, 1, 0.0, y.p, 0, z.p, 0)
    np.testing.assert_array_equal(z.x, Ab.T @ xt.x[1:])


@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("m,ai,aj", [(7, 3, 2), (8, 5, 5), (1, 2, 0), (10, 0, 0)])
def test_trmv_exact_in_place_and_out(ps, m, ai, aj):
    rng = np.random.default_rng(m + ps + 17)
    A = _ints(rng, (ai + m + 1, aj + m + 1))
    lay, hb = _mat(A, ps)
    Lb = np.tril(A[ai:ai + m, aj:aj + m])
    x0 = _ints(rng, m)
    for name, want in (("trmv_lnn", Lb @ x0), ("trmv_ltn", Lb.T @ x0)):
        x, z = _Vec(x0), _Vec(np.full(m, -7.0))
        oracle.fn("d", name)(ps, m, hb.desc_addr(0), ai, aj, x.p, 0, z.p, 0)
        np.testing.assert_array_equal(z.x, want, err_msg=name)
        np.testing.assert_array_equal(x.x, x0)
        x = _Vec(x0)
        oracle.fn("d", name)(ps, m, hb.desc_addr(0), ai, aj, x.p, 0, x.p, 0)
        np.testing.assert_array_equal(x.x, want, err_msg=name + " in place")


def _tri_exact(rng, m, n, ai, aj):
    """Integer lower m x n block with dyadic diagonal at (ai, aj) inside a
    larger integer matrix (upper part -3: must never be read)."""
    A = _ints(rng, (ai + m + 1, aj + n + 2))
    for j in range(n):
        A[ai + j, aj + j] = rng.choice([1.0, 2.0, 4.0, 0.5, -2.0, -1.0])
        A[ai + j, aj + j + 1: aj + n] = -3.0
    return A


@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("m,n,ai,aj", [(9, 5, 0, 0), (8, 8, 0, 0), (7, 3, 5, 2), (6, 1, 0, 0), (11, 6, 4, 0)])
def test_trsv_exact(ps, m, n, ai, aj):
    rng = np.random.default_rng(7 * m + n + ps)
    A = _tri_exact(rng, m, n, ai, aj)
    lay, hb = _mat(A, ps)
    B = A[ai:ai + m, aj:aj + n]
    L00 = np.tril(B[:n])
    sol, tail = _ints(rng, n), _ints(rng, m - n)
    # lnn_mn: z(0:n) solves L00 z = x(0:n); z(n:m) = x(n:m) - B(n:m) z(0:n)
    x = _Vec(np.r_[L00 @ sol, tail])
    z = _Vec(np.full(m, -9.0))
    oracle.fn("d", "trsv_lnn_mn")(ps, m, n, hb.desc_addr(0), ai, aj, x.p, 0, z.p, 0)
    np.testing.assert_array_equal(z.x, np.r_[sol, tail - B[n:] @ sol])
    flag = int(hb.desc[0]["use_dA"])
    dA = lay.dview(hb.buf, 0)
    if ai == 0 and aj == 0:
        assert flag == 1
        np.testing.assert_array_equal(dA[:n], 1.0 / np.diag(L00))
    else:
        assert flag == 0
    # ltn_mn: z(0:n) solves L00^T z = x(0:n) - B(n:m)^T x(n:m); z(n:m) = x(n:m), in place
    x = _Vec(np.r_[L00.T @ sol + B[n:].T @ tail, tail])
    oracle.fn("d", "trsv_ltn_mn")(ps, m, n, hb.desc_addr(0), ai, aj, x.p, 0, x.p, 0)
    np.testing.assert_array_equal(x.x, np.r_[sol, tail])


def test_trsv_uses_planted_cache():
    """use_dA = 1 at offset 0: the stored reciprocals are used as they are."""
    rng = np.random.default_rng(3)
    m = n = 6
    A = _tri_exact(rng, m, n, 0, 0)
    lay, hb = _mat(A, 4)
    L = np.tril(A[:m, :n])
    dA = lay.dview(hb.buf, 0)
    dA[:n] = 2.0 / np.diag(L)           # as if the diagonal were halved
    hb.desc[0]["use_dA"] = 1
    Lh = L.copy()
    np.fill_diagonal(Lh, np.diag(L) / 2)
    sol = _ints(rng, n)
    x = _Vec(Lh @ sol)
    oracle.fn("d", "trsv_lnn_mn")(4, m, n, hb.desc_addr(0), 0, 0, x.p, 0, x.p, 0)
    np.testing.assert_array_equal(x.x, sol)
    x = _Vec(Lh.T @ sol)
    oracle.fn("d", "trsv_ltn_mn")(4, m, n, hb.desc_addr(0), 0, 0, x.p, 0, x.p, 0)
    np.testing.assert_array_equal(x.x, sol)
    np.testing.assert_array_equal(dA[:n], 2.0 / np.diag(L))


# ---------------------------------------------------------------------------
# the Riccati solve
# ---------------------------------------------------------------------------
def host_trs(rb, inp, rng):
    """Runs the oracle's trf then trs on a RiccatiBatch in host memory.
    Returns (HostRiccati, RiccatiSolve, solve buffer, inputs copy)."""
    from blasfeo_amd.riccati import HostRiccati, RiccatiSolve
    h = HostRiccati(rb, inp)
    oracle.lib().or_dric_trf_batch(rb.ps, rb.batch, h.argv.ctypes.data)
    rs = RiccatiSolve(rb)
    sbuf = rs.fill_inputs(rng)
    sin = sbuf.copy()
    tab = rs.table(sbuf.ctypes.data)
    babt, rsq, L, work, nx, nu = h.t
    argv = rs.args(tab.ctypes.data, babt.ctypes.data, L.ctypes.data, nx.ctypes.data, nu.ctypes.data)
    oracle.lib().or_dric_trs_batch(rb.ps, rb.batch, argv.ctypes.data)
    return h, rs, sbuf, sin


def kkt_solution(N, nx, nu, BAbt, H, HN, rs, sin, i):
    """Dense KKT solve of instance i: returns (z[0..N], pi[0..N-1])."""
    nk = [nu + nx] * N + [nx]
    C = BAbt.T                                   # nx x (nu+nx)
    nz = sum(nk)
    nv = nz + N * nx
    K = np.zeros((nv, nv))
    r = np.zeros(nv)
    zo = np.concatenate([[0], np.cumsum(nk)])
    po = nz + nx * np.arange(N)
    for k in range(N + 1):
        Hk = H if k < N else HN
        sl = slice(zo[k], zo[k + 1])
        K[sl, sl] = Hk
        r[sl] = -rs.vec(sin, i, "rq", k)
        if k < N:                                 # + C_k^T pi_k
            K[sl, po[k]:po[k] + nx] = C.T
        if k > 0:                                 # - E_k^T pi_{k-1}
            xs = slice(zo[k] + (nk[k] - nx), zo[k + 1])
            K[xs, po[k - 1]:po[k - 1] + nx] -= np.eye(nx)
    for k in range(N):                            # C_k z_k - x_{k+1} = -b_k
        rows = slice(po[k], po[k] + nx)
        K[rows, zo[k]:zo[k + 1]] = C
        K[rows, zo[k + 1] + (nk[k + 1] - nx): zo[k + 2]] -= np.eye(nx)
        r[rows] = -rs.vec(sin, i, "b", k)
    sol = np.linalg.solve(K, r)
    return [sol[zo[k]:zo[k + 1]] for k in range(N + 1)], [sol[po[k]:po[k] + nx] for k in range(N)]


@pytest.mark.parametrize("ps", [4, 8])
def test_riccati_trs_solves_kkt(ps):
    from blasfeo_amd.riccati import RiccatiBatch
    rng = np.random.default_rng(90 + ps)
    nxs, nus, N = [3, 6, 5, 9], [1, 2, 4, 3], 4
    data = {}

    def make(i, nx, nu):
        A = np.eye(nx) + 0.2 * rng.standard_normal((nx, nx))
        B = rng.standard_normal((nx, nu))
        M = rng.standard_normal((nu + nx, nu + nx))
        H = M @ M.T / (nu + nx) + np.eye(nu + nx)          # S != 0
        MN = rng.standard_normal((nx, nx))
        HN = MN @ MN.T / nx + np.eye(nx)
        data[i] = (np.vstack([B.T, A.T]), H, HN)
        return data[i]

    rb = RiccatiBatch(nxs, nus, N=N, ps=ps)
    inp = rb.fill_custom(make)
    h, rs, sbuf, sin = host_trs(rb, inp, rng)
    for i, (nx, nu) in enumerate(zip(nxs, nus)):
        BAbt, H, HN = data[i]
        z, pi = kkt_solution(N, nx, nu, BAbt, H, HN, rs, sin, i)
        for k in range(N + 1):
            np.testing.assert_allclose(rs.vec(sbuf, i, "ux", k), z[k], rtol=1e-10, atol=1e-10,
                                       err_msg="instance %d ux[%d]" % (i, k))
        for k in range(N):
            np.testing.assert_allclose(rs.vec(sbuf, i, "pi", k), pi[k], rtol=1e-10, atol=1e-10,
                                       err_msg="instance %d pi[%d]" % (i, k))
            Lx = np.tril(h.L(i, k + 1))[nu if k + 1 < N else 0:, nu if k + 1 < N else 0:]
            np.testing.assert_allclose(rs.vec(sbuf, i, "Pb", k), Lx @ Lx.T @ rs.vec(sin, i, "b", k),
                                       rtol=1e-12, atol=1e-12)
        # inputs untouched
        for k in range(N):
            np.testing.assert_array_equal(rs.vec(sbuf, i, "b", k), rs.vec(sin, i, "b", k))
        for k in range(N + 1):
            np.testing.assert_array_equal(rs.vec(sbuf, i, "rq", k), rs.vec(sin, i, "rq", k))


def test_riccati_trs_zero_inputs_give_zero():
    """b = 0 and rq = 0: the minimiser is z = 0 and every multiplier is 0."""
    from blasfeo_amd.riccati import RiccatiBatch
    rng = np.random.default_rng(5)
    rb = RiccatiBatch([4, 7], [2, 5], N=3, ps=4)
    inp = rb.fill_inputs(rng)
    from blasfeo_amd.riccati import HostRiccati, RiccatiSolve
    h = HostRiccati(rb, inp)
    oracle.lib().or_dric_trf_batch(4, rb.batch, h.argv.ctypes.data)
    rs = RiccatiSolve(rb)
    sbuf = np.zeros(rs.nbytes, dtype=np.uint8)
    tab = rs.table(sbuf.ctypes.data)
    babt, rsq, L, work, nx, nu = h.t
    argv = rs.args(tab.ctypes.data, babt.ctypes.data, L.ctypes.data, nx.ctypes.data, nu.ctypes.data)
    oracle.lib().or_dric_trs_batch(4, rb.batch, argv.ctypes.data)
    assert not np.any(sbuf.view(np.float64))


@pytest.mark.parametrize("ps", [4, 8])
def test_riccati_sv_matches_trf_trs(ps):
    """The oracle's fused sweep (ric_sv, d_back_ric_sv_libstr) on matrices
    augmented by the linear-term row, against its factorisation + solve
    (ric_trf then ric_trs, d_back_ric_trf_libstr + d_back_ric_trs_libstr)
    on the same problem: the square part of every factor is the same bits
    (the extra row changes no other row's arithmetic), ux and pi agree to
    rounding, and the fused solution satisfies the KKT equations to the
    same backward error (the trs path is pinned by the dense KKT solve above)."""
    from blasfeo_amd.riccati import RiccatiBatch, RiccatiSv, HostRiccati, trs_backward_error
    rng = np.random.default_rng(880 + ps)
    rb = RiccatiBatch([3, 6, 9, 1, 17], [2, 1, 4, 3, 5], N=5, ps=ps, aug=True)
    inp = rb.fill_inputs(rng)
    sv = RiccatiSv(rb)
    rs = sv.rs
    L = oracle.lib()
    # factorisation + solve
    h1 = HostRiccati(rb, inp)
    L.or_dric_trf_batch(ps, rb.batch, h1.argv.ctypes.data)
    sin = sv.solve_inputs(inp)
    sol1 = sin.copy()
    babt, rsq, Lt, work, nx, nu = h1.t
    tab1 = rs.table(sol1.ctypes.data)
    targs = rs.args(tab1.ctypes.data, babt.ctypes.data, Lt.ctypes.data, nx.ctypes.data, nu.ctypes.data)
    L.or_dric_trs_batch(ps, rb.batch, targs.ctypes.data)
    # fused
    h2 = HostRiccati(rb, inp)
    sol2 = np.zeros(rs.nbytes, dtype=np.uint8)
    tab2 = rs.table(sol2.ctypes.data)
    babt2, rsq2, L2, work2, nx2, nu2 = h2.t
    argv = sv.args(tab2.ctypes.data, babt2.ctypes.data, rsq2.ctypes.data, L2.ctypes.data, work2.ctypes.data,
                   nx2.ctypes.data, nu2.ctypes.data)
    L.or_dric_sv_batch(ps, rb.batch, argv.ctypes.data)
    for i in range(rb.batch):
        nk = int(rb.nxs[i] + rb.nus[i])
        for k in range(rb.N + 1):
            n = nk if k < rb.N else int(rb.nxs[i])
            a = np.tril(h1.L(i, k)[:n, :n])
            b = np.tril(h2.L(i, k)[:n, :n])
            np.testing.assert_array_equal(a, b, err_msg="instance %d stage %d" % (i, k))
        for kind, ks in (("ux", range(rb.N + 1)), ("pi", range(rb.N))):
            for k in ks:
                u1, u2 = rs.vec(sol1, i, kind, k), rs.vec(sol2, i, kind, k)
                np.testing.assert_allclose(u2, u1, rtol=1e-9, atol=1e-9 * np.abs(u1).max(),
                                           err_msg="instance %d %s[%d]" % (i, kind, k))
        e1 = trs_backward_error(rb, inp, rs, sin, sol1, i)
        e2 = trs_backward_error(rb, inp, rs, sin, sol2, i)
        assert e2 <= 10 * e1 + 100, (i, e1, e2)
