"""Pins the oracle's LU-solve pieces (oracle_impl.h trsm_llnu, trsm_lunn,
rowpe; blasfeo_ref/x_blas3_ref.c:572, :920; auxiliary/d_aux_lib4.c:2767):
exact integer solves at panel-crossing offsets (the unit diagonal and the
triangle a solve must not read hold values that would change the answer),
the dA / use_dA rules, the row swaps against numpy, and the composition of
getrf_rp, rowpe, trsm_llnu and trsm_lunn solving A X = B
(examples/example_d_lu_factorization.c:163-171)."""
import numpy as np
import pytest

import blasfeo_amd as bfa
import oracle
from blasfeo_amd.device import HostBatch


def _hb(mats, ps):
    lay = bfa.Layout(np.float64, ps)
    for A in mats:
        lay.add(*np.shape(A))
    buf = lay.host_buffer()
    for i, A in enumerate(mats):
        lay.put(buf, i, A)
    return lay, HostBatch(lay, buf)


@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("m,n,ai,aj", [(9, 3, 0, 0), (8, 5, 5, 2), (1, 2, 3, 0), (13, 1, 0, 0)])
def test_trsm_llnu_lunn_exact(ps, m, n, ai, aj):
    rng = np.random.default_rng(m * 10 + n + ps)
    A = rng.integers(-3, 4, (ai + m, aj + m)).astype(float)
    Lu = np.tril(A[ai:, aj:], -1) + np.eye(m)             # unit lower part
    U = np.triu(A[ai:, aj:], 1)
    dg = rng.choice([1.0, 2.0, 4.0, -0.5, 0.5], m)
    for t in range(m):
        A[ai + t, aj + t] = dg[t]
    U = U + np.diag(dg)
    X = rng.integers(-4, 5, (m, n)).astype(float)
    lay, hb = _hb([A, np.zeros((m, n)), np.full((m, n), -1.0)], ps)
    # llnu: D = Lu^{-1} (2 Lu X) / 2 with alpha = 0.5 -> X; A's diagonal must not be read
    lay.put(hb.buf, 1, 2 * Lu @ X)
    hb.desc[0]["use_dA"] = 0
    oracle.fn("d", "trsm_llnu")(ps, m, n, 0.5, hb.desc_addr(0), ai, aj, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0)
    np.testing.assert_array_equal(lay.get(hb.buf, 2), X)
    assert hb.desc[0]["use_dA"] == 0 and hb.desc[2]["use_dA"] == 0
    # lunn: D = U^{-1} (U X) -> X; the strictly lower part must not be read
    lay.put(hb.buf, 1, U @ X)
    oracle.fn("d", "trsm_lunn")(ps, m, n, 1.0, hb.desc_addr(0), ai, aj, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0)
    np.testing.assert_array_equal(lay.get(hb.buf, 2), X)
    if ai == 0 and aj == 0:
        assert hb.desc[0]["use_dA"] == m
        np.testing.assert_array_equal(lay.dview(hb.buf, 0)[:m], 1.0 / dg)
    else:
        assert hb.desc[0]["use_dA"] == 0


def test_rowpe_matches_sequential_swaps():
    rng = np.random.default_rng(2)
    for ps in (4, 8):
        m, n = 11, 6
        A = rng.standard_normal((m, n))
        ipiv = np.array([3, 1, 10, 4, 9, 5, 7, 9, 8, 10], dtype=np.int32)  # repeats allowed
        lay, hb = _hb([A], ps)
        hb.desc[0]["use_dA"] = 4
        oracle.fn("d", "rowpe")(ps, len(ipiv), ipiv.ctypes.data, hb.desc_addr(0))
        want = A.copy()
        for i, p in enumerate(ipiv):
            want[[i, p]] = want[[p, i]]
        np.testing.assert_array_equal(lay.get(hb.buf, 0), want)
        assert hb.desc[0]["use_dA"] == 0


@pytest.mark.parametrize("ps", [4, 8])
def test_lu_solve_composition(ps):
    """getrf_rp + rowpe + trsm_llnu + trsm_lunn solve A X = B."""
    rng = np.random.default_rng(70 + ps)
    for n, nrhs in ((1, 1), (7, 3), (33, 4), (64, 9)):
        A = rng.uniform(-1, 1, (n, n))
        B = rng.standard_normal((n, nrhs))
        lay, hb = _hb([A, np.zeros((n, n)), B], ps)
        ipiv = np.zeros(n, dtype=np.int32)
        oracle.fn("d", "getrf_rp")(ps, n, n, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0, ipiv.ctypes.data)
        oracle.fn("d", "rowpe")(ps, n, ipiv.ctypes.data, hb.desc_addr(2))
        oracle.fn("d", "trsm_llnu")(ps, n, nrhs, 1.0, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0, hb.desc_addr(2), 0, 0)
        oracle.fn("d", "trsm_lunn")(ps, n, nrhs, 1.0, hb.desc_addr(1), 0, 0, hb.desc_addr(2), 0, 0, hb.desc_addr(2), 0, 0)
        X = lay.get(hb.buf, 2)
        r = np.abs(A @ X - B).max() / (np.abs(A).sum(1).max() * np.abs(X).max() + np.abs(B).max())
        assert r < 20 * n * np.finfo(float).eps, (n, r)


@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("m,n,ci,di", [(6, 6, 0, 0), (9, 5, 3, 0), (5, 9, 0, 2), (1, 1, 0, 0), (13, 13, 5, 6)])
def test_getrf_np_exact(ps, m, n, ci, di):
    """C = L U with integer unit-lower L and integer upper U with dyadic
    diagonal: every step is exact, so getrf_np returns L and U bit for bit,
    dA = 1/diag(U), and use_dA is 1 only at offset (0, 0)."""
    rng = np.random.default_rng(m * 7 + n + ps + ci)
    p = min(m, n)
    L = np.tril(rng.integers(-2, 3, (m, p)).astype(float), -1) + np.eye(m, p)
    U = np.triu(rng.integers(-3, 4, (p, n)).astype(float), 1)
    dg = rng.choice([1.0, 2.0, -4.0, 0.5], p)
    U[np.arange(p), np.arange(p)] = dg
    C = L @ U
    Cb = np.full((ci + m, n + 1), 7.0)
    Cb[ci:ci + m, :n] = C
    Db = np.full((di + m, n + 2), -1.0)
    lay, hb = _hb([Cb, Db], ps)
    oracle.fn("d", "getrf_np")(ps, m, n, hb.desc_addr(0), ci, 0, hb.desc_addr(1), di, 0)
    D = lay.get(hb.buf, 1)
    got = D[di:di + m, :n]
    np.testing.assert_array_equal(np.triu(got)[:p], np.triu(U)[:p])
    np.testing.assert_array_equal(np.tril(got, -1)[:, :p], np.tril(L, -1)[:, :p])
    np.testing.assert_array_equal(D[di:di + m, n:], -1.0)      # outside the block untouched
    assert hb.desc[1]["use_dA"] == (1 if di == 0 else 0)
    np.testing.assert_array_equal(lay.dview(hb.buf, 1)[:p], 1.0 / dg)


def test_getrf_np_zero_pivot_gives_inf():
    lay, hb = _hb([np.array([[0.0, 1.0], [1.0, 1.0]]), np.zeros((2, 2))], 4)
    with np.errstate(all="ignore"):
        oracle.fn("d", "getrf_np")(4, 2, 2, hb.desc_addr(0), 0, 0, hb.desc_addr(1), 0, 0)
    assert np.isinf(lay.dview(hb.buf, 1)[0])
