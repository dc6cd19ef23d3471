"""Pins the oracle's column-major <-> panel-major conversions (oracle_impl.h
pack / pack_tran / unpack, restating auxiliary/d_aux_lib4.c:1607, :2086,
:2238): the panel-major result equals the dense matrix laid out by the
panel-major addressing of include/blasfeo_common.h:112 through numpy
reshapes (an independent statement of the layout), the library's own host
blasfeo_pack_dmat writes the same bytes, unpack inverts pack, pack_tran of
A is pack of A^T, untouched elements keep their sentinel, and pack clears
use_dA."""
import ctypes

import numpy as np
import pytest

import blasfeo_amd as bfa
import oracle
from pack_cases import SENTINEL, make_cases, records, descriptors


def _host(c):
    cm = np.full(c["cm_len"], np.nan)
    pm = np.full(c["pm_bytes"] // 8, SENTINEL)
    desc = descriptors(c, pm.ctypes.data)
    desc["use_dA"] = 5
    rec = records(c, cm.ctypes.data, desc.ctypes.data)
    return cm, pm, desc, rec


def _dense_expect(c, i, A):
    """Panel-major bytes of instance i's destination after the pack."""
    ps = c["ps"]
    rows, cols = int(c["ais"][i] + c["R"][i] + 3), int(c["ajs"][i] + c["C"][i] + 2)
    pm = -(-rows // ps) * ps
    cn = -(-cols // (8 if ps == 8 else 4)) * (8 if ps == 8 else 4)
    D = np.full((pm, cn), SENTINEL)            # padding rows and columns included
    ai, aj = int(c["ais"][i]), int(c["ajs"][i])
    D[ai:ai + A.shape[0], aj:aj + A.shape[1]] = A
    return bfa.pm_pack(D, ps, cn)


@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("tran", [False, True])
def test_pack_layout_and_roundtrip(ps, tran):
    rng = np.random.default_rng(30 + ps + 2 * tran)
    c = make_cases(rng, 12, ps, tran, maxdim=40)
    cm, pm, desc, rec = _host(c)
    for i in range(c["batch"]):
        m, n, lda, o = int(c["ms"][i]), int(c["ns"][i]), int(c["ldas"][i]), int(c["cm_off"][i])
        cm[o:o + lda * n] = rng.standard_normal(lda * n)
    cm0 = cm.copy()
    L = oracle.lib()
    (L.or_dpack_tran_batch if tran else L.or_dpack_batch)(ps, c["batch"], rec.ctypes.data)
    assert np.all(desc["use_dA"] == 0)
    np.testing.assert_array_equal(cm, cm0)
    for i in range(c["batch"]):
        m, n, lda, o = int(c["ms"][i]), int(c["ns"][i]), int(c["ldas"][i]), int(c["cm_off"][i])
        A = cm[o:o + lda * n].reshape(n, lda).T[:m]
        want = _dense_expect(c, i, A.T if tran else A)
        p0 = (int(desc["pA"][i]) - pm.ctypes.data) // 8
        np.testing.assert_array_equal(pm[p0:p0 + len(want)], want, err_msg="instance %d" % i)
    # unpack of the same block returns it (for tran: the transposed block)
    back = np.full_like(cm, np.nan)
    rec2 = rec.copy()
    if tran:
        # the panel-major block is n x m: unpack it into an n x m column-major array
        rec2["m"], rec2["n"] = c["ns"], c["ms"]
        rec2["lda"] = c["ns"]
    rec2["A"] = np.uint64(back.ctypes.data) + (8 * c["cm_off"]).astype(np.uint64)
    L.or_dunpack_batch(ps, c["batch"], rec2.ctypes.data)
    for i in range(c["batch"]):
        m, n, lda, o = int(c["ms"][i]), int(c["ns"][i]), int(c["ldas"][i]), int(c["cm_off"][i])
        A = cm[o:o + lda * n].reshape(n, lda).T[:m]
        if tran:
            got = back[o:o + n * m].reshape(m, n).T
            np.testing.assert_array_equal(got, A.T)
        else:
            got = back[o:o + lda * n].reshape(n, lda).T[:m]
            np.testing.assert_array_equal(got, A)


def test_pack_matches_host_library():
    """The library's host blasfeo_pack_dmat (a separate implementation, host
    memory) writes the same panel-major bytes as the oracle."""
    lib = bfa.lib()
    rng = np.random.default_rng(44)
    for ps in (4, 8):
        assert lib.blasfeo_amd_set_ps(ps) == 0
        m, n, ai, aj, lda = 13, 9, 3, 2, 15
        A = rng.standard_normal((lda, n))
        Af = np.asfortranarray(A)
        lay = bfa.Layout(np.float64, ps)
        lay.add(ai + m + 2, aj + n + 1)
        b1, b2 = lay.host_buffer(), lay.host_buffer()
        d1, d2 = lay.build(b1.ctypes.data), lay.build(b2.ctypes.data)
        lib.blasfeo_pack_dmat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        lib.blasfeo_pack_dmat(m, n, Af.ctypes.data, lda, d1.ctypes.data, ai, aj)
        oracle.fn("d", "pack")(ps, m, n, Af.ctypes.data, lda, d2.ctypes.data, ai, aj)
        np.testing.assert_array_equal(b1, b2)
    lib.blasfeo_amd_set_ps(4)


def test_pm_dgecp_between_panel_sizes():
    """or_pm_dgecp: every (psA, psB) pair against numpy layouts of the same
    dense blocks, sentinels outside the block kept, use_dA cleared, and a
    4 -> 8 -> 4 round trip returning the original bytes."""
    from pack_cases import pm_layout
    rng = np.random.default_rng(61)
    for pa, pb in ((4, 4), (4, 8), (8, 4), (8, 8)):
        m, n, ai, aj, bi, bj = 13, 9, 3, 2, 5, 1
        A = rng.standard_normal((ai + m + 2, aj + n))
        d, offs, nb = pm_layout([pa, pb], [ai + m + 2, bi + m + 1], [aj + n, bj + n + 3], 0)
        buf = np.full(nb // 8, SENTINEL)
        d, offs, nb = pm_layout([pa, pb], [ai + m + 2, bi + m + 1], [aj + n, bj + n + 3], buf.ctypes.data)
        buf[offs[0] // 8: offs[0] // 8 + int(d["pm"][0] * d["cn"][0])] = bfa.pm_pack(
            np.pad(A, ((0, int(d["pm"][0]) - A.shape[0]), (0, int(d["cn"][0]) - A.shape[1]))), pa, int(d["cn"][0]))
        d["use_dA"][1] = 3
        oracle.lib().or_pm_dgecp(m, n, d.ctypes.data, ai, aj, d.ctypes.data + bfa.PM_DMAT.itemsize, bi, bj)
        assert d["use_dA"][1] == 0
        E = np.full((int(d["pm"][1]), int(d["cn"][1])), SENTINEL)
        E[bi:bi + m, bj:bj + n] = A[ai:ai + m, aj:aj + n]
        got = buf[offs[1] // 8: offs[1] // 8 + E.size]
        np.testing.assert_array_equal(got, bfa.pm_pack(E, pb, int(d["cn"][1])), err_msg="%d->%d" % (pa, pb))
    # 4 -> 8 -> 4 round trip of a whole matrix
    m, n = 21, 10
    d, offs, nb = pm_layout([4, 8, 4], [m] * 3, [n] * 3, 0)
    buf = np.zeros(nb // 8)
    d, offs, nb = pm_layout([4, 8, 4], [m] * 3, [n] * 3, buf.ctypes.data)
    ln = int(d["pm"][0] * d["cn"][0])
    buf[:ln] = rng.standard_normal(ln)
    sz = bfa.PM_DMAT.itemsize
    oracle.lib().or_pm_dgecp(m, n, d.ctypes.data, 0, 0, d.ctypes.data + sz, 0, 0)
    oracle.lib().or_pm_dgecp(m, n, d.ctypes.data + sz, 0, 0, d.ctypes.data + 2 * sz, 0, 0)
    a0 = bfa.pm_unpack(buf[offs[0] // 8:], 4, int(d["cn"][0]), m, n)
    a2 = bfa.pm_unpack(buf[offs[2] // 8:], 4, int(d["cn"][2]), m, n)
    np.testing.assert_array_equal(a0, a2)


@pytest.mark.parametrize("tran", [False, True])
def test_pack_empty_block_cache(tran):
    """An empty block (m = 0 or n = 0): blasfeo_pack_dmat returns before it
    touches the cache (auxiliary/d_aux_lib4.c:1609), blasfeo_pack_tran_dmat
    clears use_dA first (:2090); nothing else is written.  Checked for the
    oracle and for the library's host blasfeo_pack(_tran)_dmat."""
    rng = np.random.default_rng(77 + tran)
    c = make_cases(rng, 6, 4, tran, maxdim=20, empties=(1, 2, 3, 4))
    cm, pm, desc, rec = _host(c)
    pm0 = pm.copy()
    L = oracle.lib()
    (L.or_dpack_tran_batch if tran else L.or_dpack_batch)(4, c["batch"], rec.ctypes.data)
    empty = (c["ms"] == 0) | (c["ns"] == 0)
    assert empty.sum() == 4
    assert np.all(desc["use_dA"][empty] == (0 if tran else 5))
    assert np.all(desc["use_dA"][~empty] == 0)
    # the empty instances wrote nothing: only the non-empty ones' blocks changed
    for i in np.nonzero(empty)[0]:
        p0 = (int(desc["pA"][i]) - pm.ctypes.data) // 8
        n = int(desc["memsize"][i]) // 8
        np.testing.assert_array_equal(pm[p0:p0 + n], pm0[p0:p0 + n])
    # the library's host routine follows the same rule
    lib = bfa.lib()
    lib.blasfeo_amd_set_ps(4)
    d = np.zeros(1, dtype=bfa.DMAT)
    buf = np.zeros(bfa.memsize(np.float64, 4, 8, 8) // 8)
    lib.blasfeo_create_dmat(8, 8, ctypes.c_void_p(d.ctypes.data), ctypes.c_void_p(buf.ctypes.data))
    d["use_dA"] = 3
    A = np.ones(4)
    f = lib.blasfeo_pack_tran_dmat if tran else lib.blasfeo_pack_dmat
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    f(0, 4, A.ctypes.data, 1, d.ctypes.data, 0, 0)
    assert int(d["use_dA"][0]) == (0 if tran else 3)
    assert not buf.any()
