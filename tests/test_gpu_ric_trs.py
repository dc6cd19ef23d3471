"""GPU parity of the batched Riccati solve (blasfeo_amd_d_back_ric_trs_batched)
against the oracle's restatement of d_back_ric_trs_libstr
(examples/example_d_riccati_recursion.c:203-267).

The oracle runs on the GPU's own factors (copied back), so that only the
solve's arithmetic is compared, at the reference's 1e-11 rule; the
factorisation's parity is covered by test_gpu_riccati.py.  Sizes: nx and
nu + nx above 64 take the second row and column slot of each lane."""
import numpy as np
import pytest

import oracle
from blasfeo_amd import DMAT
from blasfeo_amd.riccati import (RiccatiBatch, RiccatiSolve, DeviceRiccati, DeviceRiccatiSolve,
                                 HostRiccati)
from helpers import assert_close

pytestmark = pytest.mark.gpu


def _oracle_on_gpu_factors(rb, inp, dev, rs, sin, use_dA):
    """The oracle's solve on the factors the GPU computed, with the given
    use_dA flag in every L descriptor; returns (solve buffer, L table)."""
    h = HostRiccati(rb, inp)
    h.outbuf[:] = dev.outbuf.cpu().numpy()
    babt, rsq, L, work, nx, nu = h.t
    L["use_dA"] = use_dA
    sbuf = sin.copy()
    tab = rs.table(sbuf.ctypes.data)
    argv = rs.args(tab.ctypes.data, babt.ctypes.data, L.ctypes.data, nx.ctypes.data, nu.ctypes.data)
    oracle.lib().or_dric_trs_batch(rb.ps, rb.batch, argv.ctypes.data)
    return sbuf, L, h


def _compare(rb, rs, gbuf, rbuf, tol=1e-11):
    N = rb.N
    for i in range(rb.batch):
        for kind, ks in (("ux", range(N + 1)), ("pi", range(N)), ("Pb", range(N))):
            for k in ks:
                assert_close(rs.vec(gbuf, i, kind, k), rs.vec(rbuf, i, kind, k), tol,
                             "instance %d %s[%d]" % (i, kind, k))


def _sizes(rng, n):
    nxs = list(rng.integers(1, 65, n)) + [64, 1, 100, 70]
    nus = list(rng.integers(1, 17, n)) + [16, 1, 20, 30]
    return nxs, nus


@pytest.mark.parametrize("N", [5, 1])
@pytest.mark.parametrize("ps", [4, 8])
def test_riccati_trs_parity(gpu, ps, N):
    """N = 1: the stage-0 forms only (full trsv, all of ux[0] negated)."""
    import torch
    rng = np.random.default_rng(700 + ps + 10 * N)
    nxs, nus = _sizes(rng, 20)
    rb = RiccatiBatch(nxs, nus, N=N, ps=ps)
    inp = rb.fill_inputs(rng)
    dev = DeviceRiccati(rb, inp)
    dev.run()
    rs = RiccatiSolve(rb)
    sin = rs.fill_inputs(rng)
    ds = DeviceRiccatiSolve(dev, rs, sin)
    print("trs parity ps=%d batch=%d N=%d" % (ps, rb.batch, rb.N), flush=True)
    ds.run()
    torch.cuda.synchronize()
    gbuf = ds.host()
    # syrk_dpotrf_ln leaves use_dA = 1 on L[0..N-1], potrf_l use_dA = nx on L[N]
    # (blasfeo_hp_pm/d_lapack_lib4.c:946-949): the solve uses the cached
    # reciprocals and changes no flag
    Lg = dev.t_dev[2].cpu().numpy().view(DMAT).reshape(rb.batch, rb.N + 1)
    assert np.all(Lg["use_dA"][:, :rb.N] == 1)
    np.testing.assert_array_equal(Lg["use_dA"][:, rb.N], nxs)
    rbuf, _, _ = _oracle_on_gpu_factors(rb, inp, dev, rs, sin, use_dA=Lg["use_dA"].reshape(-1))
    _compare(rb, rs, gbuf, rbuf)
    # the inputs are left as they were
    for i in range(rb.batch):
        for k in range(rb.N):
            np.testing.assert_array_equal(rs.vec(gbuf, i, "b", k), rs.vec(sin, i, "b", k))
        for k in range(rb.N + 1):
            np.testing.assert_array_equal(rs.vec(gbuf, i, "rq", k), rs.vec(sin, i, "rq", k))


def test_riccati_trs_forms_reciprocals(gpu):
    """use_dA = 0 in every factor's descriptor and NaN in dA: the solve forms
    1/L(j,j) for the columns its trsv uses (nu[k] on stages k >= 1, all on
    stage 0), stores them, sets use_dA = 1, and its results match the
    oracle's on the same start."""
    import torch
    rng = np.random.default_rng(720)
    nxs, nus = _sizes(rng, 10)
    rb = RiccatiBatch(nxs, nus, N=4, ps=4)
    inp = rb.fill_inputs(rng)
    dev = DeviceRiccati(rb, inp)
    dev.run()
    torch.cuda.synchronize()
    # invalidate the caches on the device: flags in the L table, NaN in dA
    Ltab = dev.t_dev[2]
    Lh = Ltab.cpu().numpy().view(DMAT).copy()
    Lh["use_dA"] = 0
    Ltab.copy_(torch.from_numpy(Lh.view(np.uint8).reshape(Ltab.shape)))
    ob = dev.outbuf.cpu().numpy()
    for i in range(rb.batch):
        for k in range(rb.N + 1):
            o = rb.out.offsets[rb.i_L[i][k]]
            m, n, pm, cn = rb.out.geom(rb.i_L[i][k])
            ob[o + pm * cn * 8: o + (pm * cn + min(m, n)) * 8].view(np.float64)[:] = np.nan
    dev.outbuf.copy_(torch.from_numpy(ob))
    rs = RiccatiSolve(rb)
    sin = rs.fill_inputs(rng)
    ds = DeviceRiccatiSolve(dev, rs, sin)
    print("trs reciprocals batch=%d" % rb.batch, flush=True)
    ds.run()
    torch.cuda.synchronize()
    gbuf = ds.host()
    # L[N] takes no triangular solve: its flag stays as it was
    Lg = Ltab.cpu().numpy().view(DMAT).reshape(rb.batch, rb.N + 1)
    assert np.all(Lg["use_dA"][:, :rb.N] == 1) and np.all(Lg["use_dA"][:, rb.N] == 0)
    # the factors themselves are unchanged by the solve; with use_dA = 0 the
    # oracle forms its own reciprocals over what the GPU stored
    rbuf, Lr, h = _oracle_on_gpu_factors(rb, inp, dev, rs, sin, use_dA=0)
    _compare(rb, rs, gbuf, rbuf)
    np.testing.assert_array_equal(Lr["use_dA"].reshape(rb.batch, rb.N + 1), Lg["use_dA"])
    gob = dev.outbuf.cpu().numpy()
    for i in range(rb.batch):
        for k in range(rb.N):
            ns = nxs[i] + nus[i] if k == 0 else nus[i]
            gd = rb.out.dview(gob, rb.i_L[i][k])[:ns]
            rd = rb.out.dview(h.outbuf, rb.i_L[i][k])[:ns]
            np.testing.assert_array_equal(gd, rd, err_msg="instance %d stage %d dA" % (i, k))
            assert np.all(np.isfinite(gd))


@pytest.mark.parametrize("ps", [4, 8])
def test_riccati_trs_full_horizon_accuracy(gpu, ps):
    """N = 20 on the headline size distribution: the recursion amplifies
    rounding over the horizon (the oracle's own KKT backward error reaches
    1e5 eps on weakly controllable instances), so every instance's KKT
    backward error must be within 10x the oracle's on the same factors
    (+100 eps), and the median at most a few hundred eps."""
    import torch
    from blasfeo_amd.riccati import trs_backward_error
    rng = np.random.default_rng(740 + ps)
    nxs = rng.integers(8, 65, 48)
    nus = rng.integers(2, 17, 48)
    rb = RiccatiBatch(nxs, nus, N=20, ps=ps)
    inp = rb.fill_inputs(rng)
    dev = DeviceRiccati(rb, inp)
    dev.run()
    rs = RiccatiSolve(rb)
    sin = rs.fill_inputs(rng)
    ds = DeviceRiccatiSolve(dev, rs, sin)
    print("trs full horizon batch=%d N=%d" % (rb.batch, rb.N), flush=True)
    ds.run()
    torch.cuda.synchronize()
    gbuf = ds.host()
    Lg = dev.t_dev[2].cpu().numpy().view(DMAT)
    rbuf, _, _ = _oracle_on_gpu_factors(rb, inp, dev, rs, sin, use_dA=Lg["use_dA"].copy())
    gb = [trs_backward_error(rb, inp, rs, sin, gbuf, i) for i in range(rb.batch)]
    ob = [trs_backward_error(rb, inp, rs, sin, rbuf, i) for i in range(rb.batch)]
    for i in range(rb.batch):
        assert gb[i] <= 10 * ob[i] + 100, "instance %d: %.3g eps against the oracle's %.3g" % (i, gb[i], ob[i])
    assert np.median(gb) <= 300
