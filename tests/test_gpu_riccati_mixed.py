"""GPU checks of the mixed-precision Riccati solve
(blasfeo_amd_d_ric_solve_mixed_batched): FP32 factorisation, the solve on
those factors, FP64 refinement of the KKT solution, FP64 fallback.  Against
the CPU oracle's FP64 d_back_ric_trf + d_back_ric_trs
(examples/example_d_riccati_recursion.c:115-140, 203-267) and against the
library's own FP64 entries."""
import numpy as np
import pytest

import oracle
from blasfeo_amd.riccati import (RiccatiBatch, RiccatiSolve, DeviceRiccati, DeviceRiccatiSolve,
                                 DeviceRiccatiMixed, HostRiccati, trs_backward_error)

pytestmark = pytest.mark.gpu


def _oracle_solution(rb, inp, rs, sin):
    host = HostRiccati(rb, inp)
    oracle.lib().or_dric_trf_batch(rb.ps, rb.batch, host.argv.ctypes.data)
    hsol = sin.copy()
    htab = rs.table(hsol.ctypes.data)
    babt, rsq, L, work, nx, nu = host.t
    targs = rs.args(htab.ctypes.data, babt.ctypes.data, L.ctypes.data, nx.ctypes.data, nu.ctypes.data)
    oracle.lib().or_dric_trs_batch(rb.ps, rb.batch, targs.ctypes.data)
    return hsol


@pytest.mark.parametrize("ps", [4, 8])
def test_riccati_mixed_converges(gpu, ps):
    """nx ~ U{1..64}, nu ~ U{1..16}, N = 6, tol 1e-14: every instance either
    converges on the FP32 factors (path 0, backward error <= tol, at most 3
    corrections) or is redone in FP64 (path 1); ux and pi agree with the
    oracle's FP64 factorisation + solve to 1e-8 of their norms, and the KKT
    backward error recomputed on the host matches the recorded one (1 %, or
    1e-15 where both are rounding noise)."""
    import torch
    rng = np.random.default_rng(800 + ps)
    nxs = list(rng.integers(1, 65, 30)) + [64, 1]
    nus = list(rng.integers(1, 17, 30)) + [16, 1]
    rb = RiccatiBatch(nxs, nus, N=6, ps=ps)
    inp = rb.fill_inputs(rng)
    rs = RiccatiSolve(rb)
    sin = rs.fill_inputs(rng)
    dev = DeviceRiccatiMixed(rb, inp, rs, sin)
    print("riccati mixed ps=%d batch=%d" % (ps, rb.batch), flush=True)
    dev.run(tol=1e-14, max_iter=3)
    torch.cuda.synchronize()
    info = dev.infos()
    sol = dev.vectors()
    ref = _oracle_solution(rb, inp, rs, sin)
    print("paths", np.bincount(info["path"], minlength=2), "iters", np.bincount(info["iters"]),
          "max berr %.2e" % info["berr"].max(), flush=True)
    assert set(np.unique(info["path"])) <= {0, 1}
    assert (info["berr"][info["path"] == 0] <= 1e-14).all()
    assert (info["path"] == 0).mean() >= 0.9, "most instances converge on the FP32 factors"
    for i in range(rb.batch):
        for kind, ks in (("ux", range(rb.N + 1)), ("pi", range(rb.N))):
            for k in ks:
                x, y = rs.vec(sol, i, kind, k), rs.vec(ref, i, kind, k)
                assert np.linalg.norm(x - y) <= 1e-8 * max(np.linalg.norm(y), 1.0), \
                    "%s[%d] instance %d (path %d): %g" % (kind, k, i, info[i]["path"], np.linalg.norm(x - y))
        # (below ~1e-15 both are rounding noise of different summation orders)
        e = trs_backward_error(rb, inp, rs, sin, sol, i) * np.finfo(float).eps
        assert abs(e - info[i]["berr"]) <= 0.01 * e + 1e-15, (i, e, info[i]["berr"])


def test_riccati_mixed_fallback_is_fp64(gpu):
    """tol 0 forces the FP64 fallback: every instance then equals, bit for
    bit, the library's FP64 factorisation + solve on the same inputs."""
    import torch
    rng = np.random.default_rng(820)
    rb = RiccatiBatch(list(rng.integers(1, 65, 24)), list(rng.integers(1, 17, 24)), N=5, ps=4)
    inp = rb.fill_inputs(rng)
    rs = RiccatiSolve(rb)
    sin = rs.fill_inputs(rng)
    dev = DeviceRiccatiMixed(rb, inp, rs, sin)
    dev.run(tol=0.0, max_iter=2)
    fac = DeviceRiccati(rb, inp)
    fac.run()
    ds = DeviceRiccatiSolve(fac, rs, sin)
    ds.run()
    torch.cuda.synchronize()
    info = dev.infos()
    exact = info["berr"] == 0.0
    assert ((info["path"] == 1) | exact).all(), info
    a, b = dev.vectors(), ds.host()
    for i in np.nonzero(~exact)[0]:
        for kind, ks in (("ux", range(rb.N + 1)), ("pi", range(rb.N))):
            for k in ks:
                assert np.array_equal(rs.vec(a, i, kind, k), rs.vec(b, i, kind, k)), (i, kind, k)
    assert np.array_equal(dev.fac.outbuf.cpu().numpy(), fac.outbuf.cpu().numpy()) or exact.any()


def test_riccati_mixed_horizon20(gpu):
    """The Riccati workload's shapes (nx ~ U{8..64}, nu ~ U{2..16}, N = 20),
    2048 instances, tol 1e-13: path 0 instances at or below tol; the FP64
    fallback instances carry the FP64 solution's backward error, which is
    the oracle's own on weakly controllable instances (within 10x + 100 eps)."""
    import torch
    rng = np.random.default_rng(840)
    batch = 2048
    rb = RiccatiBatch(rng.integers(8, 65, batch), rng.integers(2, 17, batch), N=20, ps=4)
    inp = rb.fill_inputs(rng)
    rs = RiccatiSolve(rb)
    sin = rs.fill_inputs(rng)
    dev = DeviceRiccatiMixed(rb, inp, rs, sin)
    print("riccati mixed batch=%d N=20" % batch, flush=True)
    dev.run(tol=1e-13, max_iter=3)
    torch.cuda.synchronize()
    info = dev.infos()
    sol = dev.vectors()
    print("paths", np.bincount(info["path"], minlength=2), "iters", np.bincount(info["iters"]),
          "berr median %.2e max %.2e" % (np.median(info["berr"]), info["berr"].max()), flush=True)
    assert (info["berr"][info["path"] == 0] <= 1e-13).all()
    fb = np.nonzero(info["path"] == 1)[0][:64]
    if len(fb):
        ref = _oracle_solution(rb, inp, rs, sin)
        eps = np.finfo(float).eps
        for i in fb:
            eo = trs_backward_error(rb, inp, rs, sin, ref, i)
            assert info[i]["berr"] / eps <= 10 * eo + 100, (i, info[i]["berr"] / eps, eo)
