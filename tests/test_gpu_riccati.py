"""GPU parity of the fused backward Riccati factorisation against the CPU
oracle's restatement of d_back_ric_trf_libstr
(examples/example_d_riccati_recursion.c:115-140), plus the closed form
L_k = blkdiag(I, sqrt(N-k+1) I) for A = I, B = 0, Q = R = I."""
import numpy as np
import pytest

import oracle
from blasfeo_amd.riccati import RiccatiBatch, DeviceRiccati, HostRiccati
from helpers import assert_close

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["lds", "general"])
@pytest.mark.parametrize("ps", [4, 8])
def test_riccati_random(gpu, ps, form):
    """'lds': nx <= 64, nu + nx <= 128 (LDS-resident sweep); 'general': one
    instance with nx = 70 selects the global-memory sweep."""
    import torch
    rng = np.random.default_rng(600 + ps)
    nxs = list(rng.integers(1, 65, 14)) + [64, 1]
    nus = list(rng.integers(1, 17, 14)) + [16, 1]
    if form == "general":
        nxs[3], nus[3] = 70, 5
    rb = RiccatiBatch(nxs, nus, N=6, ps=ps)
    inp = rb.fill_inputs(rng)
    dev = DeviceRiccati(rb, inp)
    dev.run()
    torch.cuda.synchronize()
    host = HostRiccati(rb, inp)
    oracle.lib().or_dric_trf_batch(ps, rb.batch, host.argv.ctypes.data)
    gbuf = dev.outbuf.cpu().numpy()
    for i in range(rb.batch):
        for k in range(rb.N + 1):
            g = rb.out.get(gbuf, rb.i_L[i][k])
            r = host.L(i, k)
            n = g.shape[0]
            assert_close(np.tril(g), np.tril(r), 1e-11, "instance %d stage %d" % (i, k))
            gd = rb.out.dview(gbuf, rb.i_L[i][k])
            rd = rb.out.dview(host.outbuf, rb.i_L[i][k])
            assert_close(gd, rd, 1e-11, "instance %d stage %d dA" % (i, k))


def test_riccati_closed_form(gpu):
    import torch
    N = 7
    nxs, nus = [5, 20, 33], [2, 3, 9]
    rb = RiccatiBatch(nxs, nus, N=N, ps=4)
    inp = rb.fill_custom(lambda i, nx, nu: (np.vstack([np.zeros((nu, nx)), np.eye(nx)]),
                                            np.eye(nu + nx), np.eye(nx)))
    dev = DeviceRiccati(rb, inp)
    dev.run()
    torch.cuda.synchronize()
    for i, (nx, nu) in enumerate(zip(nxs, nus)):
        for k in range(N):
            expect = np.diag(np.r_[np.ones(nu), np.full(nx, np.sqrt(N - k + 1))])
            np.testing.assert_allclose(np.tril(dev.L(i, k)), expect, rtol=1e-15, atol=1e-15)
        np.testing.assert_allclose(np.tril(dev.L(i, N)), np.eye(nx), rtol=0, atol=0)


def test_riccati_graph_replay(gpu):
    """The whole batched sweep captured in a HIP graph (zero-fill, plan and
    class launches; no memset node): replayed three times with eager work and
    synchronisation between replays, every replay equals the eager result bit
    for bit."""
    import torch
    rng = np.random.default_rng(640)
    nxs = list(rng.integers(1, 65, 24))
    nus = list(rng.integers(1, 17, 24))
    rb = RiccatiBatch(nxs, nus, N=5, ps=4)
    inp = rb.fill_inputs(rng)
    dev = DeviceRiccati(rb, inp)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        dev.run(s.cuda_stream)
    torch.cuda.synchronize()
    ref = dev.outbuf.clone()
    g = dev.capture(s)
    torch.cuda.synchronize()
    for rep in range(3):
        dev.outbuf.zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            g.replay()
        torch.cuda.synchronize()
        assert torch.equal(dev.outbuf, ref), "replay %d differs from the eager sweep" % rep


@pytest.mark.parametrize("ps", [4, 8])
def test_riccati_tuned_shapes(gpu, ps):
    """The launch shapes the tuning knobs select besides the defaults: one
    wave per instance (k_ric_class1 and k_ric_class<ps,64>: out-of-line stage
    calls, guarded partial panels) for the classes up to 40 KB, and four waves
    at one instance per CU for the small classes; every stage's factor and dA
    against the oracle at the 1e-11 rule."""
    import torch
    from blasfeo_amd import autotune
    rng = np.random.default_rng(660 + ps)
    # nu + nx <= 64 throughout, so every instance lands in a class up to 40 KB
    nxs = list(rng.integers(1, 49, 20)) + [48, 1, 16]
    nus = list(rng.integers(1, 17, 20)) + [16, 1, 15]
    rb = RiccatiBatch(nxs, nus, N=5, ps=ps)
    inp = rb.fill_inputs(rng)
    host = HostRiccati(rb, inp)
    oracle.lib().or_dric_trf_batch(ps, rb.batch, host.argv.ctypes.data)
    shapes = [(64, 12), (64, 3), (256, 1)]
    try:
        for nt, pc in shapes:
            autotune.reset()
            for c in range(5):
                autotune.set_knob("ric_threads", c, nt)
                autotune.set_knob("ric_per_cu", c, pc)
            dev = DeviceRiccati(rb, inp)
            print("riccati ps=%d threads=%d per_cu=%d" % (ps, nt, pc), flush=True)
            dev.run()
            torch.cuda.synchronize()
            gbuf = dev.outbuf.cpu().numpy()
            for i in range(rb.batch):
                for k in range(rb.N + 1):
                    g = rb.out.get(gbuf, rb.i_L[i][k])
                    r = host.L(i, k)
                    assert_close(np.tril(g), np.tril(r), 1e-11, "threads %d instance %d stage %d" % (nt, i, k))
                    assert_close(rb.out.dview(gbuf, rb.i_L[i][k]), rb.out.dview(host.outbuf, rb.i_L[i][k]),
                                 1e-11, "threads %d instance %d stage %d dA" % (nt, i, k))
    finally:
        autotune.reset()
