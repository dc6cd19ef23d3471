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


def test_riccati_graph_replay_after_growth(gpu):
    """A graph captured on a stream, then an eager sweep of a larger batch on
    the same stream (its plan needs a larger workspace, which replaces the
    stream's block), then the graph replayed: the replay still equals the
    first eager result, since a replaced workspace block stays allocated."""
    import torch
    rng = np.random.default_rng(650)
    small = RiccatiBatch(list(rng.integers(1, 65, 16)), list(rng.integers(1, 17, 16)), N=4, ps=4)
    sinp = small.fill_inputs(rng)
    nbig = 40000
    big = RiccatiBatch(list(rng.integers(1, 9, nbig)), list(rng.integers(1, 4, nbig)), N=1, ps=4)
    binp = big.fill_inputs(rng)
    dsmall = DeviceRiccati(small, sinp)
    dbig = DeviceRiccati(big, binp)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        dsmall.run(s.cuda_stream)
    torch.cuda.synchronize()
    ref = dsmall.outbuf.clone()
    g = dsmall.capture(s)
    torch.cuda.synchronize()
    print("riccati eager batch=%d after capture" % nbig, flush=True)
    with torch.cuda.stream(s):
        dbig.run(s.cuda_stream)
    torch.cuda.synchronize()
    for rep in range(2):
        dsmall.outbuf.zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            g.replay()
        torch.cuda.synchronize()
        assert torch.equal(dsmall.outbuf, ref), "replay %d after the larger eager call differs" % rep


def test_riccati_horizon20_full(gpu):
    """The Riccati workload at full size: 8192 instances, horizon N = 20,
    nx ~ U{8..64}, nu ~ U{2..16}, A = I + 0.1 N(0,1), B ~ N(0,1), Q = R = I,
    one batched call.  Every stage of every instance has backward error
    ||M_k - L_k L_k^T||_F <= 10 n_k eps ||M_k||_F with M_k rebuilt from the
    computed L_{k+1}; the first 32 instances' stage-N and stage-(N-1)
    factors (before the recursion amplifies rounding) equal the oracle's at
    the 1e-11 rule."""
    import torch
    from blasfeo_amd.riccati import stage_backward_errors
    rng = np.random.default_rng(1234)
    batch = 8192
    nxs = rng.integers(8, 65, batch)
    nus = rng.integers(2, 17, batch)
    rb = RiccatiBatch(nxs, nus, N=20, ps=4)
    inp = rb.fill_inputs(rng)
    dev = DeviceRiccati(rb, inp)
    print("riccati batch=%d N=20" % batch, flush=True)
    dev.run()
    torch.cuda.synchronize()
    gbuf = dev.outbuf.cpu().numpy()
    e = stage_backward_errors(rb, inp, gbuf, range(batch))
    print("worst stage backward error %.3f n eps" % e.max(), flush=True)
    assert e.max() <= 10.0, "instance %d stage %d: %.2f n eps" % (np.unravel_index(e.argmax(), e.shape) + (e.max(),))
    sub = RiccatiBatch(nxs[:32], nus[:32], N=20, ps=4)
    sinp = sub.inp.host_buffer()
    for i in range(32):
        for a, b in ((sub.i_babt[i], rb.i_babt[i]), (sub.i_rsq[i], rb.i_rsq[i]), (sub.i_qn[i], rb.i_qn[i])):
            sub.inp.view(sinp, a)[:] = rb.inp.view(inp, b)
    host = HostRiccati(sub, sinp)
    oracle.lib().or_dric_trf_batch(4, 32, host.argv.ctypes.data)
    for i in range(32):
        for k in (20, 19):
            g = np.tril(rb.out.get(gbuf, rb.i_L[i][k]))
            assert_close(g, np.tril(host.L(i, k)), 1e-11, "instance %d stage %d" % (i, k))
