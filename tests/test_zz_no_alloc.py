"""blasfeo_amd_d_back_ric_trf_batched_noalloc (the allocation-free Riccati
sweep: plan scratch in a fixed device array, classes in order on the
caller's stream) against blasfeo_amd_d_back_ric_trf_batched: two calls in a
row, the second with a smaller batch and a shorter horizon, each equal to
the regular entry's result bit for bit (same kernels, one workgroup per
instance); and batches beyond the scratch's bound are refused."""
import ctypes

import numpy as np
import pytest

import blasfeo_amd as bfa
from blasfeo_amd.riccati import RiccatiBatch, DeviceRiccati

pytestmark = pytest.mark.gpu


def _both(rng, nxs, nus, N):
    import torch
    rb = RiccatiBatch(nxs, nus, N=N, ps=4)
    inp = rb.fill_inputs(rng)
    ref = DeviceRiccati(rb, inp)
    ref.run()
    dev = DeviceRiccati(rb, inp)
    h = bfa.HINT(rb.max_nk, rb.max_nx, 0)
    print("noalloc riccati batch=%d N=%d" % (rb.batch, N), flush=True)
    rc = bfa.lib().blasfeo_amd_d_back_ric_trf_batched_noalloc(4, rb.batch, ctypes.c_void_p(dev.args.data_ptr()),
                                                               ctypes.byref(h), None)
    bfa.check(rc, "d_back_ric_trf_batched_noalloc")
    torch.cuda.synchronize()
    assert torch.equal(dev.outbuf, ref.outbuf)
    # the use_dA of every output descriptor as well (the tables' pointers
    # differ: each sweep writes its own buffer)
    gd = dev.t_dev[2].cpu().numpy().view(bfa.DMAT)["use_dA"]
    rd = ref.t_dev[2].cpu().numpy().view(bfa.DMAT)["use_dA"]
    np.testing.assert_array_equal(gd, rd)


def test_noalloc_matches_batched(gpu):
    rng = np.random.default_rng(4242)
    # first call: every LDS class and the global-memory class (nx = 70)
    nxs = list(rng.integers(1, 65, 300)) + [70]
    nus = list(rng.integers(1, 17, 300)) + [5]
    _both(rng, nxs, nus, N=6)
    # second call: fewer instances, shorter horizon (the scratch keeps the
    # first call's lists; its counts and heads are cleared by the call)
    _both(rng, list(rng.integers(1, 40, 37)), list(rng.integers(1, 9, 37)), N=3)


def test_noalloc_rejects_too_many(gpu):
    h = bfa.HINT(16, 8, 0)
    rc = bfa.lib().blasfeo_amd_d_back_ric_trf_batched_noalloc(4, 65537, ctypes.c_void_p(8), ctypes.byref(h), None)
    assert rc == -4
    assert b"65536" in bfa.lib().blasfeo_amd_last_error()
