"""Autotuner for the batched launches.

The reference fixes its block sizes and panel size at build time
(include/blasfeo_block_size.h).  Here the launch shape of the batched entries
is a runtime choice (``blasfeo_amd_tune_set``), and this module picks it by
measuring the caller's own workload on the GPU:

* Riccati sweep (``blasfeo_amd_d_back_ric_trf_batched``): for every populated
  LDS footprint class, the threads per instance (64, 128, 256) and the
  resident instances per CU of its persistent launch, by coordinate descent
  over the classes (largest share of the work first), and the panel size
  (4 or 8) of the batch layout;
* gemm/syrk/trmm: the most workgroups per instance (grid y).

A tuning is a plain dict, saved as JSON and applied with :func:`apply`.
"""
import json

import numpy as np

from . import lib, check

RIC_CAPS = [8192, 16384, 24576, 32768, 40960, 54272, 81920, 163840]
RIC_LDS = 163840
_RT_SZ = 16 * 17


def ric_footprint(nkmax, nxmax):
    """LDS bytes of one instance in the resident Riccati sweep (the kernel's
    rule: tiles, reciprocal pivots and the 16-byte work-queue slot)."""
    nkt = (nkmax + 15) // 16
    nxt = (nxmax + 15) // 16
    return ((nkt * (nkt + 1) // 2 + nkt * nxt) * _RT_SZ + nkt * 16) * 8 + 16


def ric_class(nkmax, nxmax):
    """Footprint class of an instance (0..7), or 8 for the global-memory sweep."""
    if nxmax > 64 or nkmax > 128:
        return 8
    f = ric_footprint(nkmax, nxmax)
    for c, cap in enumerate(RIC_CAPS):
        if f <= cap and (cap > 40960 or nkmax <= 64):
            return c
    return 8


def set_knob(name, index, value):
    check(lib().blasfeo_amd_tune_set(name.encode(), int(index), int(value)), "tune_set %s[%d]" % (name, index))


def get_knob(name, index=0):
    v = lib().blasfeo_amd_tune_get(name.encode(), int(index))
    if v < 0:
        check(v, "tune_get %s[%d]" % (name, index))
    return v


def reset():
    for c in range(8):
        set_knob("ric_per_cu", c, 0)
        set_knob("ric_threads", c, 0)
    set_knob("gemm_grid_y", 0, 0)
    set_knob("potrf_bw", 0, 0)
    set_knob("trsm_thin", 0, 0)
    set_knob("gemm_wave", 0, 0)
    for b in range(8):
        set_knob("lds_threads", b, 0)


def apply(tuning):
    """Sets the knobs of a tuning dict (as returned by the tune_* functions)."""
    for c, v in enumerate(tuning.get("ric_per_cu", [])):
        set_knob("ric_per_cu", c, v)
    for c, v in enumerate(tuning.get("ric_threads", [])):
        set_knob("ric_threads", c, v)
    if "gemm_grid_y" in tuning:
        set_knob("gemm_grid_y", 0, tuning["gemm_grid_y"])
    if "potrf_bw" in tuning:
        set_knob("potrf_bw", 0, tuning["potrf_bw"])


def save(path, tuning):
    with open(path, "w") as f:
        json.dump(tuning, f, indent=1)


def load(path):
    with open(path) as f:
        return json.load(f)


def _time(torch, stream, fn, steps, warmup=1):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def _ric_candidates(cls):
    cap = RIC_CAPS[cls]
    lds_fit = RIC_LDS // cap
    out = []
    for nt in ((64, 128, 256) if cls <= 4 else (128, 256)):
        # resident instances per CU: at most what LDS holds; more workgroups
        # than that only queue behind the resident ones
        waves_fit = 32 // (nt // 64)
        top = max(1, min(lds_fit, waves_fit))
        for pc in sorted({top, max(1, top // 2), max(1, (3 * top) // 4), min(top, 12), min(top, 8), min(top, 4)}):
            out.append((nt, pc))
    return out


def tune_riccati(nxs, nus, N, ps_list=(4, 8), steps=3, seed=1, log=None):
    """Tunes the Riccati launch shape on the batch (nx[i], nu[i], horizon N).
    Returns the tuning dict with the measured ms of the default and tuned
    shapes for each panel size tried; the knobs are left set to the result."""
    import torch
    from .riccati import RiccatiBatch, DeviceRiccati
    nxs = np.asarray(nxs)
    nus = np.asarray(nus)
    classes = np.array([ric_class(int(x + u), int(x)) for x, u in zip(nxs, nus)])
    work = np.zeros(9)
    for c, x, u in zip(classes, nxs, nus):
        nk = x + u
        work[c] += nk * x * x + nk * (nk + 1) * x + nk ** 3 / 3.0
    order = [int(c) for c in np.argsort(-work) if work[c] > 0 and c < 8]
    stream = torch.cuda.Stream()
    rng = np.random.default_rng(seed)
    best_all = None
    report = {}
    for ps in ps_list:
        reset()
        rb = RiccatiBatch(nxs, nus, N, ps)
        dev = DeviceRiccati(rb, rb.fill_inputs(rng))
        fn = lambda: dev.run(stream.cuda_stream)
        base = _time(torch, stream, fn, steps)
        per_cu = [0] * 8
        threads = [0] * 8
        best = base
        for c in order:
            for nt, pc in _ric_candidates(c):
                set_knob("ric_threads", c, nt)
                set_knob("ric_per_cu", c, pc)
                ms = _time(torch, stream, fn, steps)
                if ms < best * 0.995:
                    best, per_cu[c], threads[c] = ms, pc, nt
            set_knob("ric_threads", c, threads[c])
            set_knob("ric_per_cu", c, per_cu[c])
            if log:
                log("tune ps=%d class %d: threads %s per_cu %s -> %.3f ms (default %.3f)"
                    % (ps, c, threads[c] or "auto", per_cu[c] or "auto", best, base))
        report[ps] = {"default_ms": base, "tuned_ms": best}
        cand = {"ps": ps, "ric_per_cu": per_cu, "ric_threads": threads, "ms": best}
        if best_all is None or best < best_all["ms"]:
            best_all = cand
        del dev
    best_all["by_ps"] = report
    best_all["classes_by_work"] = order
    apply(best_all)
    return best_all


def tune_gemm(fn, steps=3, log=None):
    """Picks gemm_grid_y for the workload run by fn() (a zero-argument callable
    that enqueues the batched calls on the current stream)."""
    import torch
    stream = torch.cuda.current_stream()
    set_knob("gemm_grid_y", 0, 0)
    best_v, best = 0, _time(torch, stream, fn, steps)
    for v in (1, 2, 4, 8, 16, 32, 64):
        set_knob("gemm_grid_y", 0, v)
        ms = _time(torch, stream, fn, steps)
        if log:
            log("tune gemm_grid_y=%d: %.3f ms" % (v, ms))
        if ms < best * 0.995:
            best_v, best = v, ms
    set_knob("gemm_grid_y", 0, best_v)
    return {"gemm_grid_y": best_v, "ms": best}
