#!/usr/bin/env python
"""Benchmark of the batched BLASFEO hot path on one MI355X.

Workload (default): the batched backward Riccati factorisation of
examples/example_d_riccati_recursion.c:115-140 -- 8192 independent
instances, horizon N = 20, nx ~ U{8..64}, nu ~ U{2..16} per instance,
A = I + 0.1 N(0,1), B ~ N(0,1), Q = R = I, FP64, panel size 4, all inputs
resident in GPU memory.  One step = one launch of the fused sweep over the
whole batch.  ``value`` is solves (complete factorisations) per second.

The JSON line also carries
  roofline:     the sweep kernel's achieved FP64 rate (algorithmic flops per
                launch / average launch time from HIP events on the launch
                stream) against MI355X's 78.6 TFLOP/s FP64 peak, with the HBM
                traffic per launch measured by rocprofv3 (profiles/hbm_traffic.txt);
  cpu_baseline: the CPU oracle (a C restatement of the reference's REFERENCE
                backend, OpenMP over instances) timed on this host on a bounded
                sample of the same instances;
  checks:       backward check of a sample of instances against the oracle.

Flop conventions are the reference's (benchmarks/x_benchmark_blasfeo_api.c:557-566):
trmm m*n^2, syrk m(m+1)k, potrf n^3/3.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/sec + %FP64-MFMA peak, batched potrf/Riccati; mixed vs FP64 at equal error"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def riccati_flops(nxs, nus, N):
    nxs = np.asarray(nxs, dtype=np.float64)
    nus = np.asarray(nus, dtype=np.float64)
    nk = nxs + nus
    stage = nk * nxs ** 2 + nk * (nk + 1) * nxs + nk ** 3 / 3.0
    return float((N * stage + nxs ** 3 / 3.0).sum())


def riccati_hbm_bytes(nxs, nus, N):
    """Algorithmic HBM bytes of one launch: time-invariant inputs read once per
    instance (BAbt, lower RSQ, lower Q_N) and every factor written once (lower
    triangle + reciprocal diagonal), 8 bytes each."""
    nxs = np.asarray(nxs, dtype=np.float64)
    nus = np.asarray(nus, dtype=np.float64)
    nk = nxs + nus
    rd = nk * nxs + nk * (nk + 1) / 2 + nxs * (nxs + 1) / 2
    wr = N * (nk * (nk + 1) / 2 + nk) + nxs * (nxs + 1) / 2 + nxs
    return float(8 * (rd + wr).sum())


def roof(flops, algo_bytes, ms, fp_peak):
    """The roof a kernel is held to, from its algorithmic arithmetic
    intensity against the ridge point fp_peak / HBM peak: (bound, achieved,
    peak, unit), achieved in TFLOP/s for the matrix-core roof and in GB/s
    (algorithmic bytes) for the HBM roof."""
    ridge = fp_peak * 1e12 / (HBM_PEAK_GBS * 1e9)
    if flops / algo_bytes >= ridge:
        return "mfma", flops / (ms * 1e-3) / 1e12, fp_peak, "TFLOP/s"
    return "hbm", algo_bytes / (ms * 1e-3) / 1e9, HBM_PEAK_GBS, "GB/s"


def read_traffic(kernel, workload):
    """HBM bytes per launch from profiles/hbm_traffic.txt (rocprofv3 FETCH_SIZE
    and WRITE_SIZE runs, scaled by a copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 3 and f[0] == kernel and f[1] == workload:
            for tok in f[2:]:
                if tok.startswith("bytes_per_launch="):
                    return float(tok.split("=")[1])
    return None


def riccati_checks(rb, inp, gbuf, host, nchk):
    """Worst per-stage backward error of the first nchk instances (in units
    of n_k eps, blasfeo_amd.riccati.stage_backward_errors) and the forward
    difference of L_0 against the oracle's."""
    from blasfeo_amd.riccati import stage_backward_errors
    worst = float(stage_backward_errors(rb, inp, gbuf, range(nchk)).max())
    fdiff = 0.0
    for i in range(nchk):
        g0 = np.tril(rb.out.get(gbuf, rb.i_L[i][0]))
        r0 = np.tril(host.L(i, 0))
        fdiff = max(fdiff, float(np.abs(g0 - r0).max() / np.abs(r0).max()))
    return worst, fdiff


def cpu_baseline_riccati(rb, inp, seconds, max_inst):
    import oracle
    from blasfeo_amd.riccati import RiccatiBatch, HostRiccati
    n = min(max_inst, rb.batch)
    sub = RiccatiBatch(rb.nxs[:n], rb.nus[:n], rb.N, rb.ps)
    # same instances: copy the first n instances' input bytes
    buf = sub.inp.host_buffer()
    for i in range(n):
        for a, b in ((sub.i_babt[i], rb.i_babt[i]), (sub.i_rsq[i], rb.i_rsq[i]), (sub.i_qn[i], rb.i_qn[i])):
            sub.inp.view(buf, a)[:] = rb.inp.view(inp, b)
    h = HostRiccati(sub, buf)
    L = oracle.lib()
    threads = L.or_num_threads()
    done = 0
    t0 = time.perf_counter()
    while True:
        L.or_dric_trf_batch(rb.ps, n, h.argv.ctypes.data)
        done += n
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": done / el, "unit": "solves/s", "cores": int(threads), "kind": "port",
            "sample": "%d of the %d instances (same seeded data), repeated %d times, %.1f s; "
                      "CPU oracle (C restatement of blasfeo_ref, -O2, OpenMP over instances)"
                      % (n, rb.batch, done // n, el)}, h


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--horizon", type=int, default=20)
    ap.add_argument("--ps", type=int, default=4, choices=(4, 8))
    ap.add_argument("--seed", type=int, default=0xB1A5FE0 + 3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-max-instances", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed-tol", type=float, default=1e-13,
                    help="ric_mixed: backward-error target of the refinement")
    ap.add_argument("--tune", action="store_true",
                    help="riccati: pick launch shapes and the panel size with blasfeo_amd.autotune on "
                         "this workload before timing")
    ap.add_argument("--eager", action="store_true",
                    help="riccati: time eager calls instead of replays of the batched sweep captured "
                         "in a HIP graph (the default; every replay runs the whole call: plan, "
                         "queue reset and class launches)")
    ap.add_argument("--workload", default="riccati",
                    choices=("riccati", "ric_trs", "ric_sv", "ric_mixed", "pack", "getrs", "trsm_rltn", "gemm_nn", "trmm_rlnn", "syrk_potrf", "potrf_var", "potrf32", "gemm_syrk", "getrf", "sgetrf", "getrf_np",
                             "posv256"),
                    help="riccati (default, the headline); the others measure the remaining "
                         "batched configurations")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="only run a device-to-device copy of this size (rocprof counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the batched instances are independent and the "
              "run uses a single MI355X; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    import blasfeo_amd as bfa
    bfa.lib()

    if args.calibrate_copy:
        n = args.calibrate_copy << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copied %d bytes x3" % n)
        return

    if args.workload == "posv256":
        return run_posv(args, torch, bfa)
    if args.workload == "ric_trs":
        return run_ric_trs(args, torch, bfa)
    if args.workload == "ric_sv":
        return run_ric_sv(args, torch, bfa)
    if args.workload == "ric_mixed":
        return run_ric_mixed(args, torch, bfa)
    if args.workload == "pack":
        return run_pack(args, torch, bfa)
    if args.workload == "getrs":
        return run_getrs(args, torch, bfa)
    if args.workload == "trsm_rltn":
        return run_trsm_rltn(args, torch, bfa)
    if args.workload in ("gemm_nn", "trmm_rlnn", "syrk_potrf", "potrf_var"):
        return run_blas3_extra(args, torch, bfa)
    if args.workload != "riccati":
        return run_other(args, torch, bfa)
    return run_riccati(args, torch, bfa)


def run_riccati(args, torch, bfa):
    from blasfeo_amd.riccati import RiccatiBatch, DeviceRiccati
    rng = np.random.default_rng(args.seed)
    nxs = rng.integers(8, 65, args.batch)
    nus = rng.integers(2, 17, args.batch)
    tuning = None
    if args.tune:
        from blasfeo_amd import autotune
        tuning = autotune.tune_riccati(nxs, nus, args.horizon, log=log)
        args.ps = tuning["ps"]
        log("tuned: %s" % json.dumps(tuning))
    log("building %d Riccati instances (N=%d)" % (args.batch, args.horizon))
    rb = RiccatiBatch(nxs, nus, args.horizon, args.ps)
    inp = rb.fill_inputs(rng)
    dev = DeviceRiccati(rb, inp)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    flops = riccati_flops(nxs, nus, args.horizon)
    algo_bytes = riccati_hbm_bytes(nxs, nus, args.horizon)

    log("warmup %d" % args.warmup)
    with torch.cuda.stream(stream):
        for _ in range(args.warmup):
            dev.run(sh)
    torch.cuda.synchronize()

    step = lambda: dev.run(sh)
    args.graph = not args.eager
    if args.graph:
        graph = dev.capture(stream)
        torch.cuda.synchronize()
        step = graph.replay
    log("timing %d steps%s" % (args.steps, " (graph replays)" if args.graph else ""))
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        ev0.record(stream)
        for _ in range(args.steps):
            step()
        ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    solves = args.batch / (ms * 1e-3)
    achieved = flops / (ms * 1e-3) / 1e12
    log("%.3f ms/step, %.0f solves/s, %.2f TFLOP/s (wall %.3f s)" % (ms, solves, achieved, wall))

    # correctness of the measured results: a sample against the oracle
    checks = {}
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (%.0f s budget)" % args.cpu_seconds)
        cpu, host = cpu_baseline_riccati(rb, inp, args.cpu_seconds, args.cpu_max_instances)
        gbuf = dev.outbuf.cpu().numpy()
        nchk = min(64, host.rb.batch)
        berr, fdiff = riccati_checks(rb, inp, gbuf, host, nchk)
        checks = {"instances": nchk,
                  "max_stage_backward_error_over_n_eps": berr,
                  "max_rel_diff_vs_oracle_L0": fdiff,
                  "pass": berr <= 10.0}
        log("checked %d instances: stage backward error <= %.2f n eps, L_0 rel diff vs oracle %.1e"
            % (nchk, berr, fdiff))

    traffic = read_traffic("k_ric_class", "riccati-%dx%d" % (args.batch, args.horizon))
    rbound, rach, rpeak, runit = roof(flops, algo_bytes, ms, FP64_PEAK_TFLOPS)
    out = {
        "metric": METRIC,
        "value": solves,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic, seeded (numpy PCG64 seed %d): A = I + 0.1 N(0,1), B ~ N(0,1), Q = R = I" % args.seed,
        "config": {"workload": "riccati_trf_batched", "batch": args.batch, "horizon": args.horizon,
                   "nx": "U{8..64}", "nu": "U{2..16}", "ps": args.ps, "precision": "fp64",
                   "parallelism": "one workgroup per instance, 1 GPU",
                   "launch": "hip graph replay" if args.graph else "eager",
                   "tuning": tuning},
        "roofline": {"bound": rbound, "achieved": rach, "peak": rpeak, "unit": runit,
                     "frac": rach / rpeak, "traffic": traffic,
                     "algorithmic_bytes": algo_bytes, "flops_per_launch": flops,
                     "kernel": "k_ric_class<ps,256> + k_ric_class<ps,128> (the class launches of one batched call, run side by side)"},
        "cpu_baseline": cpu,
        "checks": checks,
    }
    print(json.dumps(out), flush=True)



# ---------------------------------------------------------------------------
# the Riccati solve on the factors of the same workload
# ---------------------------------------------------------------------------
def riccati_trs_bytes(nxs, nus, N):
    """Algorithmic HBM bytes of one solve launch, 8 bytes per double.  Each of
    the two sweeps reads the parts of the factors it uses: the trailing
    nx x nx lower block of L[k+1] for every stage, the first nu columns
    (lower part) of L[k] on stages 1..N-1 and all of L[0]; BAbt once per
    instance (time-invariant, the stages re-read it from cache); plus its
    vectors: backward rq/ux of stage N, and per stage b, rq, Pb, ux;
    forward per stage ux[k] in and out, the state part of ux[k+1] in and
    out, b and pi."""
    nx = np.asarray(nxs, dtype=np.float64)
    nu = np.asarray(nus, dtype=np.float64)
    nk = nx + nu
    mats = N * nx * (nx + 1) / 2 + (N - 1) * (nu * nk - nu * (nu - 1) / 2) + nk * (nk + 1) / 2 + nk * nx
    vecs = 2 * nx + N * (2 * nx + 2 * nk) + N * (4 * nx + 2 * nk)
    return float(8 * (2 * mats + vecs).sum())


def _trs_subset(rb, inp, gbuf, Ltab, rs, sin, n):
    """The first n instances in host memory with the GPU's factors (bytes and
    use_dA flags) and the same solve inputs, ready for the oracle's solve."""
    from blasfeo_amd.riccati import RiccatiBatch, RiccatiSolve, HostRiccati
    sub = RiccatiBatch(rb.nxs[:n], rb.nus[:n], rb.N, rb.ps)
    buf = sub.inp.host_buffer()
    for i in range(n):
        for a, b in ((sub.i_babt[i], rb.i_babt[i]), (sub.i_rsq[i], rb.i_rsq[i]), (sub.i_qn[i], rb.i_qn[i])):
            sub.inp.view(buf, a)[:] = rb.inp.view(inp, b)
    h = HostRiccati(sub, buf)
    for i in range(n):
        for k in range(rb.N + 1):
            a, b = sub.i_L[i][k], rb.i_L[i][k]
            sub.out.view(h.outbuf, a)[:] = rb.out.view(gbuf, b)
            sub.out.dview(h.outbuf, a)[:] = rb.out.dview(gbuf, b)
    h.t[2]["use_dA"] = Ltab["use_dA"][: n * (rb.N + 1)]
    srs = RiccatiSolve(sub)
    sbuf = np.zeros(srs.nbytes, dtype=np.uint8)
    for i in range(n):
        for kind, ks in (("b", range(rb.N)), ("rq", range(rb.N + 1))):
            for k in ks:
                j, jj = srs.index(i, kind, k), rs.index(i, kind, k)
                o, oo, m = int(srs.offs[j]), int(rs.offs[jj]), int(srs.sizes[j]) * 8
                sbuf[o:o + m] = sin[oo:oo + m]
    tab = srs.table(sbuf.ctypes.data)
    babt, rsq, L, work, nx, nu = h.t
    argv = srs.args(tab.ctypes.data, babt.ctypes.data, L.ctypes.data, nx.ctypes.data, nu.ctypes.data)
    return sub, srs, sbuf, (h, tab, argv)


def trs_checks(rb, inp, rs, sin, gbuf, Ltab, gsol, n, nchk=64):
    """Checks of a measured solve: the oracle's solve of the first n instances
    on the GPU's factors, then for the first nchk instances the KKT backward
    errors of both solutions and their largest difference (1e-11 rule).
    Returns (checks, the oracle's argument records, objects to keep alive)."""
    import oracle
    from blasfeo_amd.riccati import trs_backward_error
    sub, srs, sbuf, keep = _trs_subset(rb, inp, gbuf, Ltab, rs, sin, n)
    argv = keep[2]
    oracle.lib().or_dric_trs_batch(rb.ps, n, argv.ctypes.data)
    # the recursion amplifies rounding over the horizon (the oracle's own
    # KKT backward error reaches 1e5-1e6 eps on some instances at N = 20),
    # so the full-size bar is accuracy: the GPU's backward error within
    # 10x the oracle's (+100 eps) on the same instance and factors
    nchk = min(nchk, n)
    worst, gb, ob, ok = 0.0, [], [], True
    for i in range(nchk):
        for kind, ks in (("ux", range(rb.N + 1)), ("pi", range(rb.N)), ("Pb", range(rb.N))):
            for k in ks:
                g = rs.vec(gsol, i, kind, k)
                r = srs.vec(sbuf, i, kind, k)
                d = np.abs(g - r)
                rel = np.where(d > 1e-11, d / np.maximum(np.abs(g) + np.abs(r), 1e-300), 0.0)
                worst = max(worst, float(rel.max()) if rel.size else 0.0)
        gb.append(trs_backward_error(rb, inp, rs, sin, gsol, i))
        ob.append(trs_backward_error(sub, keep[0].inbuf, srs, sbuf, sbuf, i))
        ok = ok and gb[-1] <= 10 * ob[-1] + 100
    checks = {"instances": nchk, "max_kkt_backward_error_over_eps": max(gb),
              "median_kkt_backward_error_over_eps": float(np.median(gb)),
              "oracle_max_kkt_backward_error_over_eps": max(ob),
              "oracle_median_kkt_backward_error_over_eps": float(np.median(ob)),
              "max_rel_diff_vs_oracle_on_gpu_factors": worst, "pass": bool(ok)}
    log("checked %d instances: KKT backward error max %.3g / median %.3g eps (oracle %.3g / %.3g); "
        "max rel diff vs oracle %.1e" % (nchk, max(gb), np.median(gb), max(ob), np.median(ob), worst))
    return checks, argv, (sub, srs, sbuf, keep)


def run_ric_trs(args, torch, bfa):
    """The Riccati solve (d_back_ric_trs_libstr) on the factors of the
    headline workload: the factorisation runs once, untimed; each step is one
    batched solve launch."""
    from blasfeo_amd.riccati import RiccatiBatch, DeviceRiccati, RiccatiSolve, DeviceRiccatiSolve
    from blasfeo_amd import DMAT
    import oracle
    rng = np.random.default_rng(args.seed)
    nxs = rng.integers(8, 65, args.batch)
    nus = rng.integers(2, 17, args.batch)
    log("building %d Riccati instances (N=%d) and their solve vectors" % (args.batch, args.horizon))
    rb = RiccatiBatch(nxs, nus, args.horizon, args.ps)
    inp = rb.fill_inputs(rng)
    rs = RiccatiSolve(rb)
    sin = rs.fill_inputs(rng)
    dev = DeviceRiccati(rb, inp)
    ds = DeviceRiccatiSolve(dev, rs, sin)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        dev.run(sh)
    torch.cuda.synchronize()
    log("timing %d solve launches" % args.steps)
    ms = time_steps(torch, stream, lambda: ds.run(sh), args.steps, args.warmup)
    solves = args.batch / (ms * 1e-3)
    algo = riccati_trs_bytes(nxs, nus, args.horizon)
    achieved = algo / (ms * 1e-3) / 1e9
    log("%.3f ms/step, %.0f solves/s, %.0f GB/s algorithmic" % (ms, solves, achieved))

    checks, cpu = {}, None
    if not args.no_cpu:
        gbuf = dev.outbuf.cpu().numpy()
        Ltab = dev.t_dev[2].cpu().numpy().view(DMAT)
        n = min(args.cpu_max_instances, rb.batch)
        checks, argv, keep = trs_checks(rb, inp, rs, sin, gbuf, Ltab, ds.host(), n)
        L = oracle.lib()
        log("cpu baseline (%.0f s budget)" % args.cpu_seconds)
        cpu = cpu_time(lambda: L.or_dric_trs_batch(rb.ps, n, argv.ctypes.data), n, args.cpu_seconds,
                       "solves/s", "%d of the %d instances (same seeded data and factors)" % (n, rb.batch))
    traffic = read_traffic("k_ric_trs", "ric_trs-%dx%d" % (args.batch, args.horizon))
    out = {"metric": METRIC, "value": solves, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic, seeded (numpy PCG64 seed %d): the headline Riccati instances, "
                   "b and rq ~ N(0,1) per stage" % args.seed,
           "config": {"workload": "riccati_trs_batched", "batch": args.batch, "horizon": args.horizon,
                      "nx": "U{8..64}", "nu": "U{2..16}", "ps": args.ps, "precision": "fp64",
                      "parallelism": "one wavefront per instance, 1 GPU"},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                        "algorithmic_bytes": algo, "kernel": "k_ric_trs<%d>" % args.ps},
           "cpu_baseline": cpu, "checks": checks}
    print(json.dumps(out), flush=True)


def run_ric_mixed(args, torch, bfa):
    """Mixed vs FP64 at equal error on the headline Riccati instances with
    b, rq ~ N(0,1): the mixed-precision solve (FP32 factorisation, FP64
    refinement of the KKT solution to backward error <= tol, FP64 fallback
    for the instances that do not get there) against the FP64 factorisation
    + solve, each step one batched call (the FP64 pair: two)."""
    from blasfeo_amd.riccati import (RiccatiBatch, DeviceRiccati, RiccatiSolve, DeviceRiccatiSolve,
                                     DeviceRiccatiMixed, trs_backward_error)
    rng = np.random.default_rng(args.seed)
    nxs = rng.integers(8, 65, args.batch)
    nus = rng.integers(2, 17, args.batch)
    tol, maxit = args.mixed_tol, 3
    log("building %d Riccati instances (N=%d) and their solve vectors" % (args.batch, args.horizon))
    rb = RiccatiBatch(nxs, nus, args.horizon, args.ps)
    inp = rb.fill_inputs(rng)
    rs = RiccatiSolve(rb)
    sin = rs.fill_inputs(rng)
    mix = DeviceRiccatiMixed(rb, inp, rs, sin)
    fac = DeviceRiccati(rb, inp)
    ds = DeviceRiccatiSolve(fac, rs, sin)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        ms = time_steps(torch, stream, lambda: mix.run(tol, maxit, sh), args.steps, args.warmup)

        def fp64():
            fac.run(sh)
            ds.run(sh)
        ms64 = time_steps(torch, stream, fp64, args.steps, args.warmup)
    torch.cuda.synchronize()
    info = mix.infos()
    n_fb = int((info["path"] == 1).sum())
    solves = args.batch / (ms * 1e-3)
    log("mixed %.3f ms/step (%.0f solves/s), FP64 %.3f ms/step; %d of %d instances took the FP64 fallback"
        % (ms, solves, ms64, n_fb, args.batch))
    nchk = min(256, args.batch)
    g64 = ds.host()
    e64 = np.array([trs_backward_error(rb, inp, rs, sin, g64, i) for i in range(nchk)]) * np.finfo(float).eps
    checks = {"tol": tol, "max_corrections": maxit, "fp64_fallback_instances": n_fb,
              "corrections_histogram": np.bincount(info["iters"]).tolist(),
              "mixed_berr_median": float(np.median(info["berr"])), "mixed_berr_max": float(info["berr"].max()),
              "fp64_berr_median_first_%d" % nchk: float(np.median(e64)),
              "pass": bool((info["berr"][info["path"] == 0] <= tol).all())}
    flops = riccati_flops(nxs, nus, args.horizon)
    out = {"metric": METRIC, "value": solves, "unit": "solves/s (mixed precision, KKT berr <= %g or FP64)" % tol,
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 factorisation, f64 refinement",
           "data": "synthetic, seeded (numpy PCG64 seed %d): the headline Riccati instances, "
                   "b and rq ~ N(0,1) per stage" % args.seed,
           "config": {"workload": "riccati_solve_mixed_batched", "batch": args.batch, "horizon": args.horizon,
                      "nx": "U{8..64}", "nu": "U{2..16}", "ps": args.ps,
                      "fp64_trf_trs_ms_per_step": ms64, "mixed_speedup_vs_fp64": ms64 / ms,
                      "parallelism": "one workgroup per instance, 1 GPU"},
           "roofline": {"bound": "mfma", "achieved": flops / (ms * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": flops / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "traffic": None,
                        "flops_per_launch": flops,
                        "kernel": "whole call (factorisation flops only, FP64 peak as the reference roof)"},
           "cpu_baseline": None, "checks": checks}
    print(json.dumps(out), flush=True)


def run_ric_sv(args, torch, bfa):
    """The fused factorisation and solve (d_back_ric_sv_libstr) on the
    headline instances with their linear terms (b, r, q ~ N(0,1)) in the
    augmented rows: each step is one batched call; the same problem through
    the separate factorisation and solve entries is timed beside it."""
    from blasfeo_amd.riccati import (RiccatiBatch, DeviceRiccati, RiccatiSolve, DeviceRiccatiSolve,
                                     DeviceRiccatiSv, HostRiccatiSv, trs_backward_error)
    import oracle
    rng = np.random.default_rng(args.seed)
    nxs = rng.integers(8, 65, args.batch)
    nus = rng.integers(2, 17, args.batch)
    log("building %d augmented Riccati instances (N=%d)" % (args.batch, args.horizon))
    rb = RiccatiBatch(nxs, nus, args.horizon, args.ps, aug=True)
    inp = rb.fill_inputs(rng)
    dev = DeviceRiccatiSv(rb, inp)
    fac = DeviceRiccati(rb, inp)
    rs = RiccatiSolve(rb)
    sin = dev.sv.solve_inputs(inp)
    ds = DeviceRiccatiSolve(fac, rs, sin)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    torch.cuda.synchronize()
    log("timing %d fused calls, then %d factorisation + solve pairs" % (args.steps, args.steps))
    with torch.cuda.stream(stream):
        ms = time_steps(torch, stream, lambda: dev.run(sh), args.steps, args.warmup)

        def pair():
            fac.run(sh)
            ds.run(sh)
        ms_sep = time_steps(torch, stream, pair, args.steps, args.warmup)
    solves = args.batch / (ms * 1e-3)
    flops = riccati_flops(nxs, nus, args.horizon)
    algo_bytes = riccati_hbm_bytes(nxs, nus, args.horizon)
    log("fused %.3f ms/step (%.0f solves/s); separate %.3f ms/step" % (ms, solves, ms_sep))

    checks, cpu = {}, None
    if not args.no_cpu:
        n = min(args.cpu_max_instances, rb.batch)
        sub = RiccatiBatch(nxs[:n], nus[:n], rb.N, rb.ps, aug=True)
        buf = sub.inp.host_buffer()
        for i in range(n):
            for a, b in ((sub.i_babt[i], rb.i_babt[i]), (sub.i_rsq[i], rb.i_rsq[i]), (sub.i_qn[i], rb.i_qn[i])):
                sub.inp.view(buf, a)[:] = rb.inp.view(inp, b)
        host = HostRiccatiSv(sub, buf)
        L = oracle.lib()
        log("cpu baseline (%.0f s budget)" % args.cpu_seconds)
        cpu = cpu_time(lambda: L.or_dric_sv_batch(rb.ps, n, host.argv.ctypes.data), n, args.cpu_seconds,
                       "solves/s", "%d of the %d instances (same seeded data)" % (n, rb.batch))
        v = dev.vectors()
        nchk = min(256, n)
        eg = np.array([trs_backward_error(rb, inp, dev.sv.rs, sin, v, i) for i in range(nchk)])
        eo = np.array([trs_backward_error(sub, buf, host.sv.rs, host.sv.solve_inputs(buf), host.vbuf, i)
                       for i in range(nchk)])
        checks = {"instances": nchk, "kkt_backward_error_eps_median": float(np.median(eg)),
                  "kkt_backward_error_eps_max": float(eg.max()),
                  "oracle_kkt_backward_error_eps_median": float(np.median(eo)),
                  "pass": bool((eg <= 10 * eo + 100).all())}
        log("KKT backward error median %.1f eps (oracle %.1f)" % (np.median(eg), np.median(eo)))
    rbound, rach, rpeak, runit = roof(flops, algo_bytes, ms, FP64_PEAK_TFLOPS)
    out = {"metric": METRIC, "value": solves, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic, seeded (numpy PCG64 seed %d): the headline Riccati instances, "
                   "b, r, q ~ N(0,1) in the augmented rows" % args.seed,
           "config": {"workload": "riccati_sv_batched", "batch": args.batch, "horizon": args.horizon,
                      "nx": "U{8..64}", "nu": "U{2..16}", "ps": args.ps, "precision": "fp64",
                      "parallelism": "one workgroup per instance, 1 GPU", "launch": "eager",
                      "separate_trf_trs_ms_per_step": ms_sep, "fused_speedup": ms_sep / ms},
           "roofline": {"bound": rbound, "achieved": rach, "peak": rpeak, "unit": runit,
                        "frac": rach / rpeak, "traffic": None,
                        "algorithmic_bytes": algo_bytes, "flops_per_launch": flops,
                        "kernel": "k_ric_class<ps,NT,aug> class launches + k_ric_sv_fwd (factorisation flops only)"},
           "cpu_baseline": cpu, "checks": checks}
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------
# column-major <-> panel-major conversion on the device
# ---------------------------------------------------------------------------
def run_pack(args, torch, bfa):
    """blasfeo_pack_dmat for 16384 column-major n x n blocks, n ~ U{16..256}
    (the LU workload's sizes), into panel-major descriptors; each step is one
    batched pack launch.  The batched unpack is timed the same way and
    reported beside it."""
    import oracle
    from blasfeo_amd.device import to_device
    rng = np.random.default_rng(args.seed)
    batch = 16384
    nn = rng.integers(16, 257, batch)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    cm_off = np.concatenate([[0], np.cumsum(nn * nn)[:-1]]).astype(np.int64)
    cm = torch.empty(int((nn * nn).sum()), dtype=torch.float64, device="cuda")
    cm.normal_(generator=torch.Generator(device="cuda").manual_seed(args.seed))
    _, _, nb = bfa.bulk_layout(np.float64, args.ps, nn, nn, 0)
    pm = torch.zeros(nb // 8, dtype=torch.float64, device="cuda")
    d = to_device(bfa.bulk_layout(np.float64, args.ps, nn, nn, pm.data_ptr())[0])
    rec = np.zeros(batch, dtype=bfa.DPACK_ARG)
    rec["m"], rec["n"], rec["lda"] = nn, nn, nn
    rec["A"] = np.uint64(cm.data_ptr()) + (8 * cm_off).astype(np.uint64)
    rec["sA"] = np.uint64(d.data_ptr()) + np.arange(batch, dtype=np.uint64) * bfa.DMAT.itemsize
    drec = to_device(rec)
    out_cm = torch.empty_like(cm)
    rec2 = rec.copy()
    rec2["A"] = np.uint64(out_cm.data_ptr()) + (8 * cm_off).astype(np.uint64)
    drec2 = to_device(rec2)
    torch.cuda.synchronize()
    hint = (0, 256, 0)
    ms = time_steps(torch, stream, lambda: bfa.call("dpack", args.ps, batch, drec.data_ptr(), hint, sh),
                    args.steps, args.warmup)
    ms_un = time_steps(torch, stream, lambda: bfa.call("dunpack", args.ps, batch, drec2.data_ptr(), hint, sh),
                       args.steps, args.warmup)
    # the same blocks moved to the other panel size (blasfeo_pm_dmat descriptors)
    ps2 = 8 if args.ps == 4 else 4
    pmd = np.zeros(2 * batch, dtype=bfa.PM_DMAT)
    hd0 = bfa.bulk_layout(np.float64, args.ps, nn, nn, pm.data_ptr())[0]
    for f_ in ("mem", "pA", "dA", "m", "n", "pm", "cn", "use_dA", "memsize"):
        pmd[f_][:batch] = hd0[f_]
    pmd["ps"][:batch] = args.ps
    _, _, nb2 = bfa.bulk_layout(np.float64, ps2, nn, nn, 0)
    pm2 = torch.zeros(nb2 // 8, dtype=torch.float64, device="cuda")
    hd2 = bfa.bulk_layout(np.float64, ps2, nn, nn, pm2.data_ptr())[0]
    for f_ in ("mem", "pA", "dA", "m", "n", "pm", "cn", "use_dA", "memsize"):
        pmd[f_][batch:] = hd2[f_]
    pmd["ps"][batch:] = ps2
    dpmd = to_device(pmd)
    crec = np.zeros(batch, dtype=bfa.PM_DGECP_ARG)
    crec["m"], crec["n"] = nn, nn
    crec["sA"] = np.uint64(dpmd.data_ptr()) + np.arange(batch, dtype=np.uint64) * bfa.PM_DMAT.itemsize
    crec["sB"] = crec["sA"] + np.uint64(batch * bfa.PM_DMAT.itemsize)
    dcrec = to_device(crec)
    ms_cv = time_steps(torch, stream, lambda: bfa.pm_dgecp(batch, dcrec.data_ptr(), hint, sh),
                       args.steps, args.warmup)
    nf = nn.astype(np.float64)
    algo = float((16 * nf * nf).sum())
    achieved = algo / (ms * 1e-3) / 1e9
    value = batch / (ms * 1e-3)
    log("pack: %.3f ms/step, %.4g matrices/s, %.0f GB/s; unpack %.3f ms, %.0f GB/s; ps %d -> %d %.3f ms, %.0f GB/s"
        % (ms, value, achieved, ms_un, algo / (ms_un * 1e-3) / 1e9, args.ps, ps2, ms_cv,
           algo / (ms_cv * 1e-3) / 1e9))
    checks = {"roundtrip_bit_exact": bool(torch.equal(out_cm, cm))}
    # the converted blocks hold the same values (first 64)
    ok = True
    g1, g2 = pm.cpu().numpy(), pm2.cpu().numpy()
    for i in range(64):
        n_ = int(nn[i])
        a1 = bfa.pm_unpack(g1[(int(hd0["pA"][i]) - pm.data_ptr()) // 8:], args.ps, int(hd0["cn"][i]), n_, n_)
        a2 = bfa.pm_unpack(g2[(int(hd2["pA"][i]) - pm2.data_ptr()) // 8:], ps2, int(hd2["cn"][i]), n_, n_)
        ok = ok and np.array_equal(a1, a2)
    checks["ps_convert_bit_exact_64"] = bool(ok)
    cpu = None
    if not args.no_cpu:
        n = 1024
        hcm = cm[: int((nn[:n] * nn[:n]).sum())].cpu().numpy()
        hd, _, hnb = bfa.bulk_layout(np.float64, args.ps, nn[:n], nn[:n], 0)
        hpm = np.zeros(hnb // 8)
        hd = bfa.bulk_layout(np.float64, args.ps, nn[:n], nn[:n], hpm.ctypes.data)[0]
        hr = rec[:n].copy()
        hr["A"] = np.uint64(hcm.ctypes.data) + (8 * cm_off[:n]).astype(np.uint64)
        hr["sA"] = np.uint64(hd.ctypes.data) + np.arange(n, dtype=np.uint64) * bfa.DMAT.itemsize
        L = oracle.lib()
        L.or_dpack_batch(args.ps, n, hr.ctypes.data)
        gpm = pm.cpu().numpy()
        dd = d.cpu().numpy().view(bfa.DMAT)
        ok = True
        for i in range(min(n, 256)):
            g0 = (int(dd["pA"][i]) - pm.data_ptr()) // 8
            h0 = (int(hd["pA"][i]) - hpm.ctypes.data) // 8
            ln = int(hd["pm"][i]) * int(hd["cn"][i])
            ok = ok and np.array_equal(gpm[g0:g0 + ln], hpm[h0:h0 + ln])
        checks["pack_bit_exact_vs_oracle_256"] = bool(ok)
        cpu = cpu_time(lambda: L.or_dpack_batch(args.ps, n, hr.ctypes.data), n, args.cpu_seconds,
                       "matrices/s", "%d of the %d blocks (same data)" % (n, batch))
    log("checks: %s" % json.dumps(checks))
    out = {"metric": METRIC, "value": value, "unit": "matrices/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic, seeded (%d), N(0,1)" % args.seed,
           "config": {"workload": "dpack_batched", "batch": batch, "n": "U{16..256}", "ps": args.ps,
                      "unpack_ms_per_step": ms_un,
                      "unpack_GBps": algo / (ms_un * 1e-3) / 1e9,
                      "ps_convert": "%d -> %d (pm_dgecp)" % (args.ps, ps2),
                      "ps_convert_ms_per_step": ms_cv,
                      "ps_convert_GBps": algo / (ms_cv * 1e-3) / 1e9},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS,
                        "traffic": read_traffic("k_pack_pipe", "pack-16384"),
                        "algorithmic_bytes": algo, "kernel": "k_pack_pipe<%d>" % args.ps},
           "cpu_baseline": cpu, "checks": checks}
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------
# LU solve on getrf_rp factors
# ---------------------------------------------------------------------------
def run_getrs(args, torch, bfa):
    """The LU solve of examples/example_d_lu_factorization.c:163-171 for the LU
    workload's sizes: 16384 n x n factors, n ~ U{16..256}, 4 right-hand sides.
    getrf_rp runs once, untimed; each step is drowpe + trsm_llnu + trsm_lunn on
    right-hand sides restored from a copy (the copy is timed alone and
    subtracted)."""
    import oracle
    from blasfeo_amd.device import to_device
    rng = np.random.default_rng(args.seed)
    batch, nrhs = 16384, 4
    nn = rng.integers(16, 257, batch)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    gen = torch.Generator(device="cuda").manual_seed(args.seed)
    _, _, nb = bfa.bulk_layout(np.float64, args.ps, nn, nn, 0)
    Abuf = torch.empty(nb // 8, dtype=torch.float64, device="cuda").uniform_(-1, 1, generator=gen)
    dA = to_device(bfa.bulk_layout(np.float64, args.ps, nn, nn, Abuf.data_ptr())[0])
    rr = np.full(batch, nrhs)
    _, _, nbb = bfa.bulk_layout(np.float64, args.ps, nn, rr, 0)
    Bbuf = torch.empty(nbb // 8, dtype=torch.float64, device="cuda").normal_(generator=gen)
    B0 = Bbuf.clone()
    dB = to_device(bfa.bulk_layout(np.float64, args.ps, nn, rr, Bbuf.data_ptr())[0])
    idx = np.arange(batch, dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(nn)[:-1]]).astype(np.uint64)
    piv = torch.zeros(int(nn.sum()), dtype=torch.int32, device="cuda")
    f = np.zeros(batch, dtype=bfa.DFACT_ARG)
    f["m"], f["n"] = nn, nn
    f["sC"] = np.uint64(dA.data_ptr()) + idx * bfa.DMAT.itemsize
    f["sD"] = f["sC"]
    f["ipiv"] = np.uint64(piv.data_ptr()) + 4 * offs
    p = np.zeros(batch, dtype=bfa.DROWPE_ARG)
    p["kmax"], p["ipiv"] = nn, f["ipiv"]
    p["sA"] = np.uint64(dB.data_ptr()) + idx * bfa.DMAT.itemsize
    t = np.zeros(batch, dtype=bfa.DTRSM_ARG)
    t["m"], t["n"], t["alpha"] = nn, nrhs, 1.0
    t["sA"], t["sB"], t["sD"] = f["sC"], p["sA"], p["sA"]
    df, dp, dt = to_device(f), to_device(p), to_device(t)
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        bfa.call("dgetrf_rp", args.ps, batch, df.data_ptr(), (256, 256, 0), sh)
    torch.cuda.synchronize()
    hint = (256, nrhs, 0)

    def step():
        Bbuf.copy_(B0)
        bfa.call("drowpe", args.ps, batch, dp.data_ptr(), None, sh)
        bfa.call("dtrsm_llnu", args.ps, batch, dt.data_ptr(), hint, sh)
        bfa.call("dtrsm_lunn", args.ps, batch, dt.data_ptr(), hint, sh)
    with torch.cuda.stream(stream):
        ms_tot = time_steps(torch, stream, step, args.steps, args.warmup)
        ms_cp = time_steps(torch, stream, lambda: Bbuf.copy_(B0), args.steps, 1)
        step()
    torch.cuda.synchronize()
    ms = ms_tot - ms_cp
    nf = nn.astype(np.float64)
    # the factor read once (strict lower by llnu, upper by lunn) + the pivots,
    # the right-hand sides read and written by each of the three launches
    algo = float((8 * nf * nf + 4 * nf + 3 * 2 * 8 * nf * nrhs).sum())
    flops = float((2 * nf * nf * nrhs).sum())
    achieved = algo / (ms * 1e-3) / 1e9
    value = batch / (ms * 1e-3)
    log("getrs: %.3f ms/step (copy %.3f ms subtracted), %.4g solves/s, %.0f GB/s"
        % (ms, ms_cp, value, achieved))
    # checks: normwise backward error of the first 256 solves in FP64
    nchk = 256
    Ah = Abuf.cpu().numpy()
    dAh = bfa.bulk_layout(np.float64, args.ps, nn, nn, 0)[0]
    dBh = bfa.bulk_layout(np.float64, args.ps, nn, rr, 0)[0]
    # the factored A is gone; rebuild the original from the same generator
    gen2 = torch.Generator(device="cuda").manual_seed(args.seed)
    A0 = torch.empty(nb // 8, dtype=torch.float64, device="cuda").uniform_(-1, 1, generator=gen2).cpu().numpy()
    X = Bbuf.cpu().numpy()
    Bh = B0.cpu().numpy()
    worst = 0.0
    for i in range(nchk):
        n = int(nn[i])
        cnA, cnB = int(dAh["cn"][i]), int(dBh["cn"][i])
        oA, oB = int(dAh["pA"][i]) // 8, int(dBh["pA"][i]) // 8
        A = bfa.pm_unpack(A0[oA:], args.ps, cnA, n, n)
        x = bfa.pm_unpack(X[oB:], args.ps, cnB, n, nrhs)
        bb = bfa.pm_unpack(Bh[oB:], args.ps, cnB, n, nrhs)
        e = np.abs(A @ x - bb).max() / ((np.abs(A).sum(1).max() * np.abs(x).max() + np.abs(bb).max())
                                        * n * np.finfo(float).eps)
        worst = max(worst, float(e))
    checks = {"instances": nchk, "max_backward_error_over_n_eps": worst, "pass": worst <= 20}
    log("checks: %s" % json.dumps(checks))
    cpu = None
    if not args.no_cpu:
        n = 512
        hA = bfa.bulk_layout(np.float64, args.ps, nn[:n], nn[:n], 0)
        hAbuf = np.zeros(hA[2] // 8)
        hAd = bfa.bulk_layout(np.float64, args.ps, nn[:n], nn[:n], hAbuf.ctypes.data)[0]
        hB = bfa.bulk_layout(np.float64, args.ps, nn[:n], rr[:n], 0)
        hBbuf = np.zeros(hB[2] // 8)
        hBd = bfa.bulk_layout(np.float64, args.ps, nn[:n], rr[:n], hBbuf.ctypes.data)[0]
        for i in range(n):
            oA, oB = int(dAh["pA"][i]) // 8, int(dBh["pA"][i]) // 8
            ln, lb = int(hAd["pm"][i]) * int(hAd["cn"][i]), int(hBd["pm"][i]) * int(hBd["cn"][i])
            hAbuf[(int(hAd["pA"][i]) - hAbuf.ctypes.data) // 8:][:ln] = Ah[oA:oA + ln]
            hBbuf[(int(hBd["pA"][i]) - hBbuf.ctypes.data) // 8:][:lb] = Bh[oB:oB + lb]
        hpiv = piv.cpu().numpy()
        fo = oracle.fn("d", "rowpe")
        lo, uo = oracle.fn("d", "trsm_llnu"), oracle.fn("d", "trsm_lunn")
        sa = [hAd.ctypes.data + i * bfa.DMAT.itemsize for i in range(n)]
        sb = [hBd.ctypes.data + i * bfa.DMAT.itemsize for i in range(n)]
        pv = [hpiv[int(offs[i]):int(offs[i]) + int(nn[i])].copy() for i in range(n)]

        hB0 = hBbuf.copy()

        def cpu_run():
            hBbuf[:] = hB0      # right-hand sides restored, as on the GPU
            for i in range(n):
                m = int(nn[i])
                fo(args.ps, m, pv[i].ctypes.data, sb[i])
                lo(args.ps, m, nrhs, 1.0, sa[i], 0, 0, sb[i], 0, 0, sb[i], 0, 0)
                uo(args.ps, m, nrhs, 1.0, sa[i], 0, 0, sb[i], 0, 0, sb[i], 0, 0)
        cpu = cpu_time(cpu_run, n, args.cpu_seconds, "solves/s",
                       "%d of the %d factors (same data)" % (n, batch))
        cpu["cores"] = 1
        cpu["sample"] = cpu["sample"].replace("OpenMP over instances", "one host thread, instance by instance")
    out = {"metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic, seeded (%d): A ~ U[-1,1], B ~ N(0,1)" % args.seed,
           "config": {"workload": "dgetrs_batched (drowpe + dtrsm_llnu + dtrsm_lunn)", "batch": batch,
                      "n": "U{16..256}", "nrhs": nrhs, "ps": args.ps,
                      "timing": "step time minus a same-size device copy that restores B"},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic("getrs", "getrs-16384"),
                        "algorithmic_bytes": algo, "flops_per_launch": flops,
                        "kernel": "k_rowpe + k_trsm_left_thin<double,%d,TL_LLNU> + k_trsm_left_thin<double,%d,TL_LUNN>"
                                  % (args.ps, args.ps)},
           "cpu_baseline": cpu, "checks": checks}
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------
# the other batched configurations
# ---------------------------------------------------------------------------
def time_steps(torch, stream, fn, steps, warmup):
    """Average milliseconds per step from HIP events on the launch stream."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        fn()
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def _records(dtype, n):
    return np.zeros(n, dtype=dtype)


def run_trsm_rltn(args, torch, bfa):
    """16384 x dtrsm_rltn (D = B A^{-T}, A lower n x n, B m x n; m, n ~ U{16..256}, the
    distribution of the LU workload), the reference's blasfeo_dtrsm_rltn per instance.  A is
    U[-1,1] with +-n added to its diagonal, B ~ N(0,1), alpha = 1, D separate.  The first call
    forms A's reciprocal diagonal and sets use_dA (the reference's side effect); the timed
    calls use that cache."""
    from blasfeo_amd.device import to_device
    import oracle
    gen = torch.Generator(device="cuda")
    gen.manual_seed(args.seed)
    rng = np.random.default_rng(args.seed)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    ps = args.ps
    batch = 16384
    nn = rng.integers(16, 257, batch)
    mm = rng.integers(16, 257, batch)
    bufs, descs, hds = {}, {}, {}
    for name, (r, c) in {"A": (nn, nn), "B": (mm, nn), "D": (mm, nn)}.items():
        d, off, nb = bfa.bulk_layout(np.float64, ps, r, c, 0)
        buf = torch.empty(max(nb, 8), dtype=torch.uint8, device="cuda")
        _fill(torch, buf, "uniform" if name == "A" else "normal", gen)
        hd = bfa.bulk_layout(np.float64, ps, r, c, buf.data_ptr())[0]
        bufs[name], descs[name], hds[name] = buf, to_device(hd), hd
    # diagonal dominance: +-n on A's diagonal
    base = bufs["A"].data_ptr()
    di = []
    for i in range(batch):
        t = np.arange(int(nn[i]))
        di.append((int(hds["A"]["pA"][i]) - base) // 8 + (t - t % ps) * int(hds["A"]["cn"][i]) + t * ps + t % ps)
    di = torch.from_numpy(np.concatenate(di)).to("cuda")
    va = bufs["A"].view(torch.float64)
    va[di] += torch.from_numpy(np.concatenate([np.where(np.arange(k) % 2, k, -k) for k in nn]).astype(np.float64)).to("cuda")
    rec = _records(bfa.DTRSM_ARG, batch)
    rec["m"], rec["n"], rec["alpha"] = mm, nn, 1.0
    idx = np.arange(batch, dtype=np.uint64) * bfa.DMAT.itemsize
    rec["sA"] = descs["A"].data_ptr() + idx
    rec["sB"] = descs["B"].data_ptr() + idx
    rec["sD"] = descs["D"].data_ptr() + idx
    drec = to_device(rec)
    fn = lambda: bfa.call("dtrsm_rltn", ps, batch, drec.data_ptr(), (256, 256, 0), sh)
    ms = time_steps(torch, stream, fn, args.steps, args.warmup)
    mf, nf = mm.astype(np.float64), nn.astype(np.float64)
    flops = float((mf * nf * nf).sum())
    algo = float((8 * (nf * (nf + 1) / 2 + 2 * mf * nf + nf)).sum())
    value = batch / (ms * 1e-3)
    bound, achieved, peak, punit = roof(flops, algo, ms, FP64_PEAK_TFLOPS)
    # the first 64 instances against the oracle's trsm_rltn on the same blocks
    torch.cuda.synchronize()
    host = {k: bufs[k].cpu().numpy().view(np.float64) for k in bufs}
    worst = 0.0
    for i in range(64):
        n, m = int(nn[i]), int(mm[i])
        lay = bfa.Layout(np.float64, ps)
        lay.add(n, n)
        lay.add(m, n)
        lay.add(m, n)
        hb = lay.host_buffer()
        for slot, k in ((0, "A"), (1, "B")):
            o = (int(hds[k]["pA"][i]) - bufs[k].data_ptr()) // 8
            ln = int(hds[k]["pm"][i]) * int(hds[k]["cn"][i])
            lay.view(hb, slot)[:ln] = host[k][o:o + ln]
        desc = lay.build(hb.ctypes.data)
        oracle.fn("d", "trsm_rltn")(ps, m, n, 1.0, desc.ctypes.data, 0, 0,
                                    desc.ctypes.data + bfa.DMAT.itemsize, 0, 0,
                                    desc.ctypes.data + 2 * bfa.DMAT.itemsize, 0, 0)
        r = lay.get(hb, 2)
        o = (int(hds["D"]["pA"][i]) - bufs["D"].data_ptr()) // 8
        g = bfa.pm_unpack(host["D"][o:o + int(hds["D"]["pm"][i]) * int(hds["D"]["cn"][i])], ps,
                          int(hds["D"]["cn"][i]), m, n)
        dd = np.abs(g - r)
        rel = np.where(dd > 1e-11, dd / np.maximum(np.abs(g) + np.abs(r), 1e-300), 0.0)
        worst = max(worst, float(rel.max()))
    checks = {"instances": 64, "max_rel_diff_vs_oracle": worst, "pass": worst <= 1e-11}
    log("checks: %s" % json.dumps(checks))
    log("trsm_rltn: %.3f ms/step, %.4g solves/s, %.2f %s" % (ms, value, achieved, punit))
    out = {"metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic, seeded (%d): A = U[-1,1] + diag(+-n), B ~ N(0,1)" % args.seed,
           "config": {"workload": "dtrsm_rltn_batched", "batch": batch, "m,n": "U{16..256}", "ps": ps,
                      "precision": "fp64"},
           "roofline": {"bound": bound, "achieved": achieved, "peak": peak, "unit": punit,
                        "frac": achieved / peak, "traffic": read_traffic("k_trsm_rltn", "trsm-16384"),
                        "algorithmic_bytes": algo, "flops_per_launch": flops,
                        "kernel": "k_trsm_rltn<double,%d> (one persistent launch per n bucket)" % ps},
           "cpu_baseline": None, "checks": checks}
    print(json.dumps(out), flush=True)


def _diag_index(hd, base, nn, ps):
    """Element indexes (in doubles from `base`) of the diagonal of every block of a bulk
    layout: row t of a panel-major block sits at (t - t%ps)*cn + t*ps + t%ps."""
    di = []
    for i in range(len(nn)):
        t = np.arange(int(nn[i]))
        di.append((int(hd["pA"][i]) - base) // 8 + (t - t % ps) * int(hd["cn"][i]) + t * ps + t % ps)
    return np.concatenate(di)


def _block(host, hd, base, i):
    """The flat panel-major block of instance i from a host copy of a bulk buffer."""
    o = (int(hd["pA"][i]) - base) // 8
    return host[o:o + int(hd["pm"][i]) * int(hd["cn"][i])]


def run_blas3_extra(args, torch, bfa):
    """The batched BLAS-3 calls without a workload of their own elsewhere: dgemm_nn (the
    reference's blasfeo_dgemm_nn), dtrmm_rlnn (blasfeo_dtrmm_rlnn) and dsyrk_dpotrf_ln
    (blasfeo_dsyrk_dpotrf_ln), each as one batched call per step on separate inputs and
    outputs, so that every step does the same work.  The first 64 instances of the last call
    are checked against the oracle on the same blocks (relative difference of the entries
    that differ by more than 1e-11)."""
    from blasfeo_amd.device import to_device
    import oracle
    gen = torch.Generator(device="cuda")
    gen.manual_seed(args.seed)
    rng = np.random.default_rng(args.seed)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    ps = args.ps
    wl = args.workload
    if wl == "gemm_nn":
        # 100000 x D = C + A B, A m x k, B k x n, m, n, k ~ U{4..128}, N(0,1)
        batch = 100000
        m = rng.integers(4, 129, batch); n = rng.integers(4, 129, batch); k = rng.integers(4, 129, batch)
        shapes = {"A": (m, k), "B": (k, n), "C": (m, n), "D": (m, n)}
        fills = {"A": "normal", "B": "normal", "C": "normal", "D": None}
    elif wl == "trmm_rlnn":
        # 16384 x D = B A, A lower n x n U[-1,1], B m x n N(0,1), m, n ~ U{16..256}
        batch = 16384
        n = rng.integers(16, 257, batch); m = rng.integers(16, 257, batch)
        shapes = {"A": (n, n), "B": (m, n), "D": (m, n)}
        fills = {"A": "uniform", "B": "normal", "D": None}
    elif wl == "potrf_var":
        # 65536 x D = chol(C), C = 0.01 N(0,1) + m I (lower triangle read), m ~ U{8..128}
        batch = 65536
        m = rng.integers(8, 129, batch)
        k = np.zeros(batch, dtype=np.int64)
        shapes = {"C": (m, m), "D": (m, m)}
        fills = {"C": "normal", "D": None}
    else:
        # 65536 x D = chol(C + A A^T), A m x k N(0,1), C = (m + k) I, m, k ~ U{8..128}
        batch = 65536
        m = rng.integers(8, 129, batch); k = rng.integers(8, 129, batch)
        shapes = {"A": (m, k), "C": (m, m), "D": (m, m)}
        fills = {"A": "normal", "C": "zero", "D": None}
    bufs, descs, hds = {}, {}, {}
    for name, (r, c) in shapes.items():
        nb = bfa.bulk_layout(np.float64, ps, r, c, 0)[2]
        buf = torch.empty(max(nb, 8), dtype=torch.uint8, device="cuda")
        if fills[name] == "zero" or fills[name] is None:
            buf.zero_()
        else:
            _fill(torch, buf, fills[name], gen)
        hd = bfa.bulk_layout(np.float64, ps, r, c, buf.data_ptr())[0]
        bufs[name], descs[name], hds[name] = buf, to_device(hd), hd
    idx = np.arange(batch, dtype=np.uint64) * bfa.DMAT.itemsize
    sp = {name: descs[name].data_ptr() + idx for name in shapes}
    mf = m.astype(np.float64)
    if wl == "gemm_nn":
        rec = _records(bfa.DGEMM_ARG, batch)
        rec["m"], rec["n"], rec["k"] = m, n, k
        rec["alpha"], rec["beta"] = 1.0, 1.0
        rec["sA"], rec["sB"], rec["sC"], rec["sD"] = sp["A"], sp["B"], sp["C"], sp["D"]
        call, hint = "dgemm_nn", (128, 128, 128)
        nf, kf = n.astype(np.float64), k.astype(np.float64)
        flops = float((2 * mf * nf * kf).sum())
        algo = float((8 * (mf * kf + kf * nf + 2 * mf * nf)).sum())
        unit = "ops/s (dgemm_nn calls)"
        kernel = "k_gemm_wave<double,%d,OP_NN>" % ps
        traffic = read_traffic("k_gemm_wave", "gemm-nn-100000")
        cfg = {"workload": "dgemm_nn_batched", "batch": batch, "m,n,k": "U{4..128}"}
    elif wl == "trmm_rlnn":
        rec = _records(bfa.DTRSM_ARG, batch)
        rec["m"], rec["n"], rec["alpha"] = m, n, 1.0
        rec["sA"], rec["sB"], rec["sD"] = sp["A"], sp["B"], sp["D"]
        call, hint = "dtrmm_rlnn", (256, 256, 0)
        nf = n.astype(np.float64)
        flops = float((mf * nf * (nf + 1)).sum())
        algo = float((8 * (nf * (nf + 1) / 2 + 2 * mf * nf)).sum())
        unit = "ops/s (dtrmm_rlnn calls)"
        kernel = "k_trmm_rlnn<double,%d>" % ps
        traffic = read_traffic("k_trmm_rlnn", "trmm-16384")
        cfg = {"workload": "dtrmm_rlnn_batched", "batch": batch, "m,n": "U{16..256}"}
    elif wl == "potrf_var":
        base = bufs["C"].data_ptr()
        bufs["C"].view(torch.float64).mul_(0.01)
        di = torch.from_numpy(_diag_index(hds["C"], base, m, ps)).to("cuda")
        bufs["C"].view(torch.float64)[di] += torch.from_numpy(np.repeat(m.astype(np.float64), m)).to("cuda")
        rec = _records(bfa.DFACT_ARG, batch)
        rec["m"], rec["n"] = m, m
        rec["sC"], rec["sD"] = sp["C"], sp["D"]
        call, hint = "dpotrf_l", (128, 128, 0)
        flops = float((mf ** 3 / 3).sum())
        algo = float((8 * (mf * (mf + 1) + mf)).sum())
        unit = "factorizations/s (dpotrf_l calls)"
        kernel = ("k_potrf_lds_q<double,%d,64|256,false> (one persistent launch per 16-row bucket, "
                  "packed lower tiles in LDS sized for the bucket)" % ps)
        traffic = read_traffic("k_potrf_lds", "potrf-var-65536")
        cfg = {"workload": "dpotrf_l_batched", "batch": batch, "n": "U{8..128}"}
    else:
        base = bufs["C"].data_ptr()
        di = torch.from_numpy(_diag_index(hds["C"], base, m, ps)).to("cuda")
        bufs["C"].view(torch.float64)[di] = torch.from_numpy(
            np.repeat((m + k).astype(np.float64), m)).to("cuda")
        rec = _records(bfa.DSYRK_DPOTRF_ARG, batch)
        rec["m"], rec["n"], rec["k"] = m, m, k
        rec["sA"], rec["sB"], rec["sC"], rec["sD"] = sp["A"], sp["A"], sp["C"], sp["D"]
        call, hint = "dsyrk_dpotrf_ln", (128, 128, 128)
        kf = k.astype(np.float64)
        flops = float((mf * (mf + 1) * kf + mf ** 3 / 3).sum())
        algo = float((8 * (mf * kf + mf * (mf + 1) + mf)).sum())
        unit = "factorizations/s (dsyrk_dpotrf_ln calls)"
        kernel = ("k_syrk_potrf_lds_q<double,%d,64|256> (one persistent launch per 16-row "
                  "bucket, packed lower tiles in LDS sized for the bucket)" % ps)
        traffic = read_traffic("k_syrk_potrf_lds", "syrk-potrf-65536")
        cfg = {"workload": "dsyrk_dpotrf_ln_batched", "batch": batch, "m,k": "U{8..128}"}
    drec = to_device(rec)
    fn = lambda: bfa.call(call, ps, batch, drec.data_ptr(), hint, sh)
    ms = time_steps(torch, stream, fn, args.steps, args.warmup)
    value = batch / (ms * 1e-3)
    bound, achieved, peak, punit = roof(flops, algo, ms, FP64_PEAK_TFLOPS)
    # the first 64 instances against the oracle on copies of the same blocks
    torch.cuda.synchronize()
    host = {name: bufs[name][:int(hds[name]["pA"][64]) - bufs[name].data_ptr()].cpu().numpy().view(np.float64)
            for name in bufs}                     # the blocks of the first 64 instances
    names = list(shapes)
    worst = 0.0
    for i in range(64):
        lay = bfa.Layout(np.float64, ps)
        for name in names:
            lay.add(int(shapes[name][0][i]), int(shapes[name][1][i]))
        hb = lay.host_buffer()
        for s, name in enumerate(names):
            if name != "D":
                blk = _block(host[name], hds[name], bufs[name].data_ptr(), i)
                lay.view(hb, s)[:len(blk)] = blk
        desc = lay.build(hb.ctypes.data)
        a = {name: desc.ctypes.data + s * bfa.DMAT.itemsize for s, name in enumerate(names)}
        mi = int(m[i])
        if wl == "gemm_nn":
            oracle.fn("d", "gemm_nn")(ps, mi, int(n[i]), int(k[i]), 1.0, a["A"], 0, 0, a["B"], 0, 0,
                                      1.0, a["C"], 0, 0, a["D"], 0, 0)
        elif wl == "trmm_rlnn":
            oracle.fn("d", "trmm_rlnn")(ps, mi, int(n[i]), 1.0, a["A"], 0, 0, a["B"], 0, 0,
                                        a["D"], 0, 0)
        elif wl == "potrf_var":
            oracle.fn("d", "potrf_l")(ps, mi, a["C"], 0, 0, a["D"], 0, 0)
        else:
            oracle.fn("d", "syrk_potrf_ln")(ps, mi, int(k[i]), a["A"], 0, 0, a["A"], 0, 0,
                                            a["C"], 0, 0, a["D"], 0, 0)
        r = lay.get(hb, names.index("D"))
        cn = int(hds["D"]["cn"][i])
        g = bfa.pm_unpack(_block(host["D"], hds["D"], bufs["D"].data_ptr(), i), ps, cn, r.shape[0], r.shape[1])
        if wl in ("syrk_potrf", "potrf_var"):
            g, r = np.tril(g), np.tril(r)
        dd = np.abs(g - r)
        rel = np.where(dd > 1e-11, dd / np.maximum(np.abs(g) + np.abs(r), 1e-300), 0.0)
        worst = max(worst, float(rel.max()))
    checks = {"instances": 64, "max_rel_diff_vs_oracle": worst, "pass": worst <= 1e-11}
    log("checks: %s" % json.dumps(checks))
    log("%s: %.3f ms/step, %.4g %s, %.2f %s" % (wl, ms, value, unit, achieved, punit))
    cfg.update({"ps": ps, "precision": "fp64"})
    out = {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic, seeded (%d)" % args.seed,
           "config": cfg,
           "roofline": {"bound": bound, "achieved": achieved, "peak": peak, "unit": punit,
                        "frac": achieved / peak, "traffic": traffic, "algorithmic_bytes": algo,
                        "flops_per_launch": flops, "kernel": kernel},
           "cpu_baseline": None, "checks": checks}
    print(json.dumps(out), flush=True)


def _fill(torch, buf, kind, gen):
    """Random fill of a whole device byte buffer viewed as doubles."""
    v = buf.view(torch.float64)
    chunk = 1 << 27
    for s0 in range(0, v.numel(), chunk):
        part = v[s0:s0 + chunk]
        if kind == "normal":
            part.normal_(generator=gen)
        else:
            part.uniform_(-1.0, 1.0, generator=gen)


def run_other(args, torch, bfa):
    from blasfeo_amd.device import to_device
    import oracle
    gen = torch.Generator(device="cuda")
    gen.manual_seed(args.seed)
    rng = np.random.default_rng(args.seed)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    ps = args.ps
    cpu = None
    checks = {}

    def dbuf(nbytes):
        return torch.empty(max(nbytes, 8), dtype=torch.uint8, device="cuda")

    if args.workload == "potrf32":
        # 4096 x A = B B^T + 32 I, B ~ U[-1,1], n = 32 (the reference's CPU-runnable case)
        batch, n = 4096, 32
        dC, offC, nbC = bfa.bulk_layout(np.float64, ps, [n] * batch, [n] * batch, 0)
        bufC = dbuf(nbC)
        bufD = dbuf(nbC)
        B = torch.rand((batch, n, n), dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
        A = torch.bmm(B, B.transpose(1, 2).contiguous()) + n * torch.eye(n, dtype=torch.float64, device="cuda")
        stride = int(offC[1]) // 8
        pm = n
        packed = A.view(batch, pm // ps, ps, n).permute(0, 1, 3, 2).reshape(batch, pm * n)
        bufC.view(torch.float64).view(batch, stride)[:, : pm * n] = packed
        descC = to_device(bfa.bulk_layout(np.float64, ps, [n] * batch, [n] * batch, bufC.data_ptr())[0])
        descD = to_device(bfa.bulk_layout(np.float64, ps, [n] * batch, [n] * batch, bufD.data_ptr())[0])
        rec = _records(bfa.DFACT_ARG, batch)
        rec["m"] = n
        rec["n"] = n
        rec["sC"] = descC.data_ptr() + np.arange(batch, dtype=np.uint64) * bfa.DMAT.itemsize
        rec["sD"] = descD.data_ptr() + np.arange(batch, dtype=np.uint64) * bfa.DMAT.itemsize
        drec = to_device(rec)
        fn = lambda: bfa.call("dpotrf_l", ps, batch, drec.data_ptr(), (n, n, 0), sh)
        ms = time_steps(torch, stream, fn, args.steps, args.warmup)
        flops = batch * n ** 3 / 3.0
        algo = batch * 8704.0
        value = batch / (ms * 1e-3)
        unit = "factorizations/s"
        bound, achieved, peak, punit = roof(flops, algo, ms, FP64_PEAK_TFLOPS)
        traffic = read_traffic("k_potrf_lds", "potrf-4096x32")
        kernel = "k_potrf_lds<double,%d,64>" % ps
        wl = {"workload": "dpotrf_l_batched", "batch": batch, "n": n, "ps": ps, "precision": "fp64"}
        # check: backward error of every factor
        L = bufD.view(torch.float64).view(batch, stride)[:, : pm * n].view(batch, pm // ps, n, ps)
        L = L.permute(0, 1, 3, 2).reshape(batch, n, n).tril()
        berr = ((A - torch.bmm(L, L.transpose(1, 2).contiguous())).flatten(1).norm(dim=1)
                / A.flatten(1).norm(dim=1)).max().item()
        checks = {"max_backward_error": berr, "bound_10_n_eps": 10 * n * 2.22e-16,
                  "pass": berr <= 10 * n * 2.22e-16}
        if not args.no_cpu:
            hb = bufC.cpu().numpy()
            hd = np.zeros_like(hb)
            hC = bfa.bulk_layout(np.float64, ps, [n] * batch, [n] * batch, hb.ctypes.data)[0]
            hD = bfa.bulk_layout(np.float64, ps, [n] * batch, [n] * batch, hd.ctypes.data)[0]
            hr = _records(bfa.DFACT_ARG, batch)
            hr["m"] = n
            hr["n"] = n
            hr["sC"] = hC.ctypes.data + np.arange(batch, dtype=np.uint64) * bfa.DMAT.itemsize
            hr["sD"] = hD.ctypes.data + np.arange(batch, dtype=np.uint64) * bfa.DMAT.itemsize
            cpu = cpu_time(lambda: oracle.lib().or_dpotrf_l_batch(ps, batch, hr.ctypes.data), batch,
                           args.cpu_seconds, "factorizations/s",
                           "all %d instances per repetition" % batch)
    elif args.workload == "gemm_syrk":
        # 100000 dgemm_nt + 100000 dsyrk_ln, m, n, k ~ U{4..128}, N(0,1), alpha = beta = 1, D = C
        batch = 100000
        m = rng.integers(4, 129, batch); n = rng.integers(4, 129, batch); k = rng.integers(4, 129, batch)
        ms_ = rng.integers(4, 129, batch); ks_ = rng.integers(4, 129, batch)
        lay = {}
        for name, (r, c) in {"gA": (m, k), "gB": (n, k), "gC": (m, n), "sA": (ms_, ks_), "sC": (ms_, ms_)}.items():
            d, off, nb = bfa.bulk_layout(np.float64, ps, r, c, 0)
            buf = dbuf(nb)
            _fill(torch, buf, "normal", gen)
            d = bfa.bulk_layout(np.float64, ps, r, c, buf.data_ptr())[0]
            lay[name] = (buf, to_device(d))
        idx = np.arange(batch, dtype=np.uint64) * bfa.DMAT.itemsize
        g = _records(bfa.DGEMM_ARG, batch)
        g["m"], g["n"], g["k"] = m, n, k
        g["alpha"], g["beta"] = 1.0, 1.0
        g["sA"] = lay["gA"][1].data_ptr() + idx
        g["sB"] = lay["gB"][1].data_ptr() + idx
        g["sC"] = lay["gC"][1].data_ptr() + idx
        g["sD"] = g["sC"]
        sy = _records(bfa.DGEMM_ARG, batch)
        sy["m"], sy["n"], sy["k"] = ms_, ms_, ks_
        sy["alpha"], sy["beta"] = 1.0, 1.0
        sy["sA"] = lay["sA"][1].data_ptr() + idx
        sy["sB"] = sy["sA"]
        sy["sC"] = lay["sC"][1].data_ptr() + idx
        sy["sD"] = sy["sC"]
        dg, ds = to_device(g), to_device(sy)

        def fn():
            bfa.call("dgemm_nt", ps, batch, dg.data_ptr(), (128, 128, 128), sh)
            bfa.call("dsyrk_ln", ps, batch, ds.data_ptr(), (128, 128, 128), sh)
        ms = time_steps(torch, stream, fn, args.steps, args.warmup)
        mf, nf, kf, sm, sk = (x.astype(np.float64) for x in (m, n, k, ms_, ks_))
        flops = float((2 * mf * nf * kf).sum() + (sm * (sm + 1) * sk).sum())
        algo = float(8 * ((mf * kf + nf * kf + 2 * mf * nf).sum() + (sm * sk + sm * (sm + 1)).sum()))
        value = 2 * batch / (ms * 1e-3)
        unit = "ops/s (dgemm_nt + dsyrk_ln calls)"
        bound, achieved, peak, punit = roof(flops, algo, ms, FP64_PEAK_TFLOPS)
        traffic = read_traffic("k_gemm_wave", "gemm-syrk-100000")
        kernel = "k_gemm_wave<double,%d,OP_NT> + k_gemm_wave<double,%d,OP_SYRK>" % (ps, ps)
        wl = {"workload": "dgemm_nt+dsyrk_ln_batched", "batch": "100000 + 100000",
              "m,n,k": "U{4..128}", "ps": ps, "precision": "fp64"}
    else:
        # 16384 x getrf_rp, n ~ U{16..256}, U[-1,1]
        prec = "s" if args.workload == "sgetrf" else "d"
        dt = np.float32 if prec == "s" else np.float64
        batch = 16384
        nn = rng.integers(16, 257, batch)
        d, off, nb = bfa.bulk_layout(dt, ps, nn, nn, 0)
        buf = dbuf(nb)
        if prec == "s":
            v = buf.view(torch.float32)
            v.uniform_(-1.0, 1.0, generator=gen)
        else:
            _fill(torch, buf, "uniform", gen)
        hd = bfa.bulk_layout(dt, ps, nn, nn, buf.data_ptr())[0]
        d = to_device(hd)
        if args.workload == "getrf_np":
            # no pivoting: make every block diagonally dominant (+-n on the diagonal)
            v = buf.view(torch.float64)
            base = buf.data_ptr()
            di = []
            for i in range(batch):
                t = np.arange(int(nn[i]))
                di.append((int(hd["pA"][i]) - base) // 8 + (t - t % ps) * int(hd["cn"][i]) + t * ps + t % ps)
            di = torch.from_numpy(np.concatenate(di)).to("cuda")
            v[di] += torch.from_numpy(np.concatenate([np.where(np.arange(k) % 2, k, -k) for k in nn]).astype(np.float64)).to("cuda")
        piv_off = np.concatenate([[0], np.cumsum(nn)[:-1]]).astype(np.uint64)
        piv = torch.zeros(int(nn.sum()), dtype=torch.int32, device="cuda")
        rec = _records(bfa.DFACT_ARG, batch)
        rec["m"], rec["n"] = nn, nn
        idx = np.arange(batch, dtype=np.uint64) * bfa.DMAT.itemsize
        rec["sC"] = d.data_ptr() + idx
        rec["sD"] = rec["sC"]
        rec["ipiv"] = piv.data_ptr() + 4 * piv_off
        drec = to_device(rec)
        orig = buf.clone()
        name = "dgetrf_np" if args.workload == "getrf_np" else prec + "getrf_rp"

        def fn():
            buf.copy_(orig)
            bfa.call(name, ps, batch, drec.data_ptr(), (256, 256, 0), sh)
        # the in-place factorisation needs its input restored every step: time the
        # restore copy alone and subtract it
        ms_total = time_steps(torch, stream, fn, args.steps, args.warmup)
        ms_copy = time_steps(torch, stream, lambda: buf.copy_(orig), args.steps, 1)
        ms = ms_total - ms_copy
        nf = nn.astype(np.float64)
        flops = float((2.0 / 3.0 * nf ** 3).sum())
        algo = float((np.dtype(dt).itemsize * 2 * nf ** 2 + 12 * nf).sum())
        value = batch / (ms * 1e-3)
        unit = "factorizations/s"
        bound, achieved, peak, punit = roof(flops, algo, ms, FP64_PEAK_TFLOPS if prec == "d" else 157.3)
        traffic = read_traffic("k_getrf", "getrf-16384") if args.workload == "getrf" else None
        kernel = "k_getrf<%s,%d%s>" % ("double" if prec == "d" else "float", ps,
                                       ",PIV=false" if args.workload == "getrf_np" else "")
        if args.workload == "getrf_np":
            # the first 64 factors against the oracle's getrf_np on the same blocks
            # (the copy timing above left the input in buf: factor it once more)
            with torch.cuda.stream(stream):
                fn()
            torch.cuda.synchronize()
            gout, gin = buf.cpu().numpy().view(np.float64), orig.cpu().numpy().view(np.float64)
            worst = 0.0
            for i in range(64):
                n = int(nn[i])
                o = (int(hd["pA"][i]) - buf.data_ptr()) // 8
                ln = int(hd["pm"][i]) * int(hd["cn"][i])
                lay = bfa.Layout(np.float64, ps)
                lay.add(n, n)
                lay.add(n, n)
                hb = lay.host_buffer()
                lay.view(hb, 0)[:] = gin[o:o + ln]
                desc = lay.build(hb.ctypes.data)
                oracle.fn("d", "getrf_np")(ps, n, n, desc.ctypes.data, 0, 0,
                                           desc.ctypes.data + bfa.DMAT.itemsize, 0, 0)
                r = lay.get(hb, 1)                                  # the n x n block only
                g = bfa.pm_unpack(gout[o:o + ln], ps, int(hd["cn"][i]), n, n)
                dd = np.abs(g - r)
                rel = np.where(dd > 1e-11, dd / np.maximum(np.abs(g) + np.abs(r), 1e-300), 0.0)
                worst = max(worst, float(rel.max()))
            checks = {"instances": 64, "max_rel_diff_vs_oracle": worst, "pass": worst <= 1e-11}
            log("checks: %s" % json.dumps(checks))
        wl = {"workload": "%s_batched" % name, "batch": batch, "n": "U{16..256}", "ps": ps,
              "precision": "fp64" if prec == "d" else "fp32",
              "timing": "step time minus a same-size device copy that restores the input"}
    log("%s: %.3f ms/step, %.4g %s, %.2f %s" % (args.workload, ms, value, unit, achieved, punit))
    out = {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32" if args.workload == "sgetrf" else "f64",
           "data": "synthetic, seeded (%d)" % args.seed, "config": wl,
           "roofline": {"bound": bound, "achieved": achieved, "peak": peak, "unit": punit,
                        "frac": achieved / peak, "traffic": traffic, "algorithmic_bytes": algo,
                        "flops_per_launch": flops, "kernel": kernel},
           "cpu_baseline": cpu, "checks": checks}
    print(json.dumps(out), flush=True)


def _cn(ps, n):
    pld = 8 if ps == 8 else 4
    return (n + pld - 1) // pld * pld


def _pm_pack(torch, A, ps):
    """(batch, m, n) dense -> (batch, m*cn) panel-major FP64 payload (m a multiple of ps)."""
    b, m, n = A.shape
    cn = _cn(ps, n)
    if cn != n:
        A = torch.nn.functional.pad(A, (0, cn - n))
    return A.reshape(b, m // ps, ps, cn).permute(0, 1, 3, 2).reshape(b, m * cn)


def _pm_unpack(torch, flat, ps, m, n):
    b = flat.shape[0]
    cn = _cn(ps, n)
    return flat[:, : m * cn].reshape(b, m // ps, cn, ps).permute(0, 1, 3, 2).reshape(b, m, cn)[:, :, :n]


def _posv_berr(torch, A, X, B):
    """Normwise backward error per instance: max_c |b_c - A x_c|_inf / (|A|_inf |x_c|_inf + |b_c|_inf)."""
    R = B - torch.bmm(A, X.contiguous())
    an = A.abs().sum(2).amax(1, keepdim=True)
    den = an * X.abs().amax(1) + B.abs().amax(1)
    return (R.abs().amax(1) / den).amax(1)


def run_posv(args, torch, bfa):
    """2048 SPD solves, n = 256, 4 right-hand sides: the mixed-precision entry
    (FP32 factor in LDS + FP64 refinement) against the plain FP64 sequence
    dpotrf_l + dtrsm_llnn + dtrsm_lltn on the same inputs, both to the
    backward error each reaches (the mixed one refines to tol = 1e-14)."""
    from blasfeo_amd.device import to_device
    import oracle
    ps = args.ps
    batch, n, nrhs = 2048, 256, 4
    tol, max_iter = 1e-14, 30
    gen = torch.Generator(device="cuda")
    gen.manual_seed(args.seed)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    dev = "cuda"
    f64 = torch.float64
    # A = Q diag(lambda) Q^T, Q Haar (QR of N(0,1) with the sign fix), lambda
    # geometric from 1 to 1/cond, cond = 10^U[2,8]; B ~ N(0,1)
    A = torch.empty((batch, n, n), dtype=f64, device=dev)
    for b0 in range(0, batch, 256):
        G = torch.randn((256, n, n), dtype=f64, device=dev, generator=gen)
        # the QR on the host (no GPU library kernels in a profiled run)
        Q, R = torch.linalg.qr(G.cpu())
        Q, R = Q.to(dev), R.to(dev)
        Q = Q * torch.sign(torch.diagonal(R, dim1=1, dim2=2)).unsqueeze(1)
        lc = torch.rand(256, dtype=f64, device=dev, generator=gen) * 6 + 2
        lam = torch.pow(10.0, -lc[:, None] * torch.arange(n, dtype=f64, device=dev)[None, :] / (n - 1))
        Ab = torch.bmm(Q * lam[:, None, :], Q.transpose(1, 2).contiguous())
        A[b0:b0 + 256] = (Ab + Ab.transpose(1, 2)) / 2
        del G, Q, R, Ab
    Bd = torch.randn((batch, n, nrhs), dtype=f64, device=dev, generator=gen)
    log("posv256: inputs generated")

    def pm_buffer(rows, cols, dense=None):
        d, off, nb = bfa.bulk_layout(np.float64, ps, [rows] * batch, [cols] * batch, 0)
        buf = torch.zeros(max(nb, 8) // 8, dtype=f64, device=dev)
        stride = int(off[1]) // 8 if batch > 1 else nb // 8
        if dense is not None:
            buf.view(batch, stride)[:, : rows * _cn(ps, cols)] = _pm_pack(torch, dense, ps)
        desc = to_device(bfa.bulk_layout(np.float64, ps, [rows] * batch, [cols] * batch, buf.data_ptr())[0])
        return buf, desc, stride

    bufA, dA, sA = pm_buffer(n, n, A)
    bufB, dB, sB = pm_buffer(n, nrhs, Bd)
    bufX, dX, sX = pm_buffer(n, nrhs)
    bufL, dL, _ = pm_buffer(n, n)
    bufX2, dX2, _ = pm_buffer(n, nrhs)
    bufL2, dL2, _ = pm_buffer(n, n)
    idx = np.arange(batch, dtype=np.uint64) * bfa.DMAT.itemsize
    # FP64 sequence
    fa = _records(bfa.DFACT_ARG, batch)
    fa["m"], fa["n"] = n, n
    fa["sC"], fa["sD"] = dA.data_ptr() + idx, dL.data_ptr() + idx
    t1 = _records(bfa.DTRSM_ARG, batch)
    t1["m"], t1["n"], t1["alpha"] = n, nrhs, 1.0
    t1["sA"], t1["sB"], t1["sD"] = dL.data_ptr() + idx, dB.data_ptr() + idx, dX.data_ptr() + idx
    t2 = t1.copy()
    t2["sB"] = t2["sD"]
    dfa, dt1, dt2 = to_device(fa), to_device(t1), to_device(t2)
    # mixed entry
    info = torch.zeros(batch * bfa.SOLVE_INFO.itemsize, dtype=torch.uint8, device=dev)
    pr = _records(bfa.DPOSV_ARG, batch)
    pr["n"], pr["nrhs"] = n, nrhs
    pr["sA"], pr["sB"] = dA.data_ptr() + idx, dB.data_ptr() + idx
    pr["sX"], pr["sL"] = dX2.data_ptr() + idx, dL2.data_ptr() + idx
    pr["info"] = info.data_ptr() + np.arange(batch, dtype=np.uint64) * bfa.SOLVE_INFO.itemsize
    dpr = to_device(pr)
    hint = (n, nrhs, 0)

    def fp64():
        bfa.call("dpotrf_l", ps, batch, dfa.data_ptr(), (n, n, 0), sh)
        bfa.call("dtrsm_llnn", ps, batch, dt1.data_ptr(), hint, sh)
        bfa.call("dtrsm_lltn", ps, batch, dt2.data_ptr(), hint, sh)

    def mixed():
        bfa.posv_mixed(ps, batch, dpr.data_ptr(), tol, max_iter, hint, sh)

    ms64 = time_steps(torch, stream, fp64, args.steps, args.warmup)
    log("posv256 fp64 sequence: %.3f ms/step" % ms64)
    msmx = time_steps(torch, stream, mixed, args.steps, args.warmup)
    log("posv256 mixed: %.3f ms/step" % msmx)
    X64 = _pm_unpack(torch, bufX.view(batch, sX), ps, n, nrhs)
    Xmx = _pm_unpack(torch, bufX2.view(batch, sX), ps, n, nrhs)
    be64 = _posv_berr(torch, A, X64, Bd)
    bemx = _posv_berr(torch, A, Xmx, Bd)
    inf = info.cpu().numpy().view(bfa.SOLVE_INFO)
    paths = np.bincount(inf["path"], minlength=2)
    iters = inf["iters"][inf["path"] == 0]
    checks = {"berr_max_mixed": float(bemx.max()), "berr_max_fp64": float(be64.max()),
              "berr_target": tol, "mixed_path_counts": {"fp32_refined": int(paths[0]), "fp64_fallback": int(paths[1])},
              "refinement_iters_mean": float(iters.mean()) if iters.size else None,
              "refinement_iters_max": int(iters.max()) if iters.size else None,
              "pass": bool(bemx.max() <= 1.5e-14 and be64.max() <= 1.5e-14)}
    log("posv256 checks: %s" % json.dumps(checks))
    flops_fact = batch * n ** 3 / 3.0
    value = batch / (msmx * 1e-3)
    v64 = batch / (ms64 * 1e-3)
    achieved = flops_fact / (msmx * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        ns = min(batch, 64)
        hA = bufA.view(batch, sA)[:ns].cpu().numpy().copy()
        hB = bufB.view(batch, sB)[:ns].cpu().numpy().copy()
        hX = np.zeros_like(hB)
        hL = np.zeros_like(hA)
        def hdesc(arr, rows, cols):
            base = arr.ctypes.data
            return bfa.bulk_layout(np.float64, ps, [rows] * ns, [cols] * ns, base)[0]
        DA, DB, DX, DL = hdesc(hA, n, n), hdesc(hB, n, nrhs), hdesc(hX, n, nrhs), hdesc(hL, n, n)
        hi = np.arange(ns, dtype=np.uint64) * bfa.DMAT.itemsize
        hr = _records(bfa.DPOSV_ARG, ns)
        hr["n"], hr["nrhs"] = n, nrhs
        hr["sA"], hr["sB"], hr["sX"], hr["sL"] = (DA.ctypes.data + hi, DB.ctypes.data + hi,
                                                  DX.ctypes.data + hi, DL.ctypes.data + hi)
        cpu = cpu_time(lambda: oracle.lib().or_dposv_batch(ps, ns, hr.ctypes.data), ns, args.cpu_seconds,
                       "solves/s", "the first %d instances (potrf_l + trsm_llnn + trsm_lltn, FP64)" % ns)
    wl = {"workload": "dposv_mixed_batched vs dpotrf_l+dtrsm_llnn+dtrsm_lltn", "batch": batch, "n": n,
          "nrhs": nrhs, "ps": ps, "cond": "10^U[2,8]", "tol": tol, "max_iter": max_iter}
    out = {"metric": METRIC, "value": value, "unit": "solves/s (mixed precision, berr <= 1e-14)",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": msmx,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "synthetic, seeded (%d)" % args.seed, "config": wl,
           "fp64_solves_per_s": v64, "fp64_ms_per_step": ms64, "mixed_over_fp64": value / v64,
           "roofline": {"bound": "mfma", "achieved": achieved, "peak": 157.3, "unit": "TFLOP/s",
                        "frac": achieved / 157.3, "traffic": read_traffic("k_posv_mixed", "posv-2048x256"),
                        "flops_per_launch": flops_fact, "kernel": "k_posv_mixed<%d>" % ps,
                        "note": "factorisation flops n^3/3 per instance over the whole mixed step, against the FP32 matrix-core peak"},
           "cpu_baseline": cpu, "checks": checks}
    print(json.dumps(out), flush=True)


def cpu_time(fn, units, seconds, unit, sample):
    import oracle
    threads = oracle.lib().or_num_threads()
    done = 0
    t0 = time.perf_counter()
    while True:
        fn()
        done += units
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": done / el, "unit": unit, "cores": int(threads), "kind": "port",
            "sample": "%s, %d repetitions, %.1f s; CPU oracle (C restatement of blasfeo_ref, OpenMP over instances)"
                      % (sample, done // units, el)}


if __name__ == "__main__":
    main()
