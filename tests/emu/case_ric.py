"""Emulator cases (TEST INFRASTRUCTURE ONLY) for the Riccati sweeps on
batches whose sizes reach every LDS footprint class up to nu+nx = 127,
nx = 64, every matrix and every vector in its own exactly-sized block:
  sv <ps>   blasfeo_amd_d_back_ric_sv_batched on an augmented batch: the
            factors (with their last, linear-term row), the reciprocal
            diagonals, ux and pi against the oracle's d_back_ric_sv;
  trf <ps>  blasfeo_amd_d_back_ric_trf_batched (every class launch shape,
            the one-wave kernels included): factors and reciprocal
            diagonals against the oracle's d_back_ric_trf;
  mixed <ps> blasfeo_amd_d_ric_solve_mixed_batched (the FP32 sweep, the
            solve, the FP64 refinement): ux and pi against the oracle's
            d_back_ric_trf + d_back_ric_trs, and every instance's backward
            error at or below the tolerance.
Usage: python case_ric.py sv|trf|mixed <ps>"""
import ctypes
import sys

import numpy as np

from emu_harness import Block, compare, emu_lib, bfa, oracle
from blasfeo_amd.riccati import RiccatiBatch, RiccatiSv

HINT = bfa.HINT

NXS = [1, 3, 8, 17, 30, 40, 64, 64, 50]
NUS = [0, 2, 5, 16, 1, 33, 20, 63, 9]


def split(layout, buf, desc_fields):
    """Each matrix of ``layout`` copied from ``buf`` into its own block; returns
    (blocks, fix) where fix(d) rewrites a descriptor array built at base 0 to
    point at the blocks."""
    blocks = []
    for i in range(len(layout)):
        m, n, pm, cn = layout.geom(i)
        b = Block(bfa.memsize(np.float64, layout.ps, m, n))
        b.write(buf[layout.offsets[i]: layout.offsets[i] + b.nbytes])
        blocks.append(b)
    by_off = {off: b for off, b in zip(layout.offsets, blocks)}

    def fix(d):
        for r in d:
            b = by_off[int(r["mem"])]
            shift = b.addr - int(r["mem"])
            for f in desc_fields:
                r[f] = int(r[f]) + shift
        return d
    return blocks, fix


def run_trf(ps):
    from blasfeo_amd.riccati import HostRiccati
    nxs = NXS + [2, 5, 12, 30, 47]
    nus = NUS + [1, 4, 3, 2, 16]
    rb = RiccatiBatch(nxs, nus, N=3, ps=ps)
    rng = np.random.default_rng(950 + ps)
    inp = rb.fill_inputs(rng)
    in_blocks, fix_in = split(rb.inp, inp, ("mem", "pA", "dA"))
    out_blocks, fix_out = split(rb.out, rb.out.host_buffer(), ("mem", "pA", "dA"))
    babt, rsq, L, work, nx, nu = rb.tables(0, 0)
    fix_in(babt), fix_in(rsq), fix_out(L), fix_out(work)
    tbl = {}
    for k, v in (("babt", babt), ("rsq", rsq), ("L", L), ("work", work), ("nx", nx), ("nu", nu)):
        tbl[k] = Block(v.nbytes)
        tbl[k].write(v)
    a = rb.args(tbl["babt"].addr, tbl["rsq"].addr, tbl["L"].addr, tbl["work"].addr, tbl["nx"].addr, tbl["nu"].addr)
    ablk = Block(a.nbytes)
    ablk.write(a)
    E = emu_lib("emu_lapack")
    host = HostRiccati(rb, inp)
    oracle.lib().or_dric_trf_batch(ps, rb.batch, host.argv.ctypes.data)
    # the built-in shapes, then one wave per instance on the classes up to 40 KB
    for threads in (0, 64):
        for c in range(5):
            E.blasfeo_amd_tune_set(b"ric_threads", c, threads)
        print("emu d_back_ric_trf ps=%d batch=%d N=%d threads=%d" % (ps, rb.batch, rb.N, threads), flush=True)
        rc = E.blasfeo_amd_d_back_ric_trf_batched(ps, rb.batch, ctypes.c_void_p(ablk.addr),
                                                  ctypes.byref(HINT(rb.max_nk, rb.max_nx, 0)), None)
        assert rc == 0, E.emu_last_error()
        Lout = tbl["L"].read(bfa.DMAT)
        for i in range(rb.batch):
            for k in range(rb.N + 1):
                j = rb.i_L[i][k]
                m, n, pm, cn = rb.out.geom(j)
                got = out_blocks[j].read(np.float64)
                want = np.frombuffer(host.outbuf[rb.out.offsets[j]: rb.out.offsets[j] + out_blocks[j].nbytes], np.float64)
                what = "threads %d instance %d stage %d (%dx%d)" % (threads, i, k, m, n)
                G = bfa.pm_unpack(got[: pm * cn], ps, cn, m, n)
                W = bfa.pm_unpack(want[: pm * cn], ps, cn, m, n)
                compare(np.tril(G).reshape(-1), np.tril(W).reshape(-1), 1e-11, "factor " + what)
                ud = int(Lout[i * (rb.N + 1) + k]["use_dA"])
                assert ud == int(host.t[2][i * (rb.N + 1) + k]["use_dA"]), ("use_dA", what, ud)
                compare(got[pm * cn: pm * cn + n], want[pm * cn: pm * cn + n], 1e-11, "dA " + what)
        for b in out_blocks:
            b.write(np.zeros(b.nbytes, dtype=np.uint8))
    print("ok", flush=True)


def run(ps):
    rb = RiccatiBatch(NXS, NUS, N=3, ps=ps, aug=True)
    sv = RiccatiSv(rb)
    rs = sv.rs
    rng = np.random.default_rng(900 + ps)
    inp = rb.fill_inputs(rng)

    # emulator copy: one block per matrix, per vector, per table
    in_blocks, fix_in = split(rb.inp, inp, ("mem", "pA", "dA"))
    out_blocks, fix_out = split(rb.out, rb.out.host_buffer(), ("mem", "pA", "dA"))
    babt, rsq, L, work, nx, nu = rb.tables(0, 0)
    fix_in(babt), fix_in(rsq), fix_out(L), fix_out(work)
    vtab = rs.table(0)
    vblocks = []
    for r in vtab:
        b = Block(int(r["memsize"]))
        b.write(np.zeros(int(r["memsize"]), dtype=np.uint8))
        r["mem"] = r["pa"] = b.addr
        vblocks.append(b)
    tbl = {k: Block(v.nbytes) for k, v in (("babt", babt), ("rsq", rsq), ("L", L), ("work", work),
                                           ("nx", nx), ("nu", nu), ("v", vtab))}
    for k, v in (("babt", babt), ("rsq", rsq), ("L", L), ("work", work), ("nx", nx), ("nu", nu), ("v", vtab)):
        tbl[k].write(v)
    a = sv.args(tbl["v"].addr, tbl["babt"].addr, tbl["rsq"].addr, tbl["L"].addr, tbl["work"].addr,
                tbl["nx"].addr, tbl["nu"].addr)
    ablk = Block(a.nbytes)
    ablk.write(a)

    E = emu_lib("emu_lapack")
    print("emu d_back_ric_sv ps=%d batch=%d N=%d" % (ps, rb.batch, rb.N), flush=True)
    rc = E.blasfeo_amd_d_back_ric_sv_batched(ps, rb.batch, ctypes.c_void_p(ablk.addr),
                                             ctypes.byref(HINT(rb.max_nk, rb.max_nx, 0)), None)
    assert rc == 0, E.emu_last_error()

    # oracle copy: one buffer per kind
    hin = np.array(inp, copy=True)
    hout = rb.out.host_buffer()
    hvec = np.zeros(rs.nbytes, dtype=np.uint8)
    t = rb.tables(hin.ctypes.data, hout.ctypes.data)
    hbabt, hrsq, hL, hwork, hnx, hnu = t
    htab = rs.table(hvec.ctypes.data)
    ha = sv.args(htab.ctypes.data, hbabt.ctypes.data, hrsq.ctypes.data, hL.ctypes.data, hwork.ctypes.data,
                 hnx.ctypes.data, hnu.ctypes.data)
    oracle.lib().or_dric_sv_batch(ps, rb.batch, ha.ctypes.data)

    Lout = tbl["L"].read(bfa.DMAT)
    for i in range(rb.batch):
        for k in range(rb.N + 1):
            j = rb.i_L[i][k]
            m, n, pm, cn = rb.out.geom(j)
            got = out_blocks[j].read(np.float64)
            want = np.frombuffer(hout[rb.out.offsets[j]: rb.out.offsets[j] + out_blocks[j].nbytes], np.float64)
            what = "instance %d stage %d (%dx%d)" % (i, k, m, n)
            G = bfa.pm_unpack(got[: pm * cn], ps, cn, m, n)
            W = bfa.pm_unpack(want[: pm * cn], ps, cn, m, n)
            compare(np.tril(G).reshape(-1), np.tril(W).reshape(-1), 1e-11, "factor " + what)
            ud = int(Lout[i * (rb.N + 1) + k]["use_dA"])
            assert ud == int(hL[i * (rb.N + 1) + k]["use_dA"]), ("use_dA", what, ud)
            if ud:
                compare(got[pm * cn: pm * cn + n], want[pm * cn: pm * cn + n], 1e-11, "dA " + what)
        for kind, ks in (("ux", range(rb.N + 1)), ("pi", range(rb.N))):
            for k in ks:
                j = rs.index(i, kind, k)
                m = int(rs.sizes[j])
                got = vblocks[j].read(np.float64, m)
                want = rs.vec(hvec, i, kind, k)
                compare(got, want, 1e-11, "%s[%d] instance %d" % (kind, k, i))
    print("ok", flush=True)


def vec_blocks(rs, buf):
    """Every vector of a RiccatiSolve buffer in its own block; returns the
    blocks and the descriptor table pointing at them."""
    tab = rs.table(0)
    blocks = []
    for j, r in enumerate(tab):
        b = Block(int(r["memsize"]))
        o = int(rs.offs[j])
        b.write(np.asarray(buf)[o: o + int(r["memsize"])])
        r["mem"] = r["pa"] = b.addr
        blocks.append(b)
    return blocks, tab


def run_mixed(ps):
    from blasfeo_amd.riccati import HostRiccati, RiccatiSolve
    nxs = NXS + [2, 12]
    nus = NUS + [1, 3]
    rb = RiccatiBatch(nxs, nus, N=3, ps=ps)
    rs = RiccatiSolve(rb)
    rng = np.random.default_rng(970 + ps)
    inp = rb.fill_inputs(rng)
    sin = rs.fill_inputs(rng)
    in_blocks, fix_in = split(rb.inp, inp, ("mem", "pA", "dA"))
    out_blocks, fix_out = split(rb.out, rb.out.host_buffer(), ("mem", "pA", "dA"))
    babt, rsq, L, work, nx, nu = rb.tables(0, 0)
    fix_in(babt), fix_in(rsq), fix_out(L), fix_out(work)
    vblocks, vtab = vec_blocks(rs, sin)
    cblocks, ctab = vec_blocks(rs, np.zeros(rs.nbytes, dtype=np.uint8))
    tbl = {}
    for k, v in (("babt", babt), ("rsq", rsq), ("L", L), ("work", work), ("nx", nx), ("nu", nu), ("v", vtab),
                 ("c", ctab)):
        tbl[k] = Block(v.nbytes)
        tbl[k].write(v)
    info = Block(bfa.SOLVE_INFO.itemsize * rb.batch)
    N = rb.N
    a = np.zeros(rb.batch, dtype=bfa.RICCATI_MIXED_ARG)
    idx = np.arange(rb.batch, dtype=np.uint64)
    a["N"] = N
    a["nx"] = tbl["nx"].addr + idx * 4 * (N + 1)
    a["nu"] = tbl["nu"].addr + idx * 4 * (N + 1)
    a["BAbt"] = tbl["babt"].addr + idx * bfa.DMAT.itemsize * N
    a["RSQrq"] = tbl["rsq"].addr + idx * bfa.DMAT.itemsize * (N + 1)
    a["L"] = tbl["L"].addr + idx * bfa.DMAT.itemsize * (N + 1)
    a["work_mat"] = tbl["work"].addr + idx * bfa.DMAT.itemsize
    base = np.uint64(tbl["v"].addr) + idx * np.uint64(bfa.DVEC.itemsize * rs.per)
    cbase = np.uint64(tbl["c"].addr) + idx * np.uint64(bfa.DVEC.itemsize * rs.per)
    for f, kind in (("b", "b"), ("rq", "rq"), ("Pb", "Pb"), ("ux", "ux"), ("pi", "pi"), ("work_vec", "work")):
        a[f] = base + np.uint64(bfa.DVEC.itemsize * rs.start[kind])
    for f, kind in (("res_b", "b"), ("res_rq", "rq"), ("d_Pb", "Pb"), ("d_ux", "ux"), ("d_pi", "pi")):
        a[f] = cbase + np.uint64(bfa.DVEC.itemsize * rs.start[kind])
    a["info"] = np.uint64(info.addr) + idx * np.uint64(bfa.SOLVE_INFO.itemsize)
    ablk = Block(a.nbytes)
    ablk.write(a)
    E = emu_lib("emu_lapack")
    E.blasfeo_amd_d_ric_solve_mixed_batched.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_double,
                                                        ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    # tol 0: no instance can converge, every one takes the FP64 fallback (then
    # the results are the FP64 factorisation and solve, 1e-11 rule); then
    # tol 1e-14 from the same inputs: the mixed path
    for tol, rule in ((0.0, 1e-11), (1e-14, 1e-9)):
        for b, src in zip(vblocks, vec_blocks(rs, sin)[0]):
            b.write(src.read())
        run_mixed_once(E, ps, rb, rs, N, ablk, info, tol, rule, vblocks, inp, sin)
    print("ok", flush=True)


def run_mixed_once(E, ps, rb, rs, N, ablk, info, tol, rule, vblocks, inp, sin):
    from blasfeo_amd.riccati import HostRiccati
    print("emu d_ric_solve_mixed ps=%d batch=%d N=%d tol=%g" % (ps, rb.batch, N, tol), flush=True)
    rc = E.blasfeo_amd_d_ric_solve_mixed_batched(ps, rb.batch, ablk.addr, tol, 4,
                                                 ctypes.addressof(HINT(rb.max_nk, rb.max_nx, 0)), None)
    assert rc == 0, E.emu_last_error()
    infos = info.read(bfa.SOLVE_INFO)
    if tol == 0.0:
        # an instance may reach a zero residual exactly (then it is closed)
        assert ((infos["path"] == 1) & (infos["iters"] == 4) | (infos["berr"] == 0)).all(), infos
    else:
        assert (infos["berr"] <= tol).all() and (infos["path"] == 0).all(), infos
    # oracle: FP64 factorisation and solve
    host = HostRiccati(rb, inp)
    oracle.lib().or_dric_trf_batch(ps, rb.batch, host.argv.ctypes.data)
    hsol = sin.copy()
    htab = rs.table(hsol.ctypes.data)
    hb, hr, hL, hw, hnx, hnu = host.t
    targs = rs.args(htab.ctypes.data, hb.ctypes.data, hL.ctypes.data, hnx.ctypes.data, hnu.ctypes.data)
    oracle.lib().or_dric_trs_batch(ps, rb.batch, targs.ctypes.data)
    for i in range(rb.batch):
        for kind, ks in (("ux", range(N + 1)), ("pi", range(N))):
            for k in ks:
                j = rs.index(i, kind, k)
                got = vblocks[j].read(np.float64, int(rs.sizes[j]))
                want = rs.vec(hsol, i, kind, k)
                compare(got, want, rule, "%s[%d] instance %d (path %d, iters %d, berr %.2e)"
                        % (kind, k, i, infos[i]["path"], infos[i]["iters"], infos[i]["berr"]))
    print("paths", np.bincount(infos["path"], minlength=2).tolist(), "iters", infos["iters"].tolist(),
          "max berr %.2e" % infos["berr"].max(), flush=True)


if __name__ == "__main__":
    {"sv": run, "trf": run_trf, "mixed": run_mixed}[sys.argv[1]](int(sys.argv[2]))
