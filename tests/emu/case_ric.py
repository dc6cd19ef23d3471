"""Emulator cases (TEST INFRASTRUCTURE ONLY) for the Riccati sweeps on
batches whose sizes reach every LDS footprint class up to nu+nx = 127,
nx = 64, every matrix and every vector in its own exactly-sized block:
  sv <ps>   blasfeo_amd_d_back_ric_sv_batched on an augmented batch: the
            factors (with their last, linear-term row), the reciprocal
            diagonals, ux and pi against the oracle's d_back_ric_sv;
  trf <ps>  blasfeo_amd_d_back_ric_trf_batched (every class launch shape,
            the one-wave kernels included): factors and reciprocal
            diagonals against the oracle's d_back_ric_trf.
Usage: python case_ric.py sv|trf <ps>"""
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


if __name__ == "__main__":
    {"sv": run, "trf": run_trf}[sys.argv[1]](int(sys.argv[2]))
