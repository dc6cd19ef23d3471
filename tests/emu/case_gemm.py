"""Emulator cases (TEST INFRASTRUCTURE ONLY) for the gemm family on the
record sets of tests/gemm_cases.py, checked against the oracle at the
reference's rule (tol 1e-11).
Usage: python case_gemm.py <op> <ps> <wave: 1 = one wavefront per tile, 2 = the LDS engine>"""
import ctypes
import sys

from emu_harness import records, compare_outputs, emu_lib, to_mats, bfa, oracle
import gemm_cases


def run(op, ps, wave):
    E = emu_lib("emu_lapack")
    re, ro = to_mats(ps, gemm_cases.uneven(ps, op, 90 + ps + 7 * len(op)))
    blk = records(bfa.DGEMM_ARG, re)
    print("emu %s ps=%d wave=%d batch=%d" % (op, ps, wave, len(re)), flush=True)
    assert E.blasfeo_amd_tune_set(b"gemm_wave", 0, wave) == 0
    rc = getattr(E, "blasfeo_amd_d%s_batched" % op)(ps, len(re), ctypes.c_void_p(blk.addr),
                                                    ctypes.byref(bfa.HINT(140, 140, 140)), None)
    assert rc == 0, E.emu_last_error()
    f = oracle.fn("d", op)
    for r in ro:
        if op == "syrk_ln":
            f(ps, r["m"], r["k"], r["alpha"], r["sA"].desc.addr, r["ai"], r["aj"], r["sB"].desc.addr, r["bi"], r["bj"],
              r["beta"], r["sC"].desc.addr, r["ci"], r["cj"], r["sD"].desc.addr, r["di"], r["dj"])
        else:
            f(ps, r["m"], r["n"], r["k"], r["alpha"], r["sA"].desc.addr, r["ai"], r["aj"], r["sB"].desc.addr,
              r["bi"], r["bj"], r["beta"], r["sC"].desc.addr, r["ci"], r["cj"], r["sD"].desc.addr, r["di"], r["dj"])
    compare_outputs(re, ro, "sD", 1e-11, "%s ps=%d" % (op, ps))
    print("ok", flush=True)


if __name__ == "__main__":
    run(sys.argv[1], int(sys.argv[2]), int(sys.argv[3]))
