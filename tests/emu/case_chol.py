"""Emulator cases (TEST INFRASTRUCTURE ONLY) for the Cholesky family on the
record sets of the GPU tests (tests/chol_cases.py), each matrix in its own
exactly-sized block, checked against the oracle:
  potrf_l_mn <ps>   blasfeo_amd_dpotrf_l_mn_batched, test_potrf[potrf_l_mn]'s records
  syrk_potrf <ps>   blasfeo_amd_dsyrk_dpotrf_ln_batched, test_syrk_dpotrf_ln's records
  potrf_large <ps>  blasfeo_amd_dpotrf_l_batched for n > 256, test_potrf_large's records
  potrf_lds <ps>    blasfeo_amd_dpotrf_l(_mn)_batched with a bound of 128 rows: the
                    size-bucketed LDS-resident launches (queue mode)
  potrf_lds_mn <ps> the same for dpotrf_l_mn (m = n + U{0..19} <= 127)
  syrk_potrf_lds <ps>  blasfeo_amd_dsyrk_dpotrf_ln_batched with m <= 128, the same launches
  <case>_wide <ps>  potrf_lds, potrf_lds_mn or syrk_potrf_lds with eight-wave workgroups
                    for the buckets above 32 rows (tuning knob lds_threads = 512)
Usage: python case_chol.py <case> <ps>"""
import ctypes
import sys

from emu_harness import records, compare_outputs, emu_lib, to_mats, bfa, oracle
import chol_cases

HINT = bfa.HINT


def run(which, ps):
    E = emu_lib("emu_lapack")
    if which.endswith("_wide"):
        # the eight-wave form of the queued bucket launches (rows 33 .. 128)
        which = which[:-5]
        for b in range(2, 8):
            assert E.blasfeo_amd_tune_set(b"lds_threads", b, 512) == 0, E.emu_last_error()
    if which in ("potrf_l_mn", "potrf_large", "potrf_lds", "potrf_lds_mn"):
        small = [n for n in chol_cases.POTRF_SIZES if n <= 128]
        if which == "potrf_l_mn":
            op, cases, hint = "potrf_l_mn", chol_cases.potrf(ps, "potrf_l_mn", chol_cases.POTRF_SIZES, 300 + ps), (280, 256, 0)
        elif which == "potrf_lds":
            op, cases, hint = "potrf_l", chol_cases.potrf(ps, "potrf_l", small, 500 + ps), (128, 128, 0)
        elif which == "potrf_lds_mn":
            op, cases, hint = "potrf_l_mn", chol_cases.potrf(ps, "potrf_l_mn", [n for n in small if n <= 108], 520 + ps), (128, 128, 0)
        else:
            op, cases, hint = "potrf_l", chol_cases.potrf(ps, "potrf_l", chol_cases.POTRF_LARGE_SIZES, 760 + ps, inplace_at=5), (420, 420, 0)
        re, ro = to_mats(ps, cases)
        blk = records(bfa.DFACT_ARG, re)
        print("emu %s ps=%d batch=%d" % (op, ps, len(re)), flush=True)
        rc = getattr(E, "blasfeo_amd_d%s_batched" % op)(ps, len(re), ctypes.c_void_p(blk.addr),
                                                        ctypes.byref(HINT(*hint)), None)
        assert rc == 0, E.emu_last_error()
        f = oracle.fn("d", op)
        for r in ro:
            if op == "potrf_l":
                f(ps, r["m"], r["sC"].desc.addr, r["ci"], r["cj"], r["sD"].desc.addr, r["di"], r["dj"])
            else:
                f(ps, r["m"], r["n"], r["sC"].desc.addr, r["ci"], r["cj"], r["sD"].desc.addr, r["di"], r["dj"])
        compare_outputs(re, ro, "sD", 1e-11, "%s ps=%d" % (op, ps))
    elif which in ("syrk_potrf", "syrk_potrf_lds"):
        if which == "syrk_potrf":
            cases, hint = chol_cases.syrk_potrf(ps, 400 + ps), (300, 256, 80)
        else:
            cases, hint = chol_cases.syrk_potrf(ps, 440 + ps, count=24, mmax=129), (128, 128, 80)
        re, ro = to_mats(ps, cases)
        blk = records(bfa.DSYRK_DPOTRF_ARG, re)
        print("emu syrk_dpotrf_ln ps=%d batch=%d" % (ps, len(re)), flush=True)
        rc = E.blasfeo_amd_dsyrk_dpotrf_ln_batched(ps, len(re), ctypes.c_void_p(blk.addr),
                                                   ctypes.byref(HINT(*hint)), None)
        assert rc == 0, E.emu_last_error()
        f = oracle.fn("d", "syrk_potrf_ln_mn")
        for r in ro:
            f(ps, r["m"], r["n"], r["k"], r["sA"].desc.addr, r["ai"], r["aj"], r["sB"].desc.addr, r["bi"], r["bj"],
              r["sC"].desc.addr, r["ci"], r["cj"], r["sD"].desc.addr, r["di"], r["dj"])
        compare_outputs(re, ro, "sD", 1e-11, "syrk_dpotrf_ln ps=%d" % ps)
    else:
        raise SystemExit("unknown case " + which)
    print("ok", flush=True)


if __name__ == "__main__":
    run(sys.argv[1], int(sys.argv[2]))
