"""Emulator cases (TEST INFRASTRUCTURE ONLY) for the left-side solves on the
record sets of the GPU tests (tests/trsm_cases.py), checked against the
oracle element for element (NaN and Inf included).
Usage: python case_trsm.py nonfinite <op> <thin: 0|1>"""
import ctypes
import sys

import numpy as np

from emu_harness import records, emu_lib, to_mats, bfa, oracle
import trsm_cases


def run(op, thin):
    ps = 4
    E = emu_lib("emu_lapack")
    re, ro = to_mats(ps, trsm_cases.nonfinite(ps, op, 970 + 3 * len(op) + thin))
    blk = records(bfa.DTRSM_ARG, re)
    print("emu %s nonfinite thin=%d batch=%d" % (op, thin, len(re)), flush=True)
    assert E.blasfeo_amd_tune_set(b"trsm_thin", 0, 1 if thin else 2) == 0
    rc = getattr(E, "blasfeo_amd_d%s_batched" % op)(ps, len(re), ctypes.c_void_p(blk.addr),
                                                    ctypes.byref(bfa.HINT(256, 4, 0)), None)
    assert rc == 0, E.emu_last_error()
    f = oracle.fn("d", op)
    for r in ro:
        f(ps, r["m"], r["n"], r["alpha"], r["sA"].desc.addr, r["ai"], r["aj"],
          r["sB"].desc.addr, r["bi"], r["bj"], r["sD"].desc.addr, r["di"], r["dj"])
    for i, (a, b) in enumerate(zip(re, ro)):
        gv, rv = a["sD"].content(), b["sD"].content()
        with np.errstate(invalid="ignore"):
            same = (gv == rv) | (np.isnan(gv) & np.isnan(rv)) | (np.abs(gv - rv) <= 1e-11 * (np.abs(gv) + np.abs(rv)))
        assert same.all(), "%s instance %d: %d differ, first %s emu=%r ref=%r" % (
            op, i, int((~same).sum()), np.argwhere(~same)[:1].tolist(), gv[~same][:1], rv[~same][:1])
        assert a["sD"].use_dA() == b["sD"].use_dA()
        assert a["sA"].use_dA() == b["sA"].use_dA()
    print("ok", flush=True)


if __name__ == "__main__":
    assert sys.argv[1] == "nonfinite"
    run(sys.argv[2], int(sys.argv[3]))
