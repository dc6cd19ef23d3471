"""The library used from C the way a BLASFEO caller uses it: the program
tests/c_host/drop_in.c (built by __graft_entry__.build()) calls the
reference-named single-instance routines (blasfeo_dgemm_nt, blasfeo_dpotrf_l,
blasfeo_dtrsm_rltn with its use_dA side effects, blasfeo_dsyrk_dpotrf_ln,
blasfeo_dgetrf_rp with host pivots) and one batched entry on GPU memory, and
compares every byte with the CPU oracle's result at the reference's rule
(tol 1e-11), use_dA and pivots exactly, for panel sizes 4 and 8."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c_host", "drop_in")


def test_c_host_drop_in(gpu):
    assert os.path.exists(EXE), "tests/c_host/drop_in is not built (run __graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
    assert r.stdout.count("\nok ") == 2 * 13
