"""Kernel sources run on the CPU under AddressSanitizer (tests/emu/): the
library's .hip files compiled for the host against an emulation of the HIP
device model, each matrix in its own exactly-sized allocation, results
checked against the oracle with the reference's rule (tol 1e-11).  A kernel
access outside a matrix, descriptor, record array or the launch's LDS fails
the case with the source line.  No GPU is used."""
import glob
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")


def _asan_runtime():
    c = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return c[-1] if c else None


@pytest.fixture(scope="module")
def emu_build():
    if _asan_runtime() is None or not os.path.exists("/opt/rocm/llvm/bin/clang++"):
        pytest.skip("clang++ with the AddressSanitizer runtime is not available")
    r = subprocess.run(["make", "-s", "-j8", "-C", EMU], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return True


def run_case(script, *args):
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(), ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(EMU, script)] + [str(a) for a in args],
                       capture_output=True, text=True, env=env, timeout=600, cwd=EMU)
    assert r.returncode == 0 and r.stdout.rstrip().endswith("ok"), (r.stdout + r.stderr)[-4000:]


@pytest.mark.parametrize("ps", [4, 8])
def test_potrf_l_mn_split(emu_build, ps):
    """dpotrf_l_mn beyond the LDS-resident form (m up to 275, n up to 256):
    potrf_l of the top block, then the rows below solved with its
    reciprocals; the records of the GPU test test_potrf[potrf_l_mn]."""
    run_case("case_chol.py", "potrf_l_mn", ps)


@pytest.mark.parametrize("ps", [4, 8])
def test_syrk_dpotrf_ln(emu_build, ps):
    """dsyrk_dpotrf_ln(_mn) on test_syrk_dpotrf_ln's records: m up to 299,
    n <= 256, offsets, k up to 69 (LDS form up to 128, split form above)."""
    run_case("case_chol.py", "syrk_potrf", ps)


@pytest.mark.parametrize("case", ["potrf_lds", "potrf_lds_mn", "syrk_potrf_lds", "potrf_lds_mn_wide",
                                  "syrk_potrf_lds_wide"])
@pytest.mark.parametrize("ps", [4, 8])
def test_lds_size_buckets(emu_build, case, ps):
    """potrf_l, potrf_l_mn and dsyrk_dpotrf_ln with a bound of 128 rows: the
    plan sorts the instances into 16-row buckets and each bucket's persistent
    launch, its LDS sized for the bucket, takes them from an atomic queue;
    sizes 1 .. 128 so that every bucket holds instances; _wide: eight-wave
    workgroups for the buckets above 32 rows."""
    run_case("case_chol.py", case, ps)


@pytest.mark.parametrize("ps", [4, 8])
def test_potrf_global_kernel(emu_build, ps):
    """k_potrf (potrf_l for n > 256; 16-column panels at full height, the
    earlier panels read back from D) on test_potrf_large's records: n = 257
    .. 400, offsets, in place."""
    run_case("case_chol.py", "potrf_large", ps)


@pytest.mark.parametrize("thin", [1, 0])
@pytest.mark.parametrize("op", ["trsm_llnn", "trsm_lltn", "trsm_llnu", "trsm_lunn"])
def test_trsm_left_nonfinite(emu_build, op, thin):
    """Left solves with a zero diagonal entry or a NaN in B (records of the GPU
    test test_trsm_left_nonfinite): Inf and NaN only where the reference has
    them, on the one-wavefront (thin) and the four-wave kernel."""
    run_case("case_trsm.py", "nonfinite", op, thin)


@pytest.mark.parametrize("wave", [1, 2])
@pytest.mark.parametrize("ps", [4, 8])
@pytest.mark.parametrize("op", ["gemm_nt", "gemm_nn", "syrk_ln", "syrk_ln_mn"])
def test_gemm_family(emu_build, op, ps, wave):
    """gemm_nt / gemm_nn / syrk_ln / syrk_ln_mn on uneven sizes up to 139 with
    offsets, in place on every fifth instance, k = 0 and edge shapes:
    wave = 1 one wavefront per 32 x 32 tile, wave = 2 the LDS-staged engine."""
    run_case("case_gemm.py", op, ps, wave)


@pytest.mark.parametrize("ps", [4, 8])
def test_riccati_trf(emu_build, ps):
    """The LDS-resident Riccati factorisation sweep over every footprint class,
    with the built-in launch shapes and with one wave per instance."""
    run_case("case_ric.py", "trf", ps)


@pytest.mark.parametrize("ps", [4, 8])
def test_riccati_sv(emu_build, ps):
    """The fused Riccati factorisation and solve over every LDS footprint class
    up to nu+nx = 127, nx = 64, each matrix and vector in its own block."""
    run_case("case_ric.py", "sv", ps)


@pytest.mark.parametrize("ps", [4, 8])
def test_riccati_mixed(emu_build, ps):
    """The mixed-precision Riccati solve: the FP32 sweep, the solve on its
    factors and the FP64 refinement to backward error <= 1e-14 (ux, pi within
    1e-9 of the oracle's FP64 factorisation + solve), and with tol 0 the FP64
    fallback for every instance (the oracle's results at the 1e-11 rule)."""
    run_case("case_ric.py", "mixed", ps)
