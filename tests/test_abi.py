"""CPU checks of the C ABI: the library loads and exports every function
include/blasfeo_amd.h declares, the struct layouts seen by C match the
Python mirrors, the layout helpers follow the reference's memsize/create
rules (auxiliary/d_aux_lib4.c:1519-1570, d_aux_lib8.c:64-115) and argument
errors are reported without touching a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import blasfeo_amd as bfa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "blasfeo_amd.h")


def _declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"^[a-z][\w \*]*?\b(blasfeo_\w+)\(", txt, re.M)))


def test_exports_every_declared_symbol():
    L = bfa.lib()
    names = _declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "lay.c"
    checks = {
        "blasfeo_dmat": (bfa.DMAT, ["pA", "dA", "m", "n", "pm", "cn", "use_dA", "memsize"]),
        "blasfeo_amd_dgemm_arg": (bfa.DGEMM_ARG, ["k", "dj", "alpha", "beta", "sA", "sD"]),
        "blasfeo_amd_sgemm_arg": (bfa.SGEMM_ARG, ["alpha", "beta", "sA", "sD"]),
        "blasfeo_amd_dtrsm_arg": (bfa.DTRSM_ARG, ["dj", "alpha", "sA", "sD"]),
        "blasfeo_amd_strsm_arg": (bfa.STRSM_ARG, ["alpha", "sA", "sD"]),
        "blasfeo_amd_dfact_arg": (bfa.DFACT_ARG, ["dj", "sC", "sD", "ipiv"]),
        "blasfeo_amd_dsyrk_dpotrf_arg": (bfa.DSYRK_DPOTRF_ARG, ["dj", "sA", "sD"]),
        "blasfeo_amd_riccati_arg": (bfa.RICCATI_ARG, ["nx", "nu", "BAbt", "RSQrq", "L", "work"]),
        "blasfeo_amd_dposv_arg": (bfa.DPOSV_ARG, ["n", "nrhs", "ai", "lj", "sA", "sB", "sX", "sL", "info"]),
        "blasfeo_amd_solve_info": (bfa.SOLVE_INFO, ["iters", "path", "berr"]),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HDR, "int main(void){"]
    for s, (_, fields) in checks.items():
        lines.append('printf("%s size %%zu\\n", sizeof(struct %s));' % (s, s))
        for f in fields:
            lines.append('printf("%s %s %%zu\\n", offsetof(struct %s, %s));' % (s, f, s, f))
    lines.append('printf("blasfeo_smat size %zu\\n", sizeof(struct blasfeo_smat));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for line in out:
        if line.strip():
            a, b, c = line.split()
            got[(a, b)] = int(c)
    for s, (dt, fields) in checks.items():
        assert got[(s, "size")] == dt.itemsize, s
        for f in fields:
            assert got[(s, f)] == dt.fields[f][1], (s, f)
    assert got[("blasfeo_smat", "size")] == bfa.SMAT.itemsize


@pytest.mark.parametrize("ps,m,n,expect", [(4, 32, 32, 8448), (8, 32, 32, 8704), (4, 256, 256, 526336)])
def test_memsize_matches_reference_rule(ps, m, n, expect):
    L = bfa.lib()
    assert L.blasfeo_amd_memsize_dmat(ps, m, n) == expect
    assert bfa.memsize(np.float64, ps, m, n) == expect


def test_memsize_sweep_and_create():
    L = bfa.lib()
    for ps in (4, 8):
        for m in range(1, 40, 3):
            for n in range(1, 40, 5):
                assert L.blasfeo_amd_memsize_dmat(ps, m, n) == bfa.memsize(np.float64, ps, m, n)
    for ps in (4, 8, 16):
        assert L.blasfeo_amd_memsize_smat(ps, 13, 7) == bfa.memsize(np.float32, ps, 13, 7)
    buf = np.zeros(10000, dtype=np.uint8)
    d = np.zeros(1, dtype=bfa.DMAT)
    L.blasfeo_amd_create_dmat(8, 13, 7, d.ctypes.data, buf.ctypes.data)
    assert d[0]["pm"] == 16 and d[0]["cn"] == 8 and d[0]["m"] == 13 and d[0]["n"] == 7
    assert d[0]["dA"] - d[0]["pA"] == 16 * 8 * 8 and d[0]["use_dA"] == 0
    assert d[0]["memsize"] == bfa.memsize(np.float64, 8, 13, 7)
    lay = bfa.Layout(np.float64, 8)
    lay.add(13, 7)
    ld = lay.build(buf.ctypes.data)
    for f in ("pA", "dA", "m", "n", "pm", "cn", "memsize"):
        assert ld[0][f] == d[0][f]


def test_pm_pack_roundtrip():
    rng = np.random.default_rng(0)
    for ps in (4, 8, 16):
        A = rng.standard_normal((23, 9))
        flat = bfa.pm_pack(A, ps, 12)
        np.testing.assert_array_equal(bfa.pm_unpack(flat, ps, 12, 23, 9), A)
        # element (i,j) at (i - i%ps)*cn + j*ps + i%ps (include/blasfeo_common.h:112)
        i, j = 17, 5
        assert flat[(i - i % ps) * 12 + j * ps + i % ps] == A[i, j]


def test_host_pack_unpack_on_host_memory():
    L = bfa.lib()
    L.blasfeo_pack_dmat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.blasfeo_unpack_dmat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    assert L.blasfeo_amd_set_ps(4) == 0
    buf = np.zeros(L.blasfeo_memsize_dmat(11, 6) + 64, dtype=np.uint8)
    d = np.zeros(1, dtype=bfa.DMAT)
    L.blasfeo_create_dmat(11, 6, d.ctypes.data, buf.ctypes.data)
    A = np.asfortranarray(np.arange(30.0).reshape(6, 5))
    L.blasfeo_pack_dmat(6, 5, A.ctypes.data, 6, d.ctypes.data, 3, 1)
    flat = buf[: 12 * 8 * 8].view(np.float64)
    dense = bfa.pm_unpack(flat, 4, 8, 11, 6)
    np.testing.assert_array_equal(dense[3:9, 1:6], A)
    B = np.zeros((6, 5), order="F")
    L.blasfeo_unpack_dmat(6, 5, d.ctypes.data, 3, 1, B.ctypes.data, 6)
    np.testing.assert_array_equal(B, A)


def test_argument_errors_without_gpu():
    L = bfa.lib()
    assert L.blasfeo_amd_set_ps(5) == -1
    assert L.blasfeo_amd_set_ps(4) == 0
    f = L.blasfeo_amd_dgemm_nt_batched
    assert f(4, -1, None, None, None) == -1
    assert b"batch" in L.blasfeo_amd_last_error()
    assert f(4, 0, None, None, None) == 0
    assert f(4, 3, None, None, None) == -1
    assert f(5, 1, ctypes.c_void_p(8), None, None) == -1
    assert b"panel size" in L.blasfeo_amd_last_error()
    assert L.blasfeo_amd_dtrsm_rltn_batched(4, 1, ctypes.c_void_p(8), None, None) == -1
    h = bfa.HINT(100, 300, 0)
    assert L.blasfeo_amd_dtrsm_rltn_batched(4, 1, ctypes.c_void_p(8), ctypes.byref(h), None) == -4
    h = bfa.HINT(5000, 0, 0)
    assert L.blasfeo_amd_dpotrf_l_batched(4, 1, ctypes.c_void_p(8), ctypes.byref(h), None) == -4
    assert L.blasfeo_amd_d_back_ric_trf_batched(4, 1, ctypes.c_void_p(8), None, None) == -1
    # left-side solves need the bound of m; the mixed solve validates tol, max_iter and the hint
    assert L.blasfeo_amd_dtrsm_llnn_batched(4, 1, ctypes.c_void_p(8), None, None) == -1
    h = bfa.HINT(3000, 4, 0)
    assert L.blasfeo_amd_dtrsm_lltn_batched(4, 1, ctypes.c_void_p(8), ctypes.byref(h), None) == -4
    ok = bfa.HINT(256, 4, 0)
    f = L.blasfeo_amd_dposv_mixed_batched
    assert f(4, 1, ctypes.c_void_p(8), 0.0, 10, ctypes.byref(ok), None) == -1
    assert f(4, 1, ctypes.c_void_p(8), 1e-14, -1, ctypes.byref(ok), None) == -1
    assert f(4, 1, ctypes.c_void_p(8), 1e-14, 10, None, None) == -1
    assert f(6, 1, ctypes.c_void_p(8), 1e-14, 10, ctypes.byref(ok), None) == -1
    assert f(4, 0, None, 1e-14, 10, ctypes.byref(ok), None) == 0


def test_tuning_knobs_without_gpu():
    """Launch-shape knobs: defaults, validation, and the Python footprint
    rule used by the autotuner agreeing with the class rule of the kernel
    (classes by LDS bytes; one-wave classes only for nu+nx <= 64)."""
    from blasfeo_amd import autotune as at
    assert at.get_knob("ric_per_cu", 3) == 0
    at.set_knob("ric_per_cu", 3, 7)
    assert at.get_knob("ric_per_cu", 3) == 7
    at.set_knob("ric_threads", 2, 64)
    with pytest.raises(RuntimeError):
        at.set_knob("ric_threads", 6, 64)      # one wave needs nu+nx <= 64
    with pytest.raises(RuntimeError):
        at.set_knob("ric_per_cu", 8, 1)        # no such class
    with pytest.raises(RuntimeError):
        at.set_knob("gemm_grid_y", 0, 65)
    with pytest.raises(RuntimeError):
        at.set_knob("nonsense", 0, 1)
    at.set_knob("potrf_bw", 0, 2)
    assert at.get_knob("potrf_bw", 0) == 2
    with pytest.raises(RuntimeError):
        at.set_knob("potrf_bw", 0, 3)          # 1, 2 or 4 tile columns per block panel
    at.set_knob("trsm_thin", 0, 2)
    assert at.get_knob("trsm_thin", 0) == 2
    with pytest.raises(RuntimeError):
        at.set_knob("trsm_thin", 0, 3)         # 0 built-in, 1 always, 2 never
    at.set_knob("lds_threads", 7, 512)
    assert at.get_knob("lds_threads", 7) == 512
    with pytest.raises(RuntimeError):
        at.set_knob("lds_threads", 1, 256)      # up to 32 rows: one wave, no knob
    with pytest.raises(RuntimeError):
        at.set_knob("lds_threads", 4, 128)      # 256 or 512 threads
    at.reset()
    assert at.get_knob("lds_threads", 7) == 0
    assert at.get_knob("potrf_bw", 0) == 0 and at.get_knob("trsm_thin", 0) == 0
    assert at.get_knob("ric_per_cu", 3) == 0 and at.get_knob("ric_threads", 2) == 0
    # footprints of the resident form: largest bench stage (nu+nx = 80, nx = 64)
    # fits two per CU, nu+nx = 64, nx = 48 three per CU
    assert at.ric_footprint(80, 64) == 76816 and at.ric_class(80, 64) == 6
    assert at.ric_class(64, 48) == 5
    assert at.ric_class(65, 8) >= 5          # nu+nx > 64 never takes a one-wave class
    assert at.ric_class(60, 70) == 8 and at.ric_class(130, 10) == 8



def test_trapezoid_entries_reject_large_bounds():
    """Trapezoid potrf and fused syrk+potrf: m up to 1024 (the LDS-resident
    form up to 128, above it potrf_l of the top block plus the solve of the
    rows below), n up to 256 when m > 128; larger bounds are refused before
    any launch with BLASFEO_AMD_ETOOBIG."""
    L = bfa.lib()
    for mx in ((300, 0, 0), (300, 280, 0), (1100, 100, 0)):
        h = bfa.HINT(*mx)
        assert L.blasfeo_amd_dpotrf_l_mn_batched(4, 1, ctypes.c_void_p(8), ctypes.byref(h), None) == -4
        assert L.blasfeo_amd_dsyrk_dpotrf_ln_batched(4, 1, ctypes.c_void_p(8), ctypes.byref(h), None) == -4


def test_riccati_sv_bounds():
    """The fused Riccati factorisation and solve keeps every stage resident in
    LDS: nu+nx up to 127 (its augmented stage, 128 rows, still fits the
    largest footprint class) and nx up to 64; larger bounds are refused
    before any launch with BLASFEO_AMD_ETOOBIG."""
    from blasfeo_amd import autotune as at
    assert at.ric_class(128, 64) < 8
    L = bfa.lib()
    for mx in ((128, 64, 0), (100, 65, 0)):
        h = bfa.HINT(*mx)
        assert L.blasfeo_amd_d_back_ric_sv_batched(4, 1, ctypes.c_void_p(8), ctypes.byref(h), None) == -4
    assert L.blasfeo_amd_d_back_ric_sv_batched(4, 1, ctypes.c_void_p(8), None, None) == -1


def test_riccati_mixed_bounds():
    """The mixed-precision Riccati solve keeps the FP32 sweep in LDS: nu+nx up
    to 128 and nx up to 64; larger bounds and a negative maxit or tol are
    refused before any launch."""
    L = bfa.lib()
    f = L.blasfeo_amd_d_ric_solve_mixed_batched
    for mx in ((129, 64, 0), (100, 65, 0)):
        assert f(4, 1, ctypes.c_void_p(8), 1e-14, 3, ctypes.byref(bfa.HINT(*mx)), None) == -4
    h = bfa.HINT(80, 64, 0)
    assert f(4, 1, ctypes.c_void_p(8), 1e-14, -1, ctypes.byref(h), None) == -1
    assert f(4, 1, ctypes.c_void_p(8), -1.0, 3, ctypes.byref(h), None) == -1
    assert f(4, 1, ctypes.c_void_p(8), 1e-14, 3, None, None) == -1
