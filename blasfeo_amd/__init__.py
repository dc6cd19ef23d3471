"""Python host side of the batched BLASFEO hot path for MI355X (gfx950).

The compute lives in the HIP shared library ``libblasfeo_amd.so`` built from
``blasfeo_amd/csrc`` (C ABI in ``include/blasfeo_amd.h``).  This module

* loads that library and declares its entry points (``lib()``),
* mirrors the C descriptor and argument-record structs as numpy dtypes,
* lays out batches of panel-major matrices in one byte buffer (``Layout``),
  so that the same layout can live in GPU memory (through torch, used only
  for device memory and streams) or in host memory, and
* converts between dense column-major arrays and panel-major storage,
  following ``BLASFEO_DMATEL`` (reference include/blasfeo_common.h:112).

There is no CPU fallback: every compute entry point needs the HIP library
and a GPU, and fails loudly without them.
"""
import ctypes
import os

import numpy as np

__all__ = ["lib", "lib_path", "Layout", "DMAT", "SMAT", "DGEMM_ARG", "SGEMM_ARG",
           "DTRSM_ARG", "STRSM_ARG", "DFACT_ARG", "SFACT_ARG", "DSYRK_DPOTRF_ARG",
           "RICCATI_ARG", "RICCATI_TRS_ARG", "RICCATI_SV_ARG", "RICCATI_MIXED_ARG", "DVEC", "DPACK_ARG", "DROWPE_ARG", "PM_DMAT", "PM_DGECP_ARG", "pm_dgecp", "DPOSV_ARG", "SOLVE_INFO", "HINT", "posv_mixed", "ric_solve_mixed", "pm_pack", "pm_unpack", "memsize", "check"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libblasfeo_amd.so")
_lib = None

# ---------------------------------------------------------------------------
# struct mirrors (include/blasfeo_amd.h)
# ---------------------------------------------------------------------------
_mat_fields = [("mem", "u8"), ("pA", "u8"), ("dA", "u8"), ("m", "i4"), ("n", "i4"),
               ("pm", "i4"), ("cn", "i4"), ("use_dA", "i4"), ("memsize", "i4")]
DMAT = np.dtype(_mat_fields, align=True)
SMAT = DMAT  # same byte layout (pointers + ints)
DVEC = np.dtype([("mem", "u8"), ("pa", "u8"), ("m", "i4"), ("pm", "i4"), ("memsize", "i4")],
                align=True)

_offs = [(f, "i4") for f in ("ai", "aj", "bi", "bj", "ci", "cj", "di", "dj")]
DGEMM_ARG = np.dtype([("m", "i4"), ("n", "i4"), ("k", "i4")] + _offs +
                     [("alpha", "f8"), ("beta", "f8"),
                      ("sA", "u8"), ("sB", "u8"), ("sC", "u8"), ("sD", "u8")], align=True)
SGEMM_ARG = np.dtype([("m", "i4"), ("n", "i4"), ("k", "i4")] + _offs +
                     [("alpha", "f4"), ("beta", "f4"),
                      ("sA", "u8"), ("sB", "u8"), ("sC", "u8"), ("sD", "u8")], align=True)
_toffs = [(f, "i4") for f in ("ai", "aj", "bi", "bj", "di", "dj")]
DTRSM_ARG = np.dtype([("m", "i4"), ("n", "i4")] + _toffs +
                     [("alpha", "f8"), ("sA", "u8"), ("sB", "u8"), ("sD", "u8")], align=True)
STRSM_ARG = np.dtype([("m", "i4"), ("n", "i4")] + _toffs +
                     [("alpha", "f4"), ("sA", "u8"), ("sB", "u8"), ("sD", "u8")], align=True)
DFACT_ARG = np.dtype([("m", "i4"), ("n", "i4"), ("ci", "i4"), ("cj", "i4"), ("di", "i4"),
                      ("dj", "i4"), ("sC", "u8"), ("sD", "u8"), ("ipiv", "u8")], align=True)
SFACT_ARG = DFACT_ARG
DSYRK_DPOTRF_ARG = np.dtype([("m", "i4"), ("n", "i4"), ("k", "i4")] + _offs +
                            [("sA", "u8"), ("sB", "u8"), ("sC", "u8"), ("sD", "u8")], align=True)
RICCATI_ARG = np.dtype([("N", "i4"), ("nx", "u8"), ("nu", "u8"), ("BAbt", "u8"),
                        ("RSQrq", "u8"), ("L", "u8"), ("work", "u8")], align=True)
PM_DMAT = np.dtype([("mem", "u8"), ("pA", "u8"), ("dA", "u8")] +
                   [(f, "i4") for f in ("m", "n", "pm", "cn", "use_dA", "ps", "memsize")], align=True)
PM_DGECP_ARG = np.dtype([(f, "i4") for f in ("m", "n", "ai", "aj", "bi", "bj")] +
                        [("sA", "u8"), ("sB", "u8")], align=True)
DPACK_ARG = np.dtype([(f, "i4") for f in ("m", "n", "lda", "ai", "aj")] +
                     [("A", "u8"), ("sA", "u8")], align=True)
DROWPE_ARG = np.dtype([("kmax", "i4"), ("ipiv", "u8"), ("sA", "u8")], align=True)
RICCATI_TRS_ARG = np.dtype([("N", "i4")] + [(f, "u8") for f in (
    "nx", "nu", "BAbt", "b", "rq", "L", "Pb", "ux", "pi", "work")], align=True)
RICCATI_SV_ARG = np.dtype([("N", "i4")] + [(f, "u8") for f in (
    "nx", "nu", "BAbt", "RSQrq", "L", "ux", "pi", "work_mat", "work_vec")], align=True)
RICCATI_MIXED_ARG = np.dtype([("N", "i4")] + [(f, "u8") for f in (
    "nx", "nu", "BAbt", "RSQrq", "L", "work_mat", "b", "rq", "Pb", "ux", "pi", "work_vec",
    "res_b", "res_rq", "d_Pb", "d_ux", "d_pi", "info")], align=True)
DPOSV_ARG = np.dtype([("n", "i4"), ("nrhs", "i4")] +
                     [(f, "i4") for f in ("ai", "aj", "bi", "bj", "xi", "xj", "li", "lj")] +
                     [("sA", "u8"), ("sB", "u8"), ("sX", "u8"), ("sL", "u8"), ("info", "u8")], align=True)
SOLVE_INFO = np.dtype([("iters", "i4"), ("path", "i4"), ("berr", "f8")], align=True)


class HINT(ctypes.Structure):
    _fields_ = [("max_m", ctypes.c_int), ("max_n", ctypes.c_int), ("max_k", ctypes.c_int)]


_BATCHED = [
    "dgemm_nt", "dgemm_nn", "dsyrk_ln", "dsyrk_ln_mn", "dtrsm_rltn", "dtrmm_rlnn",
    "dtrsm_llnn", "dtrsm_lltn", "dtrsm_llnu", "dtrsm_lunn", "drowpe",
    "dpotrf_l", "dpotrf_l_mn", "dsyrk_dpotrf_ln", "dgetrf_rp", "dgetrf_np",
    "sgemm_nt", "sgemm_nn", "ssyrk_ln", "strsm_rltn", "spotrf_l", "sgetrf_rp",
    "d_back_ric_trf", "d_back_ric_trs", "d_back_ric_sv", "dpack", "dpack_tran", "dunpack",
]


def lib_path():
    return _LIB


def lib():
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        raise RuntimeError("blasfeo_amd: %s is missing; run blasfeo_amd/build.py "
                           "(or __graft_entry__.build()) first" % _LIB)
    # The library binds to the HIP runtime (libamdhip64) already in the process.
    # PyTorch ships its own: load it first, so that both use that one runtime
    # (loaded the other way round, the library's first device allocation
    # failed in a process that then imported torch).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(_LIB)
    vp, ip, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    for name in _BATCHED:
        f = getattr(L, "blasfeo_amd_%s_batched" % name)
        f.argtypes = [ip, ip, vp, ctypes.POINTER(HINT), vp]
        f.restype = ip
    L.blasfeo_amd_d_back_ric_trf_batched_noalloc.argtypes = [ip, ip, vp, ctypes.POINTER(HINT), vp]
    L.blasfeo_amd_d_back_ric_trf_batched_noalloc.restype = ip
    L.blasfeo_amd_dposv_mixed_batched.argtypes = [ip, ip, vp, ctypes.c_double, ip,
                                                  ctypes.POINTER(HINT), vp]
    L.blasfeo_amd_dposv_mixed_batched.restype = ip
    L.blasfeo_amd_d_ric_solve_mixed_batched.argtypes = [ip, ip, vp, ctypes.c_double, ip,
                                                        ctypes.POINTER(HINT), vp]
    L.blasfeo_amd_d_ric_solve_mixed_batched.restype = ip
    L.blasfeo_amd_tune_set.argtypes = [ctypes.c_char_p, ip, ip]
    L.blasfeo_amd_tune_set.restype = ip
    L.blasfeo_amd_tune_get.argtypes = [ctypes.c_char_p, ip]
    L.blasfeo_amd_tune_get.restype = ip
    L.blasfeo_amd_version.restype = ctypes.c_char_p
    L.blasfeo_amd_last_error.restype = ctypes.c_char_p
    L.blasfeo_amd_set_ps.argtypes = [ip]
    L.blasfeo_amd_memsize_dmat.argtypes = [ip, ip, ip]
    L.blasfeo_amd_memsize_dmat.restype = sz
    L.blasfeo_amd_memsize_smat.argtypes = [ip, ip, ip]
    L.blasfeo_amd_memsize_smat.restype = sz
    L.blasfeo_memsize_dmat.argtypes = [ip, ip]
    L.blasfeo_memsize_dmat.restype = sz
    L.blasfeo_pm_memsize_dmat.argtypes = [ip, ip, ip]
    L.blasfeo_pm_memsize_dmat.restype = sz
    L.blasfeo_amd_create_dmat.argtypes = [ip, ip, ip, vp, vp]
    L.blasfeo_amd_create_smat.argtypes = [ip, ip, ip, vp, vp]
    L.blasfeo_create_dmat.argtypes = [ip, ip, vp, vp]
    _lib = L
    return L


def check(rc, what="call"):
    if rc != 0:
        raise RuntimeError("blasfeo_amd %s failed (%d): %s"
                           % (what, rc, lib().blasfeo_amd_last_error().decode()))


def call(name, ps, batch, args_ptr, hint=None, stream=0):
    """Invokes blasfeo_amd_<name>_batched with device pointers; raises on error."""
    f = getattr(lib(), "blasfeo_amd_%s_batched" % name)
    h = None
    if hint is not None:
        h = ctypes.byref(HINT(*[int(x) for x in hint]))
    check(f(ps, batch, ctypes.c_void_p(int(args_ptr)), h, ctypes.c_void_p(int(stream))), name)


def pm_dgecp(batch, args_ptr, hint=None, stream=0):
    """blasfeo_amd_pm_dgecp_batched (copy between panel sizes); raises on error."""
    L = lib()
    f = L.blasfeo_amd_pm_dgecp_batched
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(HINT), ctypes.c_void_p]
    f.restype = ctypes.c_int
    h = ctypes.byref(HINT(*[int(x) for x in hint])) if hint is not None else None
    check(f(batch, ctypes.c_void_p(int(args_ptr)), h, ctypes.c_void_p(int(stream))), "pm_dgecp")


def posv_mixed(ps, batch, args_ptr, tol, max_iter, hint, stream=0):
    """blasfeo_amd_dposv_mixed_batched with device pointers; raises on error."""
    h = ctypes.byref(HINT(*[int(x) for x in hint]))
    check(lib().blasfeo_amd_dposv_mixed_batched(ps, batch, ctypes.c_void_p(int(args_ptr)), float(tol),
                                                int(max_iter), h, ctypes.c_void_p(int(stream))),
          "dposv_mixed")


def ric_solve_mixed(ps, batch, args_ptr, tol, max_iter, hint, stream=0):
    """blasfeo_amd_d_ric_solve_mixed_batched with device pointers; raises on error."""
    h = ctypes.byref(HINT(*[int(x) for x in hint]))
    check(lib().blasfeo_amd_d_ric_solve_mixed_batched(ps, batch, ctypes.c_void_p(int(args_ptr)), float(tol),
                                                      int(max_iter), h, ctypes.c_void_p(int(stream))),
          "d_ric_solve_mixed")


# ---------------------------------------------------------------------------
# panel-major layout
# ---------------------------------------------------------------------------
def _pld(dtype, ps):
    # column rounding of the reference builds: double ps=8 -> 8, else 4
    return 8 if (np.dtype(dtype) == np.float64 and ps == 8) else 4


def memsize(dtype, ps, m, n):
    """Bytes of a panel-major matrix (auxiliary/d_aux_lib4.c:1519-1530)."""
    nc = _pld(dtype, ps)
    al = ps * nc
    pm = -(-m // ps) * ps
    cn = -(-n // nc) * nc
    tmp = -(-min(m, n) // al) * al
    size = (pm * cn + tmp) * np.dtype(dtype).itemsize
    return -(-size // 64) * 64


def pm_pack(A, ps, cn, pm=None):
    """Dense (m x n) -> flat panel-major array with panel stride cn."""
    A = np.asarray(A)
    m, n = A.shape
    pm = pm if pm is not None else -(-m // ps) * ps
    P = np.zeros((pm, cn), dtype=A.dtype)
    P[:m, :n] = A
    return P.reshape(pm // ps, ps, cn).transpose(0, 2, 1).reshape(-1)


def pm_unpack(flat, ps, cn, m, n):
    """Flat panel-major array -> dense (m x n)."""
    pm = -(-m // ps) * ps
    P = np.asarray(flat)[: pm * cn].reshape(pm // ps, cn, ps).transpose(0, 2, 1).reshape(pm, cn)
    return P[:m, :n].copy()


class Layout:
    """A batch of panel-major matrices packed into one byte buffer.

    ``add(m, n)`` reserves a matrix and returns its index; ``build(base)``
    returns the descriptor array (numpy ``DMAT``) for a buffer starting at
    address ``base`` (a GPU or a host address).  Each matrix starts on a
    256-byte boundary; its reciprocal-diagonal block follows the panels as in
    ``blasfeo_create_dmat`` (auxiliary/d_aux_lib4.c:1548-1570).
    """

    ALIGN = 256

    def __init__(self, dtype=np.float64, ps=4):
        self.dtype = np.dtype(dtype)
        self.ps = ps
        self.nc = _pld(self.dtype, ps)
        self.shapes = []
        self.offsets = []
        self.nbytes = 0

    def add(self, m, n):
        off = -(-self.nbytes // self.ALIGN) * self.ALIGN
        self.shapes.append((m, n))
        self.offsets.append(off)
        self.nbytes = off + memsize(self.dtype, self.ps, m, n)
        return len(self.shapes) - 1

    def __len__(self):
        return len(self.shapes)

    def geom(self, i):
        m, n = self.shapes[i]
        pm = -(-m // self.ps) * self.ps
        cn = -(-n // self.nc) * self.nc
        return m, n, pm, cn

    def build(self, base):
        d = np.zeros(len(self.shapes), dtype=DMAT)
        isz = self.dtype.itemsize
        al = self.ps * self.nc
        for i, (m, n) in enumerate(self.shapes):
            _, _, pm, cn = self.geom(i)
            addr = int(base) + self.offsets[i]
            d[i]["mem"] = addr
            d[i]["pA"] = addr
            d[i]["dA"] = addr + pm * cn * isz
            d[i]["m"], d[i]["n"], d[i]["pm"], d[i]["cn"] = m, n, pm, cn
            d[i]["use_dA"] = 0
            d[i]["memsize"] = memsize(self.dtype, self.ps, m, n)
        return d

    def host_buffer(self):
        return np.zeros(self.nbytes, dtype=np.uint8)

    def view(self, buf, i):
        """Flat element view (pm*cn) of matrix i inside a host byte buffer."""
        _, _, pm, cn = self.geom(i)
        off = self.offsets[i]
        return buf[off: off + pm * cn * self.dtype.itemsize].view(self.dtype)

    def dview(self, buf, i):
        """Reciprocal-diagonal view of matrix i inside a host byte buffer."""
        m, n, pm, cn = self.geom(i)
        off = self.offsets[i] + pm * cn * self.dtype.itemsize
        return buf[off: off + min(m, n) * self.dtype.itemsize].view(self.dtype)

    def put(self, buf, i, A, ai=0, aj=0):
        """Writes dense A into matrix i at offset (ai, aj)."""
        A = np.asarray(A, dtype=self.dtype)
        m, n, pm, cn = self.geom(i)
        flat = self.view(buf, i)
        ps = self.ps
        r = np.arange(A.shape[0])[:, None] + ai
        c = np.arange(A.shape[1])[None, :] + aj
        flat[(r - r % ps) * cn + c * ps + r % ps] = A

    def get(self, buf, i, ai=0, aj=0, m=None, n=None):
        mm, nn, pm, cn = self.geom(i)
        m = mm - ai if m is None else m
        n = nn - aj if n is None else n
        flat = self.view(buf, i)
        ps = self.ps
        r = np.arange(m)[:, None] + ai
        c = np.arange(n)[None, :] + aj
        return flat[(r - r % ps) * cn + c * ps + r % ps].copy()


def bulk_layout(dtype, ps, ms, ns, base, align=256):
    """Vectorised layout of many matrices in one buffer starting at `base`.

    Returns (descriptors as a DMAT array, offsets in bytes, total bytes).  Same
    rules as ``Layout`` (reference auxiliary/d_aux_lib4.c:1519-1570)."""
    dtype = np.dtype(dtype)
    ms = np.asarray(ms, dtype=np.int64)
    ns = np.asarray(ns, dtype=np.int64)
    nc = _pld(dtype, ps)
    al = ps * nc
    pm = -(-ms // ps) * ps
    cn = -(-ns // nc) * nc
    tmp = -(-np.minimum(ms, ns) // al) * al
    size = (pm * cn + tmp) * dtype.itemsize
    size = -(-size // 64) * 64
    padded = -(-size // align) * align
    offs = np.concatenate([[0], np.cumsum(padded)[:-1]]).astype(np.int64)
    d = np.zeros(len(ms), dtype=DMAT)
    addr = np.uint64(int(base)) + offs.astype(np.uint64)
    d["mem"] = addr
    d["pA"] = addr
    d["dA"] = addr + (pm * cn * dtype.itemsize).astype(np.uint64)
    d["m"], d["n"], d["pm"], d["cn"] = ms, ns, pm, cn
    d["memsize"] = size
    total = int(offs[-1] + padded[-1]) if len(ms) else 0
    return d, offs, total
