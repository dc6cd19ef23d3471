"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the reported CPU baseline.
The product (blasfeo_amd/) never imports it.

Every routine takes the same arguments as the reference call it restates,
with descriptors passed as host addresses of ``blasfeo_amd.DMAT`` records;
see oracle_impl.h for the reference file:line each one follows.
"""
import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

_V, _I, _D, _F = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_float

_SIGS = {
    "gemm_nt": "iiiRpiipiiRpiipii", "gemm_nn": "iiiRpiipiiRpiipii",
    "syrk_ln": "iiRpiipiiRpiipii", "syrk_ln_mn": "iiiRpiipiiRpiipii",
    "trsm_rltn": "iiRpiipiipii", "trmm_rlnn": "iiRpiipiipii",
    "trsm_llnn": "iiRpiipiipii", "trsm_lltn": "iiRpiipiipii",
    "trsm_llnu": "iiRpiipiipii", "trsm_lunn": "iiRpiipiipii", "rowpe": "ipp",
    "potrf_l": "ipiipii", "potrf_l_mn": "iipiipii",
    "syrk_potrf_ln": "iipiipiipiipii", "syrk_potrf_ln_mn": "iiipiipiipiipii",
    "getrf_rp": "iipiipiip", "getrf_np": "iipiipii", "ric_trf": "ippppp",
    "gemv_n": "iiRpiipiRpipi", "gemv_t": "iiRpiipiRpipi",
    "trmv_lnn": "ipiipipi", "trmv_ltn": "ipiipipi",
    "trsv_lnn_mn": "iipiipipi", "trsv_ltn_mn": "iipiipipi",
    "ric_trs": "ippppppppppp", "ric_sv": "ipppppppppp",
    "pack": "iipipii", "pack_tran": "iipipii", "unpack": "iipiipi",
}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        src = [os.path.join(_HERE, f) for f in ("oracle.c", "oracle_impl.h")]
        if not os.path.exists(_SO) or any(os.path.getmtime(s) > os.path.getmtime(_SO) for s in src):
            build()
        L = ctypes.CDLL(_SO)
        for name, sig in _SIGS.items():
            for pfx, real in (("or_d", _D), ("or_s", _F)):
                f = getattr(L, pfx + name)
                f.argtypes = [_I] + [{"i": _I, "p": _V, "R": real}[c] for c in sig]
                f.restype = None
        for name in ("dgemm_nt", "dgemm_nn", "dsyrk_ln", "dpotrf_l", "dgetrf_rp",
                     "sgetrf_rp", "dric_trf", "dric_trs", "dric_sv", "dposv",
                     "dpack", "dpack_tran", "dunpack"):
            f = getattr(L, "or_%s_batch" % name)
            f.argtypes = [_I, _I, _V]
            f.restype = None
        L.or_pm_dgecp.argtypes = [_I, _I, _V, _I, _I, _V, _I, _I]
        L.or_pm_dgecp.restype = None
        L.or_pm_dgecp_batch.argtypes = [_I, _V]
        L.or_pm_dgecp_batch.restype = None
        L.or_num_threads.restype = _I
        _lib = L
    return _lib


def fn(prec, name):
    """The oracle routine or_<prec><name> (prec 'd' or 's')."""
    return getattr(lib(), "or_%s%s" % (prec, name))
