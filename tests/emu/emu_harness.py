"""Host side of the kernel emulator (TEST INFRASTRUCTURE ONLY).

Runs inside a process started with the AddressSanitizer runtime preloaded
(``run_case`` in tests/test_emu_lapack.py does that).  Every matrix, its
descriptor and every record array lives in its own ``malloc`` block of the
exact size the layout rules give (``blasfeo_memsize_dmat``), so an access a
kernel makes outside a block is reported by ASan with the kernel's source
line.  Each case is run twice from the same bytes: by the emulated kernel and
by the CPU oracle, and the results are compared with the reference's rule."""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import blasfeo_amd as bfa  # noqa: E402  (layouts and record dtypes only)
import oracle  # noqa: E402

_libc = ctypes.CDLL(None)
_libc.malloc.restype = ctypes.c_void_p
_libc.malloc.argtypes = [ctypes.c_size_t]
_libc.free.argtypes = [ctypes.c_void_p]


def emu_lib(name):
    L = ctypes.CDLL(os.path.join(HERE, os.environ.get("EMU_BUILD", "_build"), "lib%s.so" % name))
    L.emu_last_error.restype = ctypes.c_char_p
    return L


class Block:
    """One malloc'd block of exactly nbytes."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.addr = _libc.malloc(max(self.nbytes, 1))

    def write(self, arr):
        b = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
        assert b.size <= self.nbytes
        ctypes.memmove(self.addr, b.ctypes.data, b.size)

    def read(self, dtype=np.uint8, count=None):
        n = self.nbytes // np.dtype(dtype).itemsize if count is None else count
        out = np.empty(n, dtype=dtype)
        ctypes.memmove(out.ctypes.data, self.addr, out.nbytes)
        return out


class Mat:
    """A panel-major FP64 matrix in its own block and its descriptor in another."""

    def __init__(self, ps, m, n):
        self.ps, self.m, self.n = ps, m, n
        self.size = bfa.memsize(np.float64, ps, m, n)
        self.pm = -(-m // ps) * ps
        self.cn = -(-n // (8 if ps == 8 else 4)) * (8 if ps == 8 else 4)
        self.buf = Block(self.size)
        self.desc = Block(bfa.DMAT.itemsize)
        d = np.zeros(1, dtype=bfa.DMAT)
        d["mem"] = d["pA"] = self.buf.addr
        d["dA"] = self.buf.addr + self.pm * self.cn * 8
        d["m"], d["n"], d["pm"], d["cn"], d["memsize"] = m, n, self.pm, self.cn, self.size
        self.desc.write(d)

    def fill(self, A, fill=0.0):
        flat = np.full(self.size // 8, fill)
        P = np.zeros((self.pm, self.cn))
        P[:A.shape[0], :A.shape[1]] = A
        flat[: self.pm * self.cn] = bfa.pm_pack(P, self.ps, self.cn, self.pm)
        self.buf.write(flat)

    def content(self):
        return self.buf.read(np.float64)

    def use_dA(self):
        return int(self.desc.read(bfa.DMAT)[0]["use_dA"])


def records(dtype, recs):
    r = np.zeros(len(recs), dtype=dtype)
    for i, x in enumerate(recs):
        for k, v in x.items():
            r[i][k] = v.desc.addr if isinstance(v, Mat) else v
    blk = Block(r.nbytes)
    blk.write(r)
    return blk


def compare(got, want, tol, what):
    d = np.abs(got - want)
    bad = (d > tol * (np.abs(got) + np.abs(want))) & (d > tol)
    bad |= np.isnan(got) ^ np.isnan(want)
    if bad.any():
        i = int(np.argmax(bad))
        raise AssertionError("%s: %d mismatches, first at %d emu=%r oracle=%r"
                             % (what, int(bad.sum()), i, got[i], want[i]))


def to_mats(ps, cases):
    """Two independent copies (emulator, oracle) of a tests/chol_cases.py
    case list, each matrix in its own block.  Returns two lists of records
    whose matrix fields hold Mat objects."""
    sets = []
    for _ in range(2):
        recs = []
        for mats, rec in cases:
            obj = {}
            for f, spec in mats.items():
                if isinstance(spec, str):
                    continue
                rows, cols, data = spec
                M = Mat(ps, rows, cols)
                if np.isscalar(data):
                    M.fill(np.full((rows, cols), float(data)), fill=float(data))
                else:
                    M.fill(data)
                obj[f] = M
            for f, spec in mats.items():
                if isinstance(spec, str):
                    obj[f] = obj[spec[1:]]
            recs.append(dict(rec, **obj))
        sets.append(recs)
    return sets


def compare_outputs(recs_e, recs_o, field, tol, what):
    for i, (re, ro) in enumerate(zip(recs_e, recs_o)):
        compare(re[field].content(), ro[field].content(), tol, "%s instance %d" % (what, i))
        assert re[field].use_dA() == ro[field].use_dA(), (what, i, re[field].use_dA(), ro[field].use_dA())
