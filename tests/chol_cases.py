"""Record sets of the Cholesky-family tests, shared by the GPU tests
(tests/test_gpu_lapack.py) and the CPU kernel emulator (tests/emu/), so that
both run exactly the same instances.

Each case is (matrices, record): matrices maps a record field (sA, sC, sD,
...) to (rows, cols, data), data a dense array or a float fill value; a
field may instead name another field ("=sC") to alias that matrix."""
import numpy as np


def _spd(rng, n):
    B = rng.uniform(-1, 1, (n, n))
    return B @ B.T + n * np.eye(n)


def potrf(ps, op, sizes, seed, inplace_at=None):
    """potrf_l (m = n) or potrf_l_mn (m = n + U{0..19}) with row/column offsets
    on C and D; in place (D = C) for the instance number inplace_at."""
    rng = np.random.default_rng(seed)
    out = []
    for r, n in enumerate(sizes):
        m = n if op == "potrf_l" else n + int(rng.integers(0, 20))
        ci, cj = (0, 0) if r % 2 == 0 else (int(rng.integers(0, 2 * ps)), int(rng.integers(0, 3)))
        di, dj = (0, 0) if r % 3 == 0 else (int(rng.integers(0, 2 * ps)), int(rng.integers(0, 3)))
        S = np.zeros((ci + m, cj + n))
        S[ci:, cj:] = _spd(rng, m)[:, :n]
        mats = {"sC": (ci + m, cj + n, S)}
        if r == inplace_at:
            mats["sD"] = "=sC"
            di, dj = ci, cj
        else:
            mats["sD"] = (di + m, dj + n, -1.0)
        out.append((mats, dict(m=m, n=n, ci=ci, cj=cj, di=di, dj=dj, ipiv=0)))
    return out


POTRF_SIZES = [1, 2, 3, 15, 16, 17, 31, 32, 33, 63, 64, 65, 100, 128, 129, 200, 256]
POTRF_LARGE_SIZES = [257, 300, 3, 400, 129, 333]


def syrk_potrf(ps, seed, count=14, mmax=300):
    """dsyrk_dpotrf_ln(_mn): m up to mmax - 1 (the first two instances 129 and
    mmax - 1), n <= min(m, 256), k = 0 .. 69, A = B at row offset ai, D at
    row offset di on every fourth instance."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(count):
        m = int(rng.integers(1, mmax)) if r > 1 else (min(129, mmax - 1) if r == 0 else mmax - 1)
        n = min(256, m if r % 3 else int(rng.integers(1, m + 1)))
        k = int(rng.integers(0, 70))
        ai, _ = (int(x) for x in rng.integers(0, 2 * ps, 2))
        di = int(rng.integers(0, 2 * ps)) if r % 4 == 1 else 0
        mats = {"sA": (ai + m, k + 1, rng.standard_normal((ai + m, k + 1))),
                "sB": "=sA",
                "sC": (m, n, m * np.eye(m)[:, :n] + 0.01 * rng.standard_normal((m, n))),
                "sD": (di + m, n, -1.0)}
        out.append((mats, dict(m=m, n=n, k=k, ai=ai, aj=1, bi=ai, bj=1, ci=0, cj=0, di=di, dj=0)))
    return out


def to_batch(b, cases):
    """Adds the cases' matrices to a tests/helpers.py Batch; returns records
    whose matrix fields hold matrix indices."""
    recs = []
    for mats, rec in cases:
        idx = {}
        for f, spec in mats.items():
            if isinstance(spec, str):
                continue
            rows, cols, data = spec
            if np.isscalar(data):
                idx[f] = b.mat(rows, cols, "const", float(data))
            else:
                idx[f] = b.mat(rows, cols, "data", data)
        for f, spec in mats.items():
            if isinstance(spec, str):
                idx[f] = idx[spec[1:]]
        recs.append(dict(rec, **idx))
    return recs
