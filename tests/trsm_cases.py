"""Record sets of the left-solve tests shared by the GPU tests
(tests/test_gpu_lu_solve.py) and the CPU kernel emulator (tests/emu/); same
case format as tests/chol_cases.py."""
import numpy as np


def nonfinite(ps, op, seed):
    """trsm_llnn / lltn / llnu / lunn with a zero on the diagonal of the
    triangle (its reciprocal is Inf) in some instances and a NaN in B in
    others, m up to 256 and n <= 4, so that both the one-wavefront and the
    four-wave solve meet Inf and NaN in the rows they broadcast; the rows
    solved before them must keep their values, as in the reference."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(10):
        m = [256, 70, 5, 130, 64, 200, 1, 65, 129, 33][r]
        n = 1 + r % 4
        ai, bi, di = (int(x) for x in rng.integers(0, 2 * ps, 3))
        aj, bj, dj = (int(x) for x in rng.integers(0, 3, 3))
        if r % 3 == 0:
            ai = aj = 0
        blk = rng.standard_normal((m, m)) * 0.1
        diag = 1.0 + rng.uniform(0.5, 1.5, m)
        upper = op == "trsm_lunn"
        T = (np.triu(blk, 1) if upper else np.tril(blk, -1)) + np.diag(diag)
        z = int(rng.integers(0, m))
        if r % 2 == 0 and op != "trsm_llnu":
            T[z, z] = 0.0
        Ad = np.full((ai + m, aj + m), 1e3)
        Ad[ai:, aj:] = T + 1e3 * ((np.tril(np.ones((m, m)), -1) if upper else np.triu(np.ones((m, m)), 1)))
        Bd = rng.standard_normal((bi + m, bj + n))
        if r % 2 == 1:
            Bd[bi + int(rng.integers(0, m)), bj + int(rng.integers(0, n))] = np.nan
        mats = {"sA": (ai + m, aj + m, Ad), "sB": (bi + m, bj + n, Bd), "sD": (di + m, dj + n, -1.0)}
        out.append((mats, dict(m=m, n=n, ai=ai, aj=aj, bi=bi, bj=bj, di=di, dj=dj, alpha=1.0)))
    return out
