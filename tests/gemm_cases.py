"""Record sets of the gemm-family tests shared by the GPU tests and the CPU
kernel emulator (tests/emu/); same case format as tests/chol_cases.py."""
import numpy as np


def uneven(ps, op, seed, count=24, smax=140):
    """gemm_nt / gemm_nn / syrk_ln / syrk_ln_mn with m, n, k ~ U{1..smax-1}
    (a few edge shapes first: 1, 4, 31, 32, 33 and k = 0), offsets up to
    2 ps on every operand, alpha and beta in {1, -0.5, 2, 0}; D in place on C
    for every fifth instance."""
    rng = np.random.default_rng(seed)
    out = []
    edge = [(1, 1, 1), (4, 4, 4), (31, 33, 5), (32, 32, 32), (33, 1, 0), (64, 65, 3)]
    for r in range(count):
        if r < len(edge):
            m, n, k = edge[r]
        else:
            m, n, k = (int(x) for x in rng.integers(1, smax, 3))
        if op == "syrk_ln":
            n = m
        elif op == "syrk_ln_mn":
            n = min(n, m)
        ai, bi, ci, di = (int(x) for x in rng.integers(0, 2 * ps, 4))
        aj, bj, cj, dj = (int(x) for x in rng.integers(0, 3, 4))
        alpha = float(rng.choice([1.0, -0.5, 2.0]))
        beta = float(rng.choice([1.0, -0.5, 0.0]))
        kk = max(k, 1)
        A = rng.standard_normal((ai + m, aj + kk))
        B = rng.standard_normal((bi + kk, bj + n)) if op == "gemm_nn" else rng.standard_normal((bi + n, bj + kk))
        C = rng.standard_normal((ci + m, cj + n))
        mats = {"sA": (A.shape[0], A.shape[1], A), "sB": (B.shape[0], B.shape[1], B), "sC": (C.shape[0], C.shape[1], C)}
        if r % 5 == 4:
            mats["sD"] = "=sC"
            di, dj = ci, cj
        else:
            mats["sD"] = (di + m, dj + n, -1.0)
        out.append((mats, dict(m=m, n=n, k=k, ai=ai, aj=aj, bi=bi, bj=bj, ci=ci, cj=cj, di=di, dj=dj,
                               alpha=alpha, beta=beta)))
    return out
