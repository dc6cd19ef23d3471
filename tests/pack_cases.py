"""Shared batch builder for the column-major <-> panel-major conversion tests:
random block sizes and offsets, a column-major buffer (with leading
dimensions above the row count) and panel-major destinations larger than the
block, both filled with a sentinel outside the moved elements."""
import numpy as np

import blasfeo_amd as bfa

SENTINEL = -1.0


def make_cases(rng, batch, ps, tran=False, maxdim=200, empties=()):
    """empties: instance numbers given an empty block (m = 0 for an even
    entry of the list, n = 0 for an odd one)."""
    ms = rng.integers(1, maxdim + 1, batch)
    ns = rng.integers(1, maxdim + 1, batch)
    ms[:3] = [1, maxdim, 7]
    ns[:3] = [maxdim, 1, 5]
    for j, i in enumerate(empties):
        if j % 2 == 0:
            ms[i] = 0
        else:
            ns[i] = 0
    ais = rng.integers(0, 10, batch)
    ajs = rng.integers(0, 10, batch)
    ldas = ms + rng.integers(0, 5, batch)
    R = ns if tran else ms            # panel-major block rows x cols
    C = ms if tran else ns
    desc, offs, nbytes = bfa.bulk_layout(np.float64, ps, ais + R + 3, ajs + C + 2, 0)
    cm_off = np.concatenate([[0], np.cumsum(ldas * ns)[:-1]]).astype(np.int64)
    cm_len = int((ldas * ns).sum())
    return dict(ms=ms, ns=ns, ais=ais, ajs=ajs, ldas=ldas, R=R, C=C, pm_bytes=nbytes,
                cm_off=cm_off, cm_len=cm_len, ps=ps, batch=batch)


def records(c, cm_base, desc_base):
    rec = np.zeros(c["batch"], dtype=bfa.DPACK_ARG)
    rec["m"], rec["n"], rec["lda"] = c["ms"], c["ns"], c["ldas"]
    rec["ai"], rec["aj"] = c["ais"], c["ajs"]
    rec["A"] = np.uint64(int(cm_base)) + (8 * c["cm_off"]).astype(np.uint64)
    rec["sA"] = np.uint64(int(desc_base)) + np.arange(c["batch"], dtype=np.uint64) * bfa.DMAT.itemsize
    return rec


def descriptors(c, pm_base):
    return bfa.bulk_layout(np.float64, c["ps"], c["ais"] + c["R"] + 3, c["ajs"] + c["C"] + 2, pm_base)[0]


def pm_layout(pss, ms, ns, base):
    """blasfeo_pm_dmat descriptors (runtime panel size, panel leading
    dimension 4, as blasfeo_pm_create_dmat lays them out) for matrices of
    panel sizes pss and sizes ms x ns packed from address base, each on a
    256-byte boundary.  Returns (descriptors, byte offsets, total bytes)."""
    pss, ms, ns = (np.asarray(x, dtype=np.int64) for x in (pss, ms, ns))
    pm = -(-ms // pss) * pss
    cn = -(-ns // 4) * 4
    al = pss * 4
    tmp = -(-np.minimum(ms, ns) // al) * al
    size = (pm * cn + tmp) * 8
    padded = -(-size // 256) * 256
    offs = np.concatenate([[0], np.cumsum(padded)[:-1]]).astype(np.int64)
    d = np.zeros(len(ms), dtype=bfa.PM_DMAT)
    addr = np.uint64(int(base)) + offs.astype(np.uint64)
    d["mem"] = d["pA"] = addr
    d["dA"] = addr + (pm * cn * 8).astype(np.uint64)
    d["m"], d["n"], d["pm"], d["cn"], d["ps"], d["memsize"] = ms, ns, pm, cn, pss, size
    return d, offs, int(offs[-1] + padded[-1])
