"""Shared test machinery: build batches of panel-major instances, run them on
the GPU library and on the CPU oracle from identical bytes, and compare with
the reference's pass rule (tests/test_x_common.c:38-42: an element fails only
if |s-r| > tol*(|s|+|r|) and |s-r| > tol; tol 1e-11 FP64, 9e-4 FP32)."""
import numpy as np

import blasfeo_amd as bfa
import oracle

TOL = {"d": 1e-11, "s": 9e-4}
DT = {"d": np.float64, "s": np.float32}


def rel_mismatch(s, r, tol):
    """Boolean mask of elements failing the reference's comparison rule."""
    s = np.asarray(s, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    both_nan = np.isnan(s) & np.isnan(r)
    d = np.abs(s - r)
    bad = (d > tol * (np.abs(s) + np.abs(r))) & (d > tol)
    bad |= np.isnan(s) ^ np.isnan(r)
    return bad & ~both_nan


def assert_close(s, r, tol, what=""):
    bad = rel_mismatch(s, r, tol)
    if bad.any():
        idx = np.argwhere(bad)[:5]
        sv = np.asarray(s)[tuple(idx.T)]
        rv = np.asarray(r)[tuple(idx.T)]
        raise AssertionError("%s: %d mismatches, first at %s gpu=%s ref=%s"
                             % (what, bad.sum(), idx.tolist(), sv, rv))


class Batch:
    """Matrices of one precision/panel size; the same bytes go to the GPU and
    to the oracle."""

    def __init__(self, prec, ps, seed=0):
        self.prec = prec
        self.ps = ps
        self.layout = bfa.Layout(DT[prec], ps)
        self.fill = []
        self.rng = np.random.default_rng(seed)

    def mat(self, m, n, data="randn", value=None):
        i = self.layout.add(m, n)
        self.fill.append((data, value))
        return i

    def host(self):
        buf = self.layout.host_buffer()
        for i, (data, value) in enumerate(self.fill):
            m, n = self.layout.shapes[i]
            if data == "randn":
                A = self.rng.standard_normal((m, n))
            elif data == "uniform":
                A = self.rng.uniform(-1, 1, (m, n))
            elif data == "const":
                A = np.full((m, n), value)
            else:
                A = np.asarray(value)
            self.layout.put(buf, i, A)
        return buf


def run_both(batch, buf, argdtype, records, gpu_name, oracle_call, hint):
    """records: list of dicts whose matrix fields (sA, sB, ...) hold matrix
    indices.  Runs the GPU batched entry and the oracle; returns
    (gpu_buf, gpu_desc, ref_buf, ref_desc)."""
    from blasfeo_amd.device import DeviceBatch, HostBatch, to_device
    ptr_fields = [f for f in argdtype.names if f.startswith("s") and argdtype[f] == np.dtype("u8")]
    dev = DeviceBatch(batch.layout, buf)
    host = HostBatch(batch.layout, buf)
    recs = np.zeros(len(records), dtype=argdtype)
    for i, r in enumerate(records):
        for f, v in r.items():
            if f.startswith("_"):
                continue  # test-side annotation, not part of the record
            recs[i][f] = dev.desc_addr(v) if f in ptr_fields else v
    drecs = to_device(recs)
    bfa.call(gpu_name, batch.ps, len(records), drecs.data_ptr(), hint)
    import torch
    torch.cuda.synchronize()
    for r in records:
        oracle_call(host, r)
    return dev.download(), dev.download_desc(), host.buf, host.desc


def compare_buffers(batch, gbuf, rbuf, what, tol=None, mats=None):
    tol = TOL[batch.prec] if tol is None else tol
    lay = batch.layout
    for i in (range(len(lay)) if mats is None else mats):
        assert_close(lay.view(gbuf, i), lay.view(rbuf, i), tol, "%s matrix %d" % (what, i))
        assert_close(lay.dview(gbuf, i), lay.dview(rbuf, i), tol, "%s matrix %d dA" % (what, i))
