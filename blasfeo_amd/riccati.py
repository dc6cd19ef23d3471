"""Batches of backward Riccati factorisations (the chain of
examples/example_d_riccati_recursion.c:115-140) laid out for the fused GPU
sweep ``blasfeo_amd_d_back_ric_trf_batched``.

Instance i has horizon N, state size nx_i and input size nu_i on every stage
(nu = 0 on the last stage, as in the reference example).  Per instance the
time-invariant data are stored once and every stage's descriptor points at
them, as the reference example does (:737-747):
    BAbt = [B^T; A^T]   (nu+nx) x nx
    RSQ  = [R S; S^T Q] (nu+nx) x (nu+nx)      (lower triangle used)
    QN   = Q            nx x nx                 (last stage)
Outputs: L[0..N-1] (nu+nx) x (nu+nx), L[N] nx x nx, plus one work matrix.

``RiccatiSolve`` adds the vectors of the solve on those factors
(d_back_ric_trs_libstr, :203-267) for ``blasfeo_amd_d_back_ric_trs_batched``.
"""
import numpy as np

from . import DMAT, DVEC, Layout, RICCATI_ARG, RICCATI_TRS_ARG, RICCATI_SV_ARG


class RiccatiBatch:
    """aug: every matrix has one more row, as d_back_ric_sv_libstr takes them
    (examples/example_d_riccati_recursion.c:58-111): BAbt's last row holds
    b^T, RSQ's and QN's the linear term [r; q]^T, and the factors are
    (nu+nx+1) x (nu+nx) trapezoids.  The factorisation alone ignores the
    extra row, so one augmented batch serves both forms."""

    def __init__(self, nxs, nus, N, ps=4, aug=False):
        self.nxs = np.asarray(nxs, dtype=np.int64)
        self.nus = np.asarray(nus, dtype=np.int64)
        self.N = int(N)
        self.ps = ps
        self.aug = bool(aug)
        a = 1 if aug else 0
        self.batch = len(self.nxs)
        self.inp = Layout(np.float64, ps)
        self.out = Layout(np.float64, ps)
        self.i_babt, self.i_rsq, self.i_qn = [], [], []
        self.i_L, self.i_work = [], []
        for nx, nu in zip(self.nxs, self.nus):
            nx, nu = int(nx), int(nu)
            self.i_babt.append(self.inp.add(nu + nx + a, nx))
            self.i_rsq.append(self.inp.add(nu + nx + a, nu + nx))
            self.i_qn.append(self.inp.add(nx + a, nx))
            Ls = [self.out.add(nu + nx + a, nu + nx) for _ in range(self.N)]
            Ls.append(self.out.add(nx + a, nx))
            self.i_L.append(Ls)
            self.i_work.append(self.out.add(nu + nx + a, nx))
        self.max_nk = int((self.nxs + self.nus).max()) if self.batch else 0
        self.max_nx = int(self.nxs.max()) if self.batch else 0

    # -- data ---------------------------------------------------------------
    def fill_inputs(self, rng, A_scale=0.1, R=None, Q=None):
        """A = I + A_scale*N(0,1), B ~ N(0,1), R = Q = I, S = 0 (host buffer);
        with aug, b, r and q ~ N(0,1) in the extra rows."""
        buf = self.inp.host_buffer()
        for i in range(self.batch):
            nx, nu = int(self.nxs[i]), int(self.nus[i])
            A = np.eye(nx) + A_scale * rng.standard_normal((nx, nx))
            B = rng.standard_normal((nx, nu))
            babt, rsq, qn = np.vstack([B.T, A.T]), np.eye(nu + nx), np.eye(nx)
            if self.aug:
                babt = np.vstack([babt, rng.standard_normal((1, nx))])
                rsq = np.vstack([rsq, rng.standard_normal((1, nu + nx))])
                qn = np.vstack([qn, rng.standard_normal((1, nx))])
            self.inp.put(buf, self.i_babt[i], babt)
            self.inp.put(buf, self.i_rsq[i], rsq)
            self.inp.put(buf, self.i_qn[i], qn)
        return buf

    def fill_custom(self, fn):
        """fn(i, nx, nu) -> (BAbt, RSQ, QN) dense arrays."""
        buf = self.inp.host_buffer()
        for i in range(self.batch):
            nx, nu = int(self.nxs[i]), int(self.nus[i])
            babt, rsq, qn = fn(i, nx, nu)
            self.inp.put(buf, self.i_babt[i], babt)
            self.inp.put(buf, self.i_rsq[i], rsq)
            self.inp.put(buf, self.i_qn[i], qn)
        return buf

    # -- descriptor arrays ----------------------------------------------------
    def tables(self, in_base, out_base):
        """Per-stage descriptor tables and nx/nu arrays for buffers at the given
        addresses: returns (babt[batch*N], rsq[batch*(N+1)], L[batch*(N+1)],
        work[batch], nx[batch*(N+1)], nu[batch*(N+1)])."""
        N = self.N
        din = self.inp.build(in_base)
        dout = self.out.build(out_base)
        babt = np.zeros(self.batch * max(N, 1), dtype=DMAT)
        rsq = np.zeros(self.batch * (N + 1), dtype=DMAT)
        L = np.zeros(self.batch * (N + 1), dtype=DMAT)
        work = np.zeros(self.batch, dtype=DMAT)
        nx = np.zeros(self.batch * (N + 1), dtype=np.int32)
        nu = np.zeros(self.batch * (N + 1), dtype=np.int32)
        for i in range(self.batch):
            for k in range(N):
                babt[i * N + k] = din[self.i_babt[i]]
                rsq[i * (N + 1) + k] = din[self.i_rsq[i]]
            rsq[i * (N + 1) + N] = din[self.i_qn[i]]
            for k in range(N + 1):
                L[i * (N + 1) + k] = dout[self.i_L[i][k]]
            work[i] = dout[self.i_work[i]]
            nx[i * (N + 1): (i + 1) * (N + 1)] = self.nxs[i]
            nu[i * (N + 1): (i + 1) * (N + 1) - 1] = self.nus[i]
        return babt, rsq, L, work, nx, nu

    def args(self, addr_babt, addr_rsq, addr_L, addr_work, addr_nx, addr_nu):
        N = self.N
        a = np.zeros(self.batch, dtype=RICCATI_ARG)
        idx = np.arange(self.batch, dtype=np.uint64)
        a["N"] = N
        a["nx"] = addr_nx + idx * 4 * (N + 1)
        a["nu"] = addr_nu + idx * 4 * (N + 1)
        a["BAbt"] = addr_babt + idx * DMAT.itemsize * max(N, 1)
        a["RSQrq"] = addr_rsq + idx * DMAT.itemsize * (N + 1)
        a["L"] = addr_L + idx * DMAT.itemsize * (N + 1)
        a["work"] = addr_work + idx * DMAT.itemsize
        return a


class RiccatiSolve:
    """The vectors of the Riccati solve for every instance of a RiccatiBatch,
    packed into one byte buffer (each vector on a 64-byte boundary, its size
    rounded as blasfeo_memsize_dvec does, auxiliary/d_aux_lib4.c).

    Per instance the descriptors are stored in this order:
        b[0..N-1] (nx), rq[0..N] (nu+nx; nx on stage N), Pb[0..N-1] (nx),
        ux[0..N] (nu+nx; nx on stage N), pi[0..N-1] (nx), work (nx)
    so that each kind is a contiguous descriptor array, as the solve's
    argument record expects.  b and rq are inputs, the rest outputs.
    """

    KINDS = ("b", "rq", "Pb", "ux", "pi", "work")

    def __init__(self, rb):
        self.rb = rb
        N, ps = rb.N, rb.ps
        self.per = 5 * N + 3
        self.start = {"b": 0, "rq": N, "Pb": 2 * N + 1, "ux": 3 * N + 1, "pi": 4 * N + 2,
                      "work": 5 * N + 2}
        nx = rb.nxs[:, None]
        nk = (rb.nxs + rb.nus)[:, None]
        sz = np.zeros((rb.batch, self.per), dtype=np.int64)
        st = self.start
        sz[:, st["b"]:st["b"] + N] = nx
        sz[:, st["rq"]:st["rq"] + N] = nk
        sz[:, st["rq"] + N] = rb.nxs
        sz[:, st["Pb"]:st["Pb"] + N] = nx
        sz[:, st["ux"]:st["ux"] + N] = nk
        sz[:, st["ux"] + N] = rb.nxs
        sz[:, st["pi"]:st["pi"] + N] = nx
        sz[:, st["work"]] = rb.nxs
        self.sizes = sz.reshape(-1)
        self.pm = -(-self.sizes // ps) * ps
        self.memsize = -(-(self.pm * 8) // 64) * 64
        self.offs = np.concatenate([[0], np.cumsum(self.memsize)[:-1]]).astype(np.int64)
        self.nbytes = int(self.offs[-1] + self.memsize[-1]) if len(self.offs) else 0

    def index(self, i, kind, k=0):
        return i * self.per + self.start[kind] + k

    def table(self, base):
        d = np.zeros(len(self.sizes), dtype=DVEC)
        addr = np.uint64(int(base)) + self.offs.astype(np.uint64)
        d["mem"] = addr
        d["pa"] = addr
        d["m"] = self.sizes
        d["pm"] = self.pm
        d["memsize"] = self.memsize
        return d

    def fill_inputs(self, rng):
        """b ~ N(0,1) and rq ~ N(0,1) on every stage (host buffer)."""
        buf = np.zeros(self.nbytes, dtype=np.uint8)
        f = buf.view(np.float64)
        sel = np.zeros((self.rb.batch, self.per), dtype=bool)
        N = self.rb.N
        sel[:, self.start["b"]:self.start["b"] + N] = True
        sel[:, self.start["rq"]:self.start["rq"] + N + 1] = True
        sel = sel.reshape(-1)
        m = self.sizes[sel]
        first = self.offs[sel] // 8
        idx = np.repeat(first - np.concatenate([[0], np.cumsum(m)[:-1]]), m) + np.arange(int(m.sum()))
        f[idx] = rng.standard_normal(len(idx))
        return buf

    def vec(self, buf, i, kind, k=0):
        j = self.index(i, kind, k)
        o = int(self.offs[j])
        return np.asarray(buf)[o: o + int(self.sizes[j]) * 8].view(np.float64).copy()

    def args(self, tab, addr_babt, addr_L, addr_nx, addr_nu):
        """Argument records for descriptor table address ``tab`` and the
        factorisation's BAbt / L / nx / nu arrays (RiccatiBatch.tables)."""
        rb, N = self.rb, self.rb.N
        a = np.zeros(rb.batch, dtype=RICCATI_TRS_ARG)
        idx = np.arange(rb.batch, dtype=np.uint64)
        a["N"] = N
        a["nx"] = addr_nx + idx * 4 * (N + 1)
        a["nu"] = addr_nu + idx * 4 * (N + 1)
        a["BAbt"] = addr_babt + idx * DMAT.itemsize * max(N, 1)
        a["L"] = addr_L + idx * DMAT.itemsize * (N + 1)
        base = np.uint64(int(tab)) + idx * np.uint64(DVEC.itemsize * self.per)
        for kind in self.KINDS:
            a[kind] = base + np.uint64(DVEC.itemsize * self.start[kind])
        return a


def trs_backward_error(rb, inp, rs, sin, sol, i):
    """Normwise backward error of instance i's Riccati solve ``sol`` (a
    RiccatiSolve buffer; ``sin`` holds its b and rq, ``inp`` the RiccatiBatch
    inputs), in units of eps: the
    largest over stages of the residual of the stage's KKT equations
        H_k z_k + rq_k + BAbt_k pi_k - E_k^T pi_{k-1} = 0,   BAbt_k^T z_k + b_k - x_{k+1} = 0
    (z_k = ux_k, x_{k+1} the state part of ux_{k+1}, E_k z_k = x_k) over
    ||H_k|| ||z_k|| + ||rq_k|| + ||BAbt_k|| ||pi_k|| + ||pi_{k-1}|| and
    ||BAbt_k|| ||z_k|| + ||b_k|| + ||x_{k+1}|| (infinity norms)."""
    eps = np.finfo(np.float64).eps
    N, nx, nu = rb.N, int(rb.nxs[i]), int(rb.nus[i])
    # square parts (an augmented batch has the linear terms in one more row)
    C = rb.inp.get(inp, rb.i_babt[i])[:nu + nx]
    H = np.tril(rb.inp.get(inp, rb.i_rsq[i])[:nu + nx])
    H = H + np.tril(H, -1).T
    HN = np.tril(rb.inp.get(inp, rb.i_qn[i])[:nx])
    HN = HN + np.tril(HN, -1).T
    nrm = lambda v: float(np.abs(v).max()) if np.size(v) else 0.0
    nC, nH, nHN = np.abs(C).sum(1).max(), np.abs(H).sum(1).max(), np.abs(HN).sum(1).max()
    z = [rs.vec(sol, i, "ux", k) for k in range(N + 1)]
    pi = [rs.vec(sol, i, "pi", k) for k in range(N)]
    worst = 0.0
    for k in range(N + 1):
        g = rs.vec(sin, i, "rq", k)
        Hk, nHk = (H, nH) if k < N else (HN, nHN)
        r = Hk @ z[k] + g
        den = nHk * nrm(z[k]) + nrm(g)
        if k < N:
            r = r + C @ pi[k]
            den += nC * nrm(pi[k])
        if k > 0:
            r[len(r) - nx:] -= pi[k - 1]
            den += nrm(pi[k - 1])
        worst = max(worst, nrm(r) / (den * eps))
        if k < N:
            b = rs.vec(sin, i, "b", k)
            x1 = z[k + 1][len(z[k + 1]) - nx:]
            r = C.T @ z[k] + b - x1
            den = np.abs(C).sum(0).max() * nrm(z[k]) + nrm(b) + nrm(x1)
            worst = max(worst, nrm(r) / (den * eps))
    return worst


class RiccatiSv:
    """The output vectors ux[0..N] and pi[0..N-1] of the fused sweep
    (d_back_ric_sv_libstr) for every instance of an augmented RiccatiBatch,
    in the vector layout of a RiccatiSolve (its ux, pi and work entries)."""

    def __init__(self, rb):
        assert rb.aug, "the fused sweep takes an augmented RiccatiBatch"
        self.rb = rb
        self.rs = RiccatiSolve(rb)

    def args(self, tab, addr_babt, addr_rsq, addr_L, addr_work, addr_nx, addr_nu):
        rb, rs, N = self.rb, self.rs, self.rb.N
        a = np.zeros(rb.batch, dtype=RICCATI_SV_ARG)
        idx = np.arange(rb.batch, dtype=np.uint64)
        a["N"] = N
        a["nx"] = addr_nx + idx * 4 * (N + 1)
        a["nu"] = addr_nu + idx * 4 * (N + 1)
        a["BAbt"] = addr_babt + idx * DMAT.itemsize * max(N, 1)
        a["RSQrq"] = addr_rsq + idx * DMAT.itemsize * (N + 1)
        a["L"] = addr_L + idx * DMAT.itemsize * (N + 1)
        a["work_mat"] = addr_work + idx * DMAT.itemsize
        base = np.uint64(int(tab)) + idx * np.uint64(DVEC.itemsize * rs.per)
        a["ux"] = base + np.uint64(DVEC.itemsize * rs.start["ux"])
        a["pi"] = base + np.uint64(DVEC.itemsize * rs.start["pi"])
        a["work_vec"] = base + np.uint64(DVEC.itemsize * rs.start["work"])
        return a

    def solve_inputs(self, inp):
        """A RiccatiSolve input buffer whose b[k] and rq[k] are the extra rows
        of the augmented inputs, for d_back_ric_trs_libstr on the same problem."""
        rb, rs = self.rb, self.rs
        buf = np.zeros(rs.nbytes, dtype=np.uint8)
        f = buf.view(np.float64)
        for i in range(rb.batch):
            nx, nk = int(rb.nxs[i]), int(rb.nxs[i] + rb.nus[i])
            b = rb.inp.get(inp, rb.i_babt[i])[nk]
            rq = rb.inp.get(inp, rb.i_rsq[i])[nk]
            q = rb.inp.get(inp, rb.i_qn[i])[nx]
            for k in range(rb.N + 1):
                if k < rb.N:
                    o = int(rs.offs[rs.index(i, "b", k)]) // 8
                    f[o: o + nx] = b
                o = int(rs.offs[rs.index(i, "rq", k)]) // 8
                v = rq if k < rb.N else q
                f[o: o + len(v)] = v
        return buf


class DeviceRiccati:
    """A RiccatiBatch resident in GPU memory, ready to launch."""

    def __init__(self, rb, in_host_buf):
        import torch
        from .device import to_device
        self.rb = rb
        self.inbuf = torch.from_numpy(np.ascontiguousarray(in_host_buf)).to("cuda")
        self.outbuf = torch.zeros(rb.out.nbytes, dtype=torch.uint8, device="cuda")
        t = rb.tables(self.inbuf.data_ptr(), self.outbuf.data_ptr())
        self.t_dev = [to_device(x) for x in t]
        babt, rsq, L, work, nx, nu = self.t_dev
        self.args = to_device(rb.args(babt.data_ptr(), rsq.data_ptr(), L.data_ptr(),
                                      work.data_ptr(), nx.data_ptr(), nu.data_ptr()))

    def run(self, stream=0):
        from . import call
        call("d_back_ric_trf", self.rb.ps, self.rb.batch, self.args.data_ptr(),
             (self.rb.max_nk, self.rb.max_nx, 0), stream)

    def capture(self, stream):
        """The batched sweep (plan and class launches) captured into a CUDA/HIP
        graph on ``stream``.  Run it once eagerly on the same stream first: the
        library allocates its per-stream workspace on first use, which a
        capture cannot record."""
        import torch
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            self.run(stream.cuda_stream)
        return g

    def L(self, i, k):
        buf = self.outbuf.cpu().numpy()
        return self.rb.out.get(buf, self.rb.i_L[i][k])


class DeviceRiccatiSolve:
    """The solve vectors of a RiccatiSolve in GPU memory, on the factors of a
    DeviceRiccati (whose sweep must have run first on the same stream)."""

    def __init__(self, dev, rs, host_buf):
        import torch
        from .device import to_device
        self.dev, self.rs = dev, rs
        self.buf = torch.from_numpy(np.ascontiguousarray(host_buf)).to("cuda")
        self.tab = to_device(rs.table(self.buf.data_ptr()))
        babt, rsq, L, work, nx, nu = dev.t_dev
        self.args = to_device(rs.args(self.tab.data_ptr(), babt.data_ptr(), L.data_ptr(),
                                      nx.data_ptr(), nu.data_ptr()))

    def run(self, stream=0):
        from . import call
        rb = self.rs.rb
        call("d_back_ric_trs", rb.ps, rb.batch, self.args.data_ptr(),
             (rb.max_nk, rb.max_nx, 0), stream)

    def host(self):
        return self.buf.cpu().numpy()


class DeviceRiccatiSv:
    """An augmented RiccatiBatch and its RiccatiSv vectors resident in GPU
    memory, for ``blasfeo_amd_d_back_ric_sv_batched``."""

    def __init__(self, rb, in_host_buf):
        import torch
        from .device import to_device
        self.rb, self.sv = rb, RiccatiSv(rb)
        self.fac = DeviceRiccati(rb, in_host_buf)
        self.vbuf = torch.zeros(self.sv.rs.nbytes, dtype=torch.uint8, device="cuda")
        self.tab = to_device(self.sv.rs.table(self.vbuf.data_ptr()))
        babt, rsq, L, work, nx, nu = self.fac.t_dev
        self.args = to_device(self.sv.args(self.tab.data_ptr(), babt.data_ptr(), rsq.data_ptr(), L.data_ptr(),
                                           work.data_ptr(), nx.data_ptr(), nu.data_ptr()))

    def run(self, stream=0):
        from . import call
        call("d_back_ric_sv", self.rb.ps, self.rb.batch, self.args.data_ptr(),
             (self.rb.max_nk, self.rb.max_nx, 0), stream)

    def vectors(self):
        return self.vbuf.cpu().numpy()


class DeviceRiccatiMixed:
    """A RiccatiBatch, its RiccatiSolve vectors (b, rq in; Pb, ux, pi out) and
    a second vector set for the corrections, resident in GPU memory, for
    ``blasfeo_amd_d_ric_solve_mixed_batched`` (FP32 factorisation, FP64
    refinement, FP64 fallback)."""

    def __init__(self, rb, in_host_buf, rs, sol_host_buf):
        import torch
        from . import RICCATI_MIXED_ARG, SOLVE_INFO
        from .device import to_device
        self.rb, self.rs = rb, rs
        self.fac = DeviceRiccati(rb, in_host_buf)
        self.vbuf = torch.from_numpy(np.ascontiguousarray(sol_host_buf)).to("cuda")
        self.cbuf = torch.zeros(rs.nbytes, dtype=torch.uint8, device="cuda")
        self.tab = to_device(rs.table(self.vbuf.data_ptr()))
        self.ctab = to_device(rs.table(self.cbuf.data_ptr()))
        self.info = torch.zeros(rb.batch * SOLVE_INFO.itemsize, dtype=torch.uint8, device="cuda")
        babt, rsq, L, work, nx, nu = self.fac.t_dev
        N = rb.N
        a = np.zeros(rb.batch, dtype=RICCATI_MIXED_ARG)
        idx = np.arange(rb.batch, dtype=np.uint64)
        a["N"] = N
        a["nx"] = nx.data_ptr() + idx * 4 * (N + 1)
        a["nu"] = nu.data_ptr() + idx * 4 * (N + 1)
        a["BAbt"] = babt.data_ptr() + idx * DMAT.itemsize * max(N, 1)
        a["RSQrq"] = rsq.data_ptr() + idx * DMAT.itemsize * (N + 1)
        a["L"] = L.data_ptr() + idx * DMAT.itemsize * (N + 1)
        a["work_mat"] = work.data_ptr() + idx * DMAT.itemsize
        base = np.uint64(self.tab.data_ptr()) + idx * np.uint64(DVEC.itemsize * rs.per)
        cbase = np.uint64(self.ctab.data_ptr()) + idx * np.uint64(DVEC.itemsize * rs.per)
        for f, kind in (("b", "b"), ("rq", "rq"), ("Pb", "Pb"), ("ux", "ux"), ("pi", "pi"), ("work_vec", "work")):
            a[f] = base + np.uint64(DVEC.itemsize * rs.start[kind])
        for f, kind in (("res_b", "b"), ("res_rq", "rq"), ("d_Pb", "Pb"), ("d_ux", "ux"), ("d_pi", "pi")):
            a[f] = cbase + np.uint64(DVEC.itemsize * rs.start[kind])
        a["info"] = np.uint64(self.info.data_ptr()) + idx * np.uint64(SOLVE_INFO.itemsize)
        self.args = to_device(a)

    def run(self, tol=1e-14, max_iter=3, stream=0):
        from . import ric_solve_mixed
        rb = self.rb
        ric_solve_mixed(rb.ps, rb.batch, self.args.data_ptr(), tol, max_iter, (rb.max_nk, rb.max_nx, 0), stream)

    def vectors(self):
        return self.vbuf.cpu().numpy()

    def infos(self):
        from . import SOLVE_INFO
        return self.info.cpu().numpy().view(SOLVE_INFO)


class HostRiccatiSv:
    """The same augmented batch and vectors in host memory for the CPU oracle
    (``or_dric_sv_batch``)."""

    def __init__(self, rb, in_host_buf):
        self.rb, self.sv = rb, RiccatiSv(rb)
        self.fac = HostRiccati(rb, in_host_buf)
        self.vbuf = np.zeros(self.sv.rs.nbytes, dtype=np.uint8)
        self.tab = self.sv.rs.table(self.vbuf.ctypes.data)
        babt, rsq, L, work, nx, nu = self.fac.t
        self.argv = self.sv.args(self.tab.ctypes.data, babt.ctypes.data, rsq.ctypes.data, L.ctypes.data,
                                 work.ctypes.data, nx.ctypes.data, nu.ctypes.data)


class HostRiccati:
    """The same batch in host memory for the CPU oracle."""

    def __init__(self, rb, in_host_buf):
        self.rb = rb
        self.inbuf = np.array(in_host_buf, copy=True)
        self.outbuf = rb.out.host_buffer()
        t = rb.tables(self.inbuf.ctypes.data, self.outbuf.ctypes.data)
        self.t = t
        babt, rsq, L, work, nx, nu = t
        self.argv = rb.args(babt.ctypes.data, rsq.ctypes.data, L.ctypes.data,
                            work.ctypes.data, nx.ctypes.data, nu.ctypes.data)

    def L(self, i, k):
        return self.rb.out.get(self.outbuf, self.rb.i_L[i][k])


def stage_backward_errors(rb, inp, gbuf, instances):
    """Per-stage backward error of computed Riccati factors, in units of
    n_k eps: ||M_k - L_k L_k^T||_F / (||M_k||_F n_k eps) with
    M_N = Q_N and M_k = RSQ + W W^T, W = BAbt L_{k+1}(nu:, nu:), rebuilt from
    the computed L_{k+1} itself, so that each stage is judged on its own
    inputs (the recursion amplifies forward differences over the horizon).
    gbuf: the output buffer (host bytes) of a RiccatiBatch, inp its input
    buffer.  Returns an array (len(instances), N + 1); column k is stage k."""
    eps = np.finfo(np.float64).eps
    out = []
    for i in instances:
        nx, nu = int(rb.nxs[i]), int(rb.nus[i])
        babt = rb.inp.get(inp, rb.i_babt[i])[:nu + nx]
        rsq = np.tril(rb.inp.get(inp, rb.i_rsq[i])[:nu + nx])
        rsq = rsq + np.tril(rsq, -1).T
        qn = np.tril(rb.inp.get(inp, rb.i_qn[i])[:nx])
        qn = qn + np.tril(qn, -1).T
        L = [np.tril(rb.out.get(gbuf, rb.i_L[i][k]))[:nu + nx if k < rb.N else nx] for k in range(rb.N + 1)]
        row = np.empty(rb.N + 1)
        row[rb.N] = np.linalg.norm(qn - L[rb.N] @ L[rb.N].T) / (np.linalg.norm(qn) * max(nx, 1) * eps)
        for k in range(rb.N):
            nu1 = nu if k + 1 < rb.N else 0
            W = babt @ L[k + 1][nu1:, nu1:]
            M = rsq + W @ W.T
            row[k] = np.linalg.norm(M - L[k] @ L[k].T) / (np.linalg.norm(M) * (nx + nu) * eps)
        out.append(row)
    return np.array(out)
