// bfa_ric_trs.hip -- batched Riccati solve (d_back_ric_trs_libstr,
// examples/example_d_riccati_recursion.c:203-267) on the factors L[0..N] of
// the backward factorisation sweep.
//
// One wavefront per instance runs both sweeps in one launch.  With
// Lxx(k) = L[k](nu[k]:, nu[k]:) and nk = nu[k] + nx[k]:
//   backward, k = N-1 .. 0 (ux[N] = rq[N]):
//     Pb[k] = Lxx(k+1) Lxx(k+1)^T b[k]                       (trmv_ltn, trmv_lnn)
//     ux[k] = rq[k] + BAbt[k] (Pb[k] + ux[k+1](nu[k+1]:))     (gemv_n)
//     ux[k] = trsv_lnn_mn(nk, ns) of L[k] on ux[k]
//   forward, k = 0 .. N-1:
//     pi[k] = ux[k+1](nu[k+1]:)                               (the backward value)
//     ux[k](0:ns) = -ux[k](0:ns); ux[k] = trsv_ltn_mn(nk, ns) of L[k]
//     ux[k+1](nu[k+1]:) = b[k] + BAbt[k]^T ux[k]              (gemv_t)
//     pi[k] += Lxx(k+1) Lxx(k+1)^T ux[k+1](nu[k+1]:)
// where ns = nu[k] on stages k >= 1 and ns = nk on stage 0 (the reference's
// full trsv and full negation there).
//
// Work is level 2 and bound by HBM: each sweep reads every factor's lower
// triangle and every BAbt once.  Lanes own rows (gemv_n, trmv_lnn, the
// forward substitution) or columns (gemv_t, trmv_ltn); the vector a lane
// multiplies by is broadcast from LDS; the backward substitution's column
// sums are wave reductions with DPP moves.  Rows and columns go up to 128
// (two per lane).
#include "bfa_common.hpp"
#include "bfa_chol_core.hpp"
#include "bfa_internal.hpp"

namespace bfa {

constexpr int TRS_MAX = 2 * WAVE;   // largest nu+nx and nx of an instance

template <int PS>
struct PmRead
{
	const double *p;
	int cn;
	__device__ __forceinline__ double operator()(int i, int j) const
	{
		return p[(i & ~(PS - 1)) * cn + j * PS + (i & (PS - 1))];
	}
};

template <int PS>
__device__ __forceinline__ PmRead<PS> pm_read(const blasfeo_dmat &d)
{
	return PmRead<PS>{d.pA, d.cn};
}

// t(j) = sum_{i >= j} A(r0+i, r0+j) x(i), j < n   (A^T x, A lower; one column per lane)
template <int PS>
__device__ __forceinline__ void trs_trmv_t(const PmRead<PS> A, int r0, int n, const double *x, int j, double &t)
{
	double acc = 0.0;
#pragma unroll 8
	for (int i = j; i < n; i++)
		acc += A(r0 + i, r0 + j) * x[i];
	t = acc;
}

// y(i) = sum_{j <= i} A(r0+i, r0+j) t(j)   (A t, A lower; one row per lane)
template <int PS>
__device__ __forceinline__ double trs_trmv_n(const PmRead<PS> A, int r0, const double *t, int i)
{
	double acc = 0.0;
#pragma unroll 8
	for (int j = 0; j <= i; j++)
		acc += A(r0 + i, r0 + j) * t[j];
	return acc;
}

// L(i, j) for the lane's rows i = lane, lane + 64 with j < i < n0, else 0
template <int PS>
__device__ __forceinline__ void trs_coef(const PmRead<PS> L, int j, int n0, double (&e)[2])
{
	const int lane = threadIdx.x;
#pragma unroll
	for (int q = 0; q < 2; q++)
	{
		const int i = lane + q * WAVE;
		e[q] = (i > j && i < n0) ? L(i, j) : 0.0;
	}
}

// 1 / L(j, j): the cached reciprocal, or formed as the reference's trsv does
template <int PS>
__device__ __forceinline__ double trs_recip(const PmRead<PS> L, const double *dA, bool cached, int j)
{
	return cached ? dA[j] : 1.0 / L(j, j);
}

template <int PS>
__global__ __launch_bounds__(WAVE) void k_ric_trs(const blasfeo_amd_riccati_trs_arg *__restrict__ args, int batch)
{
	__shared__ double s_x[TRS_MAX];   // vector broadcast to the lanes
	__shared__ double s_t[TRS_MAX];   // intermediate of the trmv pairs
	__shared__ double s_p[TRS_MAX];   // ux[k+1] of the backward sweep
	const int inst = blockIdx.x;
	const int lane = threadIdx.x;
	if (inst >= batch)
		return;
	const blasfeo_amd_riccati_trs_arg a = args[inst];
	const int N = a.N;
	if (N < 0)
		return;
	// an instance the on-chip vectors cannot hold is left untouched (the host
	// rejects a hint above TRS_MAX)
	for (int k = 0; k <= N; k++)
		if (a.nu[k] + a.nx[k] > TRS_MAX || a.nx[k] > TRS_MAX || a.nu[k] < 0 || a.nx[k] < 0)
			return;

	// ---- backward substitution ----
	{
		const int n = a.nu[N] + a.nx[N];
		const double *rq = a.rq[N].pa;
		double *ux = a.ux[N].pa;
		for (int i = lane; i < n; i += WAVE)
		{
			const double v = rq[i];
			ux[i] = v;
			s_p[i] = v;
		}
	}
	__syncthreads();
	for (int k = N - 1; k >= 0; k--)
	{
		const int nu0 = a.nu[k], n0 = nu0 + a.nx[k];
		const int nu1 = a.nu[k + 1], nx1 = a.nx[k + 1];
		const PmRead<PS> L1 = pm_read<PS>(a.L[k + 1]);
		const PmRead<PS> B = pm_read<PS>(a.BAbt[k]);
		const double *b = a.b[k].pa;
		const double *rq = a.rq[k].pa;
		double *Pb = a.Pb[k].pa;
		double *ux = a.ux[k].pa;

		for (int i = lane; i < nx1; i += WAVE)
			s_x[i] = b[i];
		__syncthreads();
		for (int j = lane; j < nx1; j += WAVE)
			trs_trmv_t<PS>(L1, nu1, nx1, s_x, j, s_t[j]);
		__syncthreads();
		// Pb = Lxx t; the gemv_n operand w = Pb + ux[k+1](nu1:) goes to s_x
		for (int i = lane; i < nx1; i += WAVE)
		{
			const double pb = trs_trmv_n<PS>(L1, nu1, s_t, i);
			Pb[i] = pb;
			s_x[i] = pb + s_p[nu1 + i];
		}
		__syncthreads();

		// z = rq + BAbt w, rows lane and lane + 64
		double z[2];
#pragma unroll
		for (int q = 0; q < 2; q++)
		{
			const int i = lane + q * WAVE;
			z[q] = 0.0;
			if (i < n0)
			{
				double acc = 0.0;
#pragma unroll 8
				for (int j = 0; j < nx1; j++)
					acc += B(i, j) * s_x[j];
				z[q] = rq[i] + acc;
			}
		}

		// forward substitution with the first ns columns of L[k] (trsv_lnn_mn):
		// z(j) *= 1/L(j,j), then every later row drops L(i,j) z(j)
		const int ns = k == 0 ? n0 : nu0;
		const blasfeo_dmat *Ld = &a.L[k];
		const PmRead<PS> L0 = pm_read<PS>(*Ld);
		double *dA = Ld->dA;
		const bool cached = Ld->use_dA == 1;
		// the column's coefficients and reciprocal are loaded one column ahead,
		// off the serial chain of v_readlane broadcasts
		double e[2] = {0.0, 0.0}, d = 0.0;
		if (ns > 0)
		{
			trs_coef<PS>(L0, 0, n0, e);
			d = trs_recip<PS>(L0, dA, cached, 0);
		}
		for (int j = 0; j < ns; j++)
		{
			double en[2] = {0.0, 0.0}, dn = 0.0;
			const bool more = j + 1 < ns;
			if (more)
			{
				trs_coef<PS>(L0, j + 1, n0, en);
				dn = trs_recip<PS>(L0, dA, cached, j + 1);
			}
			const double zj = readlane(j < WAVE ? z[0] : z[1], j & (WAVE - 1)) * d;
			if (!cached && lane == 0)
				dA[j] = d;
#pragma unroll
			for (int q = 0; q < 2; q++)
			{
				const int i = lane + q * WAVE;
				if (i == j)
					z[q] = zj;
				else if (i > j && i < n0)
					z[q] -= e[q] * zj;
			}
			if (more)
			{
				e[0] = en[0];
				e[1] = en[1];
				d = dn;
			}
		}
		if (!cached && ns > 0 && n0 > 0 && lane == 0)
			a.L[k].use_dA = 1;
#pragma unroll
		for (int q = 0; q < 2; q++)
		{
			const int i = lane + q * WAVE;
			if (i < n0)
			{
				ux[i] = z[q];
				s_p[i] = z[q];
			}
		}
		__syncthreads();
	}

	// ---- forward substitution ----
	for (int k = 0; k < N; k++)
	{
		const int nu0 = a.nu[k], n0 = nu0 + a.nx[k];
		const int nu1 = a.nu[k + 1], nx1 = a.nx[k + 1];
		const PmRead<PS> L1 = pm_read<PS>(a.L[k + 1]);
		const PmRead<PS> B = pm_read<PS>(a.BAbt[k]);
		const double *b = a.b[k].pa;
		double *ux0 = a.ux[k].pa;
		double *ux1 = a.ux[k + 1].pa;
		double *pi = a.pi[k].pa;
		const int ns = k == 0 ? n0 : nu0;

		double pio[2], z[2];
#pragma unroll
		for (int q = 0; q < 2; q++)
		{
			const int i = lane + q * WAVE;
			pio[q] = i < nx1 ? ux1[nu1 + i] : 0.0;
			z[q] = i < n0 ? ux0[i] : 0.0;
			if (i < ns)
				z[q] = -z[q];
		}

		// backward substitution with the first ns columns of L[k] (trsv_ltn_mn):
		// z(j) = (z(j) - sum_{i > j} L(i,j) z(i)) / L(j,j), j = ns-1 .. 0
		const blasfeo_dmat *Ld = &a.L[k];
		const PmRead<PS> L0 = pm_read<PS>(*Ld);
		double *dA = Ld->dA;
		const bool cached = Ld->use_dA == 1;
		double e[2] = {0.0, 0.0}, d = 0.0;
		if (ns > 0)
		{
			trs_coef<PS>(L0, ns - 1, n0, e);
			d = trs_recip<PS>(L0, dA, cached, ns - 1);
		}
		for (int j = ns - 1; j >= 0; j--)
		{
			double en[2] = {0.0, 0.0}, dn = 0.0;
			const bool more = j > 0;
			if (more)
			{
				trs_coef<PS>(L0, j - 1, n0, en);
				dn = trs_recip<PS>(L0, dA, cached, j - 1);
			}
			// only rows j+1 .. n0-1 contribute (e is zero elsewhere, but a zero
			// coefficient times an Inf or NaN in an unsolved row is NaN)
			const int i0 = lane, i1 = lane + WAVE;
			const double part = (i0 > j && i0 < n0 ? e[0] * z[0] : 0.0) + (i1 > j && i1 < n0 ? e[1] * z[1] : 0.0);
			const double s = readlane(wave_sum_d(part), WAVE - 1);
			if (!cached && lane == 0)
				dA[j] = d;
#pragma unroll
			for (int q = 0; q < 2; q++)
				if (lane + q * WAVE == j)
					z[q] = (z[q] - s) * d;
			if (more)
			{
				e[0] = en[0];
				e[1] = en[1];
				d = dn;
			}
		}
		if (!cached && n0 > 0 && lane == 0)
			a.L[k].use_dA = 1;
#pragma unroll
		for (int q = 0; q < 2; q++)
		{
			const int i = lane + q * WAVE;
			if (i < n0)
			{
				ux0[i] = z[q];
				s_x[i] = z[q];
			}
		}
		__syncthreads();

		// x(k+1) = b + BAbt^T z (gemv_t), one column per lane
		for (int j = lane; j < nx1; j += WAVE)
		{
			double acc = 0.0;
#pragma unroll 8
			for (int i = 0; i < n0; i++)
				acc += B(i, j) * s_x[i];
			const double y = b[j] + acc;
			ux1[nu1 + j] = y;
			s_p[j] = y;
		}
		__syncthreads();
		for (int j = lane; j < nx1; j += WAVE)
			trs_trmv_t<PS>(L1, nu1, nx1, s_p, j, s_t[j]);
		__syncthreads();
#pragma unroll
		for (int q = 0; q < 2; q++)
		{
			const int i = lane + q * WAVE;
			if (i < nx1)
				pi[i] = pio[q] + trs_trmv_n<PS>(L1, nu1, s_t, i);
		}
		__syncthreads();
	}
}

template <int PS>
static int launch_trs(int batch, const blasfeo_amd_riccati_trs_arg *args, hipStream_t st, const char *name)
{
	hipLaunchKernelGGL(k_ric_trs<PS>, dim3(batch), dim3(WAVE), 0, st, args, batch);
	return check_launch(name);
}

} // namespace bfa

using namespace bfa;

extern "C" int blasfeo_amd_d_back_ric_trs_batched(int ps, int batch, const blasfeo_amd_riccati_trs_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream)
{
	const char *name = "blasfeo_amd_d_back_ric_trs_batched";
	if (batch < 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: batch=%d < 0", name, batch);
	if (batch == 0)
		return BLASFEO_AMD_OK;
	if (!args)
		return set_error(BLASFEO_AMD_EINVAL, "%s: args is NULL", name);
	if (ps != 4 && ps != 8)
		return set_error(BLASFEO_AMD_EINVAL, "%s: unsupported panel size %d", name, ps);
	if (!hint || hint->max_m <= 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: hint->max_m (largest nu[k]+nx[k]) is required", name);
	const int nxmax = hint->max_n > 0 ? hint->max_n : hint->max_m;
	if (hint->max_m > TRS_MAX || nxmax > TRS_MAX)
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: nu+nx and nx up to %d supported, got %d and %d",
		                 name, TRS_MAX, hint->max_m, nxmax);
	hipStream_t st = (hipStream_t)stream;
	return ps == 4 ? launch_trs<4>(batch, args, st, name) : launch_trs<8>(batch, args, st, name);
}
