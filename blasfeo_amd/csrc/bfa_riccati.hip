// bfa_riccati.hip -- fused backward Riccati factorisation sweep
// (d_back_ric_trf_libstr, examples/example_d_riccati_recursion.c:115-140).
//
// One workgroup runs one instance's whole sweep: L[N] = chol(RSQ[N]), then for
// k = N-1 .. 0
//     work = BAbt[k] * L[k+1](nu[k+1]:, nu[k+1]:)        (trmm_rlnn)
//     L[k] = chol(RSQ[k] + work * work^T)                (syrk_dpotrf_ln)
// with the matrix-core tile engine and the panel Cholesky of the single
// routines; stage k+1's factor is read back from L2 by the same workgroup, so
// the whole horizon runs in one launch with no host round trip.
#include "bfa_gemm_core.hpp"
#include "bfa_chol_core.hpp"
#include "bfa_internal.hpp"

namespace bfa {

template <int PS>
__global__ __launch_bounds__(NT3) void k_ric_trf(const blasfeo_amd_riccati_arg *__restrict__ args, int mmax)
{
	extern __shared__ unsigned char smem_raw[];
	__shared__ double As[KC * LDT];
	__shared__ double Bs[KC * LDT];
	__shared__ double L11[PW * (PW + 1)];
	__shared__ double invs[PW];
	double *P = reinterpret_cast<double *>(smem_raw);
	const int LDP = (mmax + 15) / 16 * 16 + 1;

	const blasfeo_amd_riccati_arg &a = args[blockIdx.x];
	const int N = a.N;
	if (N < 0)
		return;
	const int *nx = a.nx;
	const int *nu = a.nu;
	blasfeo_dmat *L = a.L;
	blasfeo_dmat *RSQ = a.RSQrq;
	blasfeo_dmat *BAbt = a.BAbt;
	blasfeo_dmat *W = a.work;

	// last stage: L[N] = chol(RSQ[N]) on nx[N] x nx[N]
	{
		const int m = nx[N];
		if (m > 0 && m <= mmax)
		{
			PM<double, PS> gL = view<double, PS>(&L[N]);
			chol_instance<double, PS>(P, LDP, L11, invs, m, m, 0, gL, 0, 0, gL, 0, 0,
				view<double, PS>(&RSQ[N]), 0, 0, gL, 0, 0, L[N].dA);
		}
		if (threadIdx.x == 0)
			L[N].use_dA = m;
	}
	for (int s = 0; s < N; s++)
	{
		const int k = N - 1 - s;
		const int nk = nu[k] + nx[k];
		const int nx1 = nx[k + 1];
		const int nu1 = nu[k + 1];
		if (nk > mmax)
			return;
		__syncthreads();
		// work = BAbt[k] * L[k+1](nu1.., nu1..)
		if (nk > 0 && nx1 > 0)
		{
			PM<double, PS> gL1 = view<double, PS>(&L[k + 1]);
			gemm_core<double, PS, OP_TRMM_RLNN>(As, Bs, nk, nx1, nx1, 1.0, 0.0,
				gL1, nu1, nu1, view<double, PS>(&BAbt[k]), 0, 0,
				gL1, 0, 0, view<double, PS>(W), 0, 0);
		}
		__syncthreads();
		// L[k] = chol(RSQ[k] + work * work^T)
		if (nk > 0)
		{
			PM<double, PS> gW = view<double, PS>(W);
			chol_instance<double, PS>(P, LDP, L11, invs, nk, nk, nx1, gW, 0, 0, gW, 0, 0,
				view<double, PS>(&RSQ[k]), 0, 0, view<double, PS>(&L[k]), 0, 0, L[k].dA);
		}
		if (threadIdx.x == 0)
		{
			W->use_dA = 0;
			L[k].use_dA = 1;
		}
	}
}

// ---------------------------------------------------------------------------
// LDS-resident sweep (nu[k]+nx[k] <= 128, nx[k] <= 64): the stage's working
// set never leaves the CU.  LDS holds
//   M [NKP x LDM]: L[k+1] at the start of stage k, then RSQ[k] + W W^T, then L[k]
//   W [NXP x LDM]: BAbt[k], then W = BAbt[k] * L[k+1](nu1:, nu1:)
// (column-major, LDM = NKP + 1 so that 16 lanes reading a row hit distinct
// banks).  Per stage: stage BAbt, W <- W*Lxx on the matrix cores (each wave
// keeps its 16-row strip's 16x64 result in registers and writes it back in
// place), stage RSQ (lower), M += W W^T on the lower tiles, in-LDS blocked
// Cholesky, stream L[k] and its reciprocal diagonal out.
// ---------------------------------------------------------------------------
template <int PS>
__global__ __launch_bounds__(NT3) void k_ric_trf_lds(const blasfeo_amd_riccati_arg *__restrict__ args, int nkp, int nxp)
{
	typedef d4_t V4;
	extern __shared__ unsigned char smem_raw[];
	const int LDM = nkp + 1;
	double *M = reinterpret_cast<double *>(smem_raw);
	double *W = M + (size_t)nkp * LDM;
	double *invd = W + (size_t)nxp * LDM;
	const int lane = lane_id();
	const int w = wave_id();

	const blasfeo_amd_riccati_arg &a = args[blockIdx.x];
	const int N = a.N;
	if (N < 0)
		return;
	const int *nx = a.nx;
	const int *nu = a.nu;
	blasfeo_dmat *L = a.L;
	blasfeo_dmat *RSQ = a.RSQrq;
	blasfeo_dmat *BAbt = a.BAbt;
	// sizes are checked up front so that every workgroup either runs the
	// whole sweep or returns before its first barrier
	for (int k = 0; k <= N; k++)
		if (nx[k] > nxp || nx[k] + (k < N ? nu[k] : 0) > nkp || nx[k] < 0 || nu[k] < 0)
			return;

	// last stage: L[N] = chol(Q_N)
	{
		const int m = nx[N];
		const int mt = (m + 15) / 16;
		g2s<double, PS, NT3, false, true>(view<double, PS>(&RSQ[N]), 0, 0, mt * 16, mt * 16, m, m, M, LDM);
		chol_lds<double>(M, LDM, m, invd);
		s2g<double, PS, NT3, true>(view<double, PS>(&L[N]), 0, 0, m, m, M, LDM, 0);
		for (int j = threadIdx.x; j < m; j += NT3)
			L[N].dA[j] = invd[j];
		if (threadIdx.x == 0)
			L[N].use_dA = m;
	}
	for (int s = 0; s < N; s++)
	{
		const int k = N - 1 - s;
		const int nk = nu[k] + nx[k];
		const int n1 = nx[k + 1];
		const int u1 = nu[k + 1];
		const int nkt = (nk + 15) / 16;
		const int n1t = (n1 + 15) / 16;
		// stage BAbt[k] (nk x n1) into W; M still holds L[k+1]
		g2s<double, PS, NT3, false>(view<double, PS>(&BAbt[k]), 0, 0, nkt * 16, n1t * 16, nk, n1, W, LDM);
		__syncthreads();
		// W <- W * Lxx, Lxx = L[k+1](u1:u1+n1, u1:u1+n1) lower, 16-row strips per wave
		{
			V4 acc[4];
			for (int rt = w; rt < nkt; rt += NT3 / WAVE)
			{
#pragma unroll
				for (int ct = 0; ct < 4; ct++)
					acc[ct] = V4{0, 0, 0, 0};
				for (int l0 = 0; l0 < n1; l0 += 4)
				{
					const int lk = l0 + (lane >> 4);
					const double av = W[lk * LDM + rt * 16 + (lane & 15)];
#pragma unroll
					for (int ct = 0; ct < 4; ct++)
					{
						if (ct < n1t && l0 + 3 >= ct * 16)
						{
							const int j = ct * 16 + (lane & 15);
							const double bv = (lk < n1 && lk >= j) ? M[(u1 + j) * LDM + u1 + lk] : 0.0;
							acc[ct] = mfma16(av, bv, acc[ct]);
						}
					}
				}
				// every wave reads only its own strip of W, so the strip can be
				// overwritten as soon as this wave is done with it
#pragma unroll
				for (int ct = 0; ct < 4; ct++)
					if (ct < n1t)
#pragma unroll
						for (int r = 0; r < 4; r++)
							W[(ct * 16 + (lane & 15)) * LDM + rt * 16 + acc_row<double>(lane, r)] = acc[ct][r];
			}
		}
		__syncthreads();
		// M <- lower(RSQ[k]) (L[k+1] is no longer needed)
		g2s<double, PS, NT3, false, true>(view<double, PS>(&RSQ[k]), 0, 0, nkt * 16, nkt * 16, nk, nk, M, LDM);
		__syncthreads();
		// M += W W^T on the lower tiles
		{
			const int ntiles = nkt * (nkt + 1) / 2;
			for (int t = w; t < ntiles; t += NT3 / WAVE)
			{
				int bj = 0, rem = t;
				while (rem >= nkt - bj)
				{
					rem -= nkt - bj;
					bj++;
				}
				const int bi = bj + rem;
				V4 acc;
#pragma unroll
				for (int r = 0; r < 4; r++)
					acc[r] = M[(bj * 16 + (lane & 15)) * LDM + bi * 16 + acc_row<double>(lane, r)];
				for (int l0 = 0; l0 < n1; l0 += 4)
				{
					const int lk = l0 + (lane >> 4);
					const double av = W[lk * LDM + bi * 16 + (lane & 15)];
					const double bv = W[lk * LDM + bj * 16 + (lane & 15)];
					acc = mfma16(av, bv, acc);
				}
#pragma unroll
				for (int r = 0; r < 4; r++)
					M[(bj * 16 + (lane & 15)) * LDM + bi * 16 + acc_row<double>(lane, r)] = acc[r];
			}
		}
		// L[k] = chol(M)
		chol_lds<double>(M, LDM, nk, invd);
		s2g<double, PS, NT3, true>(view<double, PS>(&L[k]), 0, 0, nk, nk, M, LDM, 0);
		for (int j = threadIdx.x; j < nk; j += NT3)
			L[k].dA[j] = invd[j];
		if (threadIdx.x == 0)
		{
			L[k].use_dA = 1;
			a.work->use_dA = 0;
		}
	}
}

} // namespace bfa

using namespace bfa;

extern "C" int blasfeo_amd_d_back_ric_trf_batched(int ps, int batch, const blasfeo_amd_riccati_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream)
{
	const char *name = "blasfeo_amd_d_back_ric_trf_batched";
	if (batch < 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: batch=%d < 0", name, batch);
	if (batch == 0)
		return BLASFEO_AMD_OK;
	if (!args)
		return set_error(BLASFEO_AMD_EINVAL, "%s: args is NULL", name);
	if (ps != 4 && ps != 8)
		return set_error(BLASFEO_AMD_EINVAL, "%s: unsupported panel size %d", name, ps);
	if (!hint || hint->max_m <= 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: hint->max_m (largest nu[k]+nx[k]) is required; hint->max_n (largest nx[k]) selects the LDS-resident form", name);
	const int mmax = hint->max_m;
	if (mmax > fact_mmax<double>())
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: nu+nx up to %d supported, got %d", name, fact_mmax<double>(), mmax);
	hipStream_t st = (hipStream_t)stream;
	const int nxmax = hint->max_n > 0 ? hint->max_n : mmax;
	if (mmax <= 128 && nxmax <= 64)
	{
		// LDS-resident sweep
		const int nkp = (mmax + 15) / 16 * 16;
		const int nxp = (nxmax + 15) / 16 * 16;
		const size_t shm = sizeof(double) * ((size_t)(nkp + nxp) * (nkp + 1) + nkp);
		if (ps == 4)
			hipLaunchKernelGGL((k_ric_trf_lds<4>), dim3(batch), dim3(NT3), shm, st, args, nkp, nxp);
		else
			hipLaunchKernelGGL((k_ric_trf_lds<8>), dim3(batch), dim3(NT3), shm, st, args, nkp, nxp);
		return check_launch(name);
	}
	const size_t shm = sizeof(double) * PW * ((mmax + 15) / 16 * 16 + 1);
	if (ps == 4)
		hipLaunchKernelGGL((k_ric_trf<4>), dim3(batch), dim3(NT3), shm, st, args, mmax);
	else
		hipLaunchKernelGGL((k_ric_trf<8>), dim3(batch), dim3(NT3), shm, st, args, mmax);
	return check_launch(name);
}
