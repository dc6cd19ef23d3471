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
		return set_error(BLASFEO_AMD_EINVAL, "%s: hint->max_m (largest nu[k]+nx[k]) is required", name);
	const int mmax = hint->max_m;
	if (mmax > fact_mmax<double>())
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: nu+nx up to %d supported, got %d", name, fact_mmax<double>(), mmax);
	const size_t shm = sizeof(double) * PW * ((mmax + 15) / 16 * 16 + 1);
	hipStream_t st = (hipStream_t)stream;
	if (ps == 4)
		hipLaunchKernelGGL((k_ric_trf<4>), dim3(batch), dim3(NT3), shm, st, args, mmax);
	else
		hipLaunchKernelGGL((k_ric_trf<8>), dim3(batch), dim3(NT3), shm, st, args, mmax);
	return check_launch(name);
}
