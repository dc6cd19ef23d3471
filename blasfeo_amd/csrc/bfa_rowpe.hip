// bfa_rowpe.hip -- batched row permutation blasfeo_drowpe (auxiliary/d_aux_lib4.c:2767):
// for ii = 0 .. kmax-1, rows ii and ipiv[ii] of sA are swapped over all of its
// columns, in that order (the LAPACK-style pivot vector getrf_rp returns).
//
// The swaps are sequential in ii but independent across columns, so each
// thread of an instance's workgroup replays the whole swap sequence on its own
// columns: no synchronisation, and every swap of a column stays in the thread
// that owns it.
#include "bfa_common.hpp"
#include "bfa_internal.hpp"

namespace bfa {

constexpr int RP_NT = 64;

template <int PS>
__global__ __launch_bounds__(RP_NT) void k_rowpe(const blasfeo_amd_drowpe_arg *__restrict__ args)
{
	const blasfeo_amd_drowpe_arg a = args[blockIdx.x];
	blasfeo_dmat *sA = a.sA;
	if (threadIdx.x == 0)
		sA->use_dA = 0;
	// entries outside the matrix (a malformed ipiv) are skipped, not followed
	const int m = sA->m, kmax = min(a.kmax, m), n = sA->n, cn = sA->cn;
	double *pA = sA->pA;
	const int *ipiv = a.ipiv;
	for (int c = threadIdx.x; c < n; c += RP_NT)
	{
		double *col = pA + (size_t)c * PS;
		for (int ii = 0; ii < kmax; ii++)
		{
			const int p = ipiv[ii];
			if (p != ii && p >= 0 && p < m)
			{
				double *x = col + (size_t)(ii & ~(PS - 1)) * cn + (ii & (PS - 1));
				double *y = col + (size_t)(p & ~(PS - 1)) * cn + (p & (PS - 1));
				const double t = *x;
				*x = *y;
				*y = t;
			}
		}
	}
}

} // namespace bfa

using namespace bfa;

extern "C" int blasfeo_amd_drowpe_batched(int ps, int batch, const blasfeo_amd_drowpe_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream)
{
	const char *name = "blasfeo_amd_drowpe_batched";
	(void)hint;
	if (batch < 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: batch=%d < 0", name, batch);
	if (batch == 0)
		return BLASFEO_AMD_OK;
	if (!args)
		return set_error(BLASFEO_AMD_EINVAL, "%s: args is NULL", name);
	if (ps != 4 && ps != 8)
		return set_error(BLASFEO_AMD_EINVAL, "%s: unsupported panel size %d", name, ps);
	hipStream_t st = (hipStream_t)stream;
	if (ps == 4)
		hipLaunchKernelGGL(k_rowpe<4>, dim3(batch), dim3(RP_NT), 0, st, args);
	else
		hipLaunchKernelGGL(k_rowpe<8>, dim3(batch), dim3(RP_NT), 0, st, args);
	return check_launch(name);
}
