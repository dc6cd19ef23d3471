// bfa_rowpe.hip -- batched row permutation blasfeo_drowpe (auxiliary/d_aux_lib4.c:2767):
// for ii = 0 .. kmax-1, rows ii and ipiv[ii] of sA are swapped over all of its
// columns, in that order (the LAPACK-style pivot vector getrf_rp returns).
//
// The swap sequence is composed first into one permutation in LDS (perm[i] =
// the original row that ends in row i; one thread replays the swaps on the
// index array, with ipiv staged in LDS by all threads), then every column of
// the matrix is gathered through it: all threads read the column's source
// rows into LDS, then write the rows back in place.  Matrices
// taller than RP_MAXM rows replay the swaps per column instead (each thread
// owns columns, so no synchronisation is needed there).
#include "bfa_common.hpp"
#include "bfa_internal.hpp"

namespace bfa {

// one wavefront per instance and a one-column gather buffer (6 KB of LDS), so
// that many instances share a CU: the composition of the swaps is one thread's
// serial chain, overlapped across instances
constexpr int RP_NT = 64;
constexpr int RP_MAXM = 512;     // rows of the LDS form
constexpr int RP_CB = 1;         // columns gathered per pass

template <int PS>
__global__ __launch_bounds__(RP_NT) void k_rowpe(const blasfeo_amd_drowpe_arg *__restrict__ args)
{
	__shared__ int perm[RP_MAXM];
	__shared__ double blk[RP_MAXM * RP_CB];
	const blasfeo_amd_drowpe_arg a = args[blockIdx.x];
	blasfeo_dmat *sA = a.sA;
	const int t = threadIdx.x;
	if (t == 0)
		sA->use_dA = 0;
	// entries outside the matrix (a malformed ipiv) are skipped, not followed
	const int m = sA->m, kmax = min(a.kmax, m), n = sA->n, cn = sA->cn;
	double *pA = sA->pA;
	const int *ipiv = a.ipiv;
	auto at = [&](int i, int c) -> double & {
		return pA[(size_t)(i & ~(PS - 1)) * cn + (size_t)c * PS + (i & (PS - 1))];
	};
	if (m > RP_MAXM)
	{
		for (int c = t; c < n; c += RP_NT)
			for (int ii = 0; ii < kmax; ii++)
			{
				const int p = ipiv[ii];
				if (p != ii && p >= 0 && p < m)
				{
					const double x = at(ii, c);
					at(ii, c) = at(p, c);
					at(p, c) = x;
				}
			}
		return;
	}
	// ipiv staged in the gather buffer (as ints), then the swaps composed
	int *piv_s = reinterpret_cast<int *>(blk);
	for (int i = t; i < m; i += RP_NT)
	{
		perm[i] = i;
		piv_s[i] = i < kmax ? ipiv[i] : i;
	}
	__syncthreads();
	if (t == 0)
	{
		for (int ii = 0; ii < kmax; ii++)
		{
			const int p = piv_s[ii];
			if (p != ii && p >= 0 && p < m)
			{
				const int x = perm[ii];
				perm[ii] = perm[p];
				perm[p] = x;
			}
		}
	}
	__syncthreads();
	for (int c0 = 0; c0 < n; c0 += RP_CB)
	{
		const int nc = min(RP_CB, n - c0);
		for (int e = t; e < m * RP_CB; e += RP_NT)
		{
			const int i = e % m, c = e / m;
			if (c < nc)
				blk[c * RP_MAXM + i] = at(perm[i], c0 + c);
		}
		__syncthreads();
		for (int e = t; e < m * RP_CB; e += RP_NT)
		{
			const int i = e % m, c = e / m;
			if (c < nc && perm[i] != i)
				at(i, c0 + c) = blk[c * RP_MAXM + i];
		}
		__syncthreads();
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
