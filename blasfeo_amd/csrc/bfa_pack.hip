// bfa_pack.hip -- batched column-major <-> panel-major conversion on the device
// (blasfeo_pack_dmat, blasfeo_pack_tran_dmat, blasfeo_unpack_dmat:
// auxiliary/d_aux_lib4.c:1607, :2086, :2238).
//
// Pure data movement, bound by HBM: every element is read once and written
// once.  The panel-major block is cut into tiles of 64 rows (aligned to the
// panel grid of the matrix) by 16 columns; a workgroup of 256 threads stages
// one tile in LDS.  The column-major side is read or written as 64-row runs of
// a column (512 contiguous bytes), the panel-major side as whole panel rows of
// the tile (16 * ps contiguous doubles), so both sides move in long
// contiguous runs.  For pack_tran the column-major side is walked along its
// columns, which are the tile's rows of the transposed block.
#include "bfa_common.hpp"
#include "bfa_internal.hpp"

namespace bfa {

constexpr int PK_R = 64;          // tile rows (multiple of every panel size)
constexpr int PK_C = 16;          // tile columns
constexpr int PK_NT = 256;
constexpr int PK_LD = PK_R + 1;   // LDS column stride (odd: distinct banks across columns)

enum PackDir { PK_PACK = 0, PK_TRAN = 1, PK_UNPACK = 2 };

template <int PS, int DIR>
__global__ __launch_bounds__(PK_NT) void k_pack(const blasfeo_amd_dpack_arg *__restrict__ args)
{
	__shared__ double T[PK_C * PK_LD];
	const blasfeo_amd_dpack_arg a = args[blockIdx.x];
	// the panel-major block: R rows x C columns at (ai, aj)
	const int R = DIR == PK_TRAN ? a.n : a.m;
	const int C = DIR == PK_TRAN ? a.m : a.n;
	blasfeo_dmat *sA = a.sA;
	// pack_tran clears the cache even for an empty block, pack returns first
	// (auxiliary/d_aux_lib4.c:2090, :1609)
	if (R <= 0 || C <= 0)
	{
		if (DIR == PK_TRAN && blockIdx.y == 0 && threadIdx.x == 0)
			sA->use_dA = 0;
		return;
	}
	double *pA = sA->pA;
	const int cn = sA->cn;
	const int lda = a.lda;
	double *A = a.A;
	const int base = a.ai - (a.ai & (PS - 1));     // first panel row of the block
	const int span = (a.ai & (PS - 1)) + R;        // rows from base to the block's end
	const int t = threadIdx.x;
	if (DIR != PK_UNPACK && blockIdx.y == 0 && t == 0)
		sA->use_dA = 0;

	for (int c0 = blockIdx.y * PK_C; c0 < C; c0 += gridDim.y * PK_C)
	{
		const int nc = min(PK_C, C - c0);
		for (int r0 = 0; r0 < span; r0 += PK_R)
		{
			// tile rows are absolute rows base + r0 + rr; block row = that - ai
			const int b0 = base + r0 - a.ai;
			if (DIR != PK_UNPACK)
			{
				// column-major -> LDS
				if (DIR == PK_PACK)
				{
#pragma unroll
					for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
					{
						const int rr = t & (PK_R - 1), cc = (t >> 6) + 4 * q;
						const int br = b0 + rr;
						if (cc < nc && br >= 0 && br < R)
							T[cc * PK_LD + rr] = A[br + (size_t)(c0 + cc) * lda];
					}
				}
				else
				{
					// transposed block (br, bc) = A[bc + br*lda]: runs along bc
#pragma unroll
					for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
					{
						const int cc = t & (PK_C - 1), rr = (t >> 4) + 16 * q;
						const int br = b0 + rr;
						if (cc < nc && br >= 0 && br < R)
							T[cc * PK_LD + rr] = A[(c0 + cc) + (size_t)br * lda];
					}
				}
				__syncthreads();
				// LDS -> panel-major: thread order (row in panel, column, panel)
#pragma unroll
				for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
				{
					const int e = t + q * PK_NT;
					const int r = e & (PS - 1);
					const int cc = (e / PS) & (PK_C - 1);
					const int p = e / (PS * PK_C);
					const int rr = p * PS + r;
					const int br = b0 + rr;
					if (cc < nc && br >= 0 && br < R)
						pA[(size_t)(base + r0 + p * PS) * cn + (size_t)(a.aj + c0 + cc) * PS + r] = T[cc * PK_LD + rr];
				}
			}
			else
			{
				// panel-major -> LDS
#pragma unroll
				for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
				{
					const int e = t + q * PK_NT;
					const int r = e & (PS - 1);
					const int cc = (e / PS) & (PK_C - 1);
					const int p = e / (PS * PK_C);
					const int rr = p * PS + r;
					const int br = b0 + rr;
					if (cc < nc && br >= 0 && br < R)
						T[cc * PK_LD + rr] = pA[(size_t)(base + r0 + p * PS) * cn + (size_t)(a.aj + c0 + cc) * PS + r];
				}
				__syncthreads();
				// LDS -> column-major
#pragma unroll
				for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
				{
					const int rr = t & (PK_R - 1), cc = (t >> 6) + 4 * q;
					const int br = b0 + rr;
					if (cc < nc && br >= 0 && br < R)
						A[br + (size_t)(c0 + cc) * lda] = T[cc * PK_LD + rr];
				}
			}
			__syncthreads();
		}
	}
}

// pack with the next tile's column-major loads issued before the current
// tile's panel-major stores (the tiles of a workgroup in one flat sequence:
// column tiles blockIdx.y, blockIdx.y + gridDim.y, ..., every row tile of each)
template <int PS>
__global__ __launch_bounds__(PK_NT) void k_pack_pipe(const blasfeo_amd_dpack_arg *__restrict__ args)
{
	__shared__ double T[PK_C * PK_LD];
	const blasfeo_amd_dpack_arg a = args[blockIdx.x];
	const int R = a.m, C = a.n;
	blasfeo_dmat *sA = a.sA;
	const int t = threadIdx.x;
	// an empty block leaves the cache alone (auxiliary/d_aux_lib4.c:1609)
	if (R <= 0 || C <= 0)
		return;
	if (blockIdx.y == 0 && t == 0)
		sA->use_dA = 0;
	double *pA = sA->pA;
	const int cn = sA->cn;
	const int lda = a.lda;
	const double *A = a.A;
	const int base = a.ai - (a.ai & (PS - 1));
	const int span = (a.ai & (PS - 1)) + R;
	const int nrt = (span + PK_R - 1) / PK_R;
	const int nct = (C + PK_C - 1) / PK_C;
	const int my_ct = nct > (int)blockIdx.y ? (nct - (int)blockIdx.y + (int)gridDim.y - 1) / (int)gridDim.y : 0;
	const int ntile = my_ct * nrt;
	double v[PK_R * PK_C / PK_NT];
	auto load = [&](int k) {
		const int c0 = ((int)blockIdx.y + (k / nrt) * (int)gridDim.y) * PK_C;
		const int r0 = (k % nrt) * PK_R;
		const int nc = min(PK_C, C - c0);
		const int b0 = base + r0 - a.ai;
#pragma unroll
		for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
		{
			const int rr = t & (PK_R - 1), cc = (t >> 6) + 4 * q;
			const int br = b0 + rr;
			v[q] = (cc < nc && br >= 0 && br < R) ? A[br + (size_t)(c0 + cc) * lda] : 0.0;
		}
	};
	if (ntile > 0)
		load(0);
	for (int k = 0; k < ntile; k++)
	{
#pragma unroll
		for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
			T[((t >> 6) + 4 * q) * PK_LD + (t & (PK_R - 1))] = v[q];
		__syncthreads();
		if (k + 1 < ntile)
			load(k + 1);
		const int c0 = ((int)blockIdx.y + (k / nrt) * (int)gridDim.y) * PK_C;
		const int r0 = (k % nrt) * PK_R;
		const int nc = min(PK_C, C - c0);
		const int b0 = base + r0 - a.ai;
#pragma unroll
		for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
		{
			const int e = t + q * PK_NT;
			const int r = e & (PS - 1);
			const int cc = (e / PS) & (PK_C - 1);
			const int p = e / (PS * PK_C);
			const int rr = p * PS + r;
			const int br = b0 + rr;
			if (cc < nc && br >= 0 && br < R)
				pA[(size_t)(base + r0 + p * PS) * cn + (size_t)(a.aj + c0 + cc) * PS + r] = T[cc * PK_LD + rr];
		}
		__syncthreads();
	}
}

// unpack with the same flat tile sequence: the next tile's panel-major loads
// are issued before the current tile's column-major stores
template <int PS>
__global__ __launch_bounds__(PK_NT) void k_unpack_pipe(const blasfeo_amd_dpack_arg *__restrict__ args)
{
	__shared__ double T[PK_C * PK_LD];
	const blasfeo_amd_dpack_arg a = args[blockIdx.x];
	const int R = a.m, C = a.n;
	if (R <= 0 || C <= 0)
		return;
	const blasfeo_dmat *sA = a.sA;
	const double *pA = sA->pA;
	const int cn = sA->cn;
	const int lda = a.lda;
	double *A = a.A;
	const int t = threadIdx.x;
	const int base = a.ai - (a.ai & (PS - 1));
	const int span = (a.ai & (PS - 1)) + R;
	const int nrt = (span + PK_R - 1) / PK_R;
	const int nct = (C + PK_C - 1) / PK_C;
	const int my_ct = nct > (int)blockIdx.y ? (nct - (int)blockIdx.y + (int)gridDim.y - 1) / (int)gridDim.y : 0;
	const int ntile = my_ct * nrt;
	double v[PK_R * PK_C / PK_NT];
	auto load = [&](int k) {
		const int c0 = ((int)blockIdx.y + (k / nrt) * (int)gridDim.y) * PK_C;
		const int r0 = (k % nrt) * PK_R;
		const int nc = min(PK_C, C - c0);
		const int b0 = base + r0 - a.ai;
#pragma unroll
		for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
		{
			const int e = t + q * PK_NT;
			const int r = e & (PS - 1);
			const int cc = (e / PS) & (PK_C - 1);
			const int p = e / (PS * PK_C);
			const int br = b0 + p * PS + r;
			v[q] = (cc < nc && br >= 0 && br < R)
			           ? pA[(size_t)(base + r0 + p * PS) * cn + (size_t)(a.aj + c0 + cc) * PS + r] : 0.0;
		}
	};
	if (ntile > 0)
		load(0);
	for (int k = 0; k < ntile; k++)
	{
#pragma unroll
		for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
		{
			const int e = t + q * PK_NT;
			const int r = e & (PS - 1);
			const int cc = (e / PS) & (PK_C - 1);
			const int p = e / (PS * PK_C);
			T[cc * PK_LD + p * PS + r] = v[q];
		}
		__syncthreads();
		if (k + 1 < ntile)
			load(k + 1);
		const int c0 = ((int)blockIdx.y + (k / nrt) * (int)gridDim.y) * PK_C;
		const int r0 = (k % nrt) * PK_R;
		const int nc = min(PK_C, C - c0);
		const int b0 = base + r0 - a.ai;
#pragma unroll
		for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
		{
			const int rr = t & (PK_R - 1), cc = (t >> 6) + 4 * q;
			const int br = b0 + rr;
			if (cc < nc && br >= 0 && br < R)
				A[br + (size_t)(c0 + cc) * lda] = T[cc * PK_LD + rr];
		}
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------
// blasfeo_dgecp (auxiliary/d_aux_lib4.c:3006) between panel-major matrices of
// any two panel sizes: B(bi+i, bj+j) = A(ai+i, aj+j), sB->use_dA = 0.  The
// panel sizes come from the blasfeo_pm_dmat descriptors; the tiles follow
// B's panel grid (writes in 16 * psB contiguous runs), A is read one column
// of a 64-row tile at a time (runs of psA contiguous doubles), the next
// tile's loads issued before the current tile's stores.
// ---------------------------------------------------------------------------
template <int PA, int PB>
__device__ __forceinline__ void pm_gecp_tiles(const blasfeo_amd_pm_dgecp_arg &a, double *T)
{
	// the flat tile sequence of k_pack_pipe on B's panel grid, the next tile's
	// loads from A issued before the current tile's stores to B
	const int m = a.m, n = a.n;
	const double *pa = a.sA->pA;
	double *pb = a.sB->pA;
	const int cna = a.sA->cn, cnb = a.sB->cn;
	const int base = a.bi - (a.bi & (PB - 1));
	const int span = (a.bi & (PB - 1)) + m;
	const int t = threadIdx.x;
	const int nrt = (span + PK_R - 1) / PK_R;
	const int nct = (n + PK_C - 1) / PK_C;
	const int my_ct = nct > (int)blockIdx.y ? (nct - (int)blockIdx.y + (int)gridDim.y - 1) / (int)gridDim.y : 0;
	const int ntile = my_ct * nrt;
	double v[PK_R * PK_C / PK_NT];
	auto load = [&](int k) {
		const int c0 = ((int)blockIdx.y + (k / nrt) * (int)gridDim.y) * PK_C;
		const int r0 = (k % nrt) * PK_R;
		const int nc = min(PK_C, n - c0);
		const int b0 = base + r0 - a.bi;
#pragma unroll
		for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
		{
			const int rr = t & (PK_R - 1), cc = (t >> 6) + 4 * q;
			const int br = b0 + rr;
			const int i = a.ai + br;
			v[q] = (cc < nc && br >= 0 && br < m)
			           ? pa[(size_t)(i & ~(PA - 1)) * cna + (size_t)(a.aj + c0 + cc) * PA + (i & (PA - 1))] : 0.0;
		}
	};
	if (ntile > 0)
		load(0);
	for (int k = 0; k < ntile; k++)
	{
#pragma unroll
		for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
			T[((t >> 6) + 4 * q) * PK_LD + (t & (PK_R - 1))] = v[q];
		__syncthreads();
		if (k + 1 < ntile)
			load(k + 1);
		const int c0 = ((int)blockIdx.y + (k / nrt) * (int)gridDim.y) * PK_C;
		const int r0 = (k % nrt) * PK_R;
		const int nc = min(PK_C, n - c0);
		const int b0 = base + r0 - a.bi;
#pragma unroll
		for (int q = 0; q < PK_R * PK_C / PK_NT; q++)
		{
			const int e = t + q * PK_NT;
			const int r = e & (PB - 1);
			const int cc = (e / PB) & (PK_C - 1);
			const int p = e / (PB * PK_C);
			const int br = b0 + p * PB + r;
			if (cc < nc && br >= 0 && br < m)
				pb[(size_t)(base + r0 + p * PB) * cnb + (size_t)(a.bj + c0 + cc) * PB + r] = T[cc * PK_LD + p * PB + r];
		}
		__syncthreads();
	}
}

__global__ __launch_bounds__(PK_NT) void k_pm_gecp(const blasfeo_amd_pm_dgecp_arg *__restrict__ args)
{
	__shared__ double T[PK_C * PK_LD];
	const blasfeo_amd_pm_dgecp_arg a = args[blockIdx.x];
	if (blockIdx.y == 0 && threadIdx.x == 0)
		a.sB->use_dA = 0;
	if (a.m <= 0 || a.n <= 0)
		return;
	const int pa = a.sA->ps, pb = a.sB->ps;
	// panel sizes other than 4 and 8 are left alone (the instance is skipped)
	if (pa == 4 && pb == 4)
		pm_gecp_tiles<4, 4>(a, T);
	else if (pa == 4 && pb == 8)
		pm_gecp_tiles<4, 8>(a, T);
	else if (pa == 8 && pb == 4)
		pm_gecp_tiles<8, 4>(a, T);
	else if (pa == 8 && pb == 8)
		pm_gecp_tiles<8, 8>(a, T);
}

template <int DIR>
static int launch_pack(int ps, int batch, const blasfeo_amd_dpack_arg *args, const blasfeo_amd_hint *hint,
	hipStream_t st, const char *name)
{
	// column tiles per instance spread over up to 16 workgroups
	const int nmax = hint && hint->max_n > 0 ? hint->max_n : PK_C;
	const int gy = max(1, min(16, (nmax + PK_C - 1) / PK_C));
	const dim3 grid(batch, gy);
	if (DIR == PK_PACK)
	{
		if (ps == 4)
			hipLaunchKernelGGL((k_pack_pipe<4>), grid, dim3(PK_NT), 0, st, args);
		else
			hipLaunchKernelGGL((k_pack_pipe<8>), grid, dim3(PK_NT), 0, st, args);
	}
	else if (DIR == PK_UNPACK)
	{
		if (ps == 4)
			hipLaunchKernelGGL((k_unpack_pipe<4>), grid, dim3(PK_NT), 0, st, args);
		else
			hipLaunchKernelGGL((k_unpack_pipe<8>), grid, dim3(PK_NT), 0, st, args);
	}
	else if (ps == 4)
		hipLaunchKernelGGL((k_pack<4, DIR>), grid, dim3(PK_NT), 0, st, args);
	else
		hipLaunchKernelGGL((k_pack<8, DIR>), grid, dim3(PK_NT), 0, st, args);
	return check_launch(name);
}

template <int DIR>
static int pack_entry(int ps, int batch, const blasfeo_amd_dpack_arg *args, const blasfeo_amd_hint *hint,
	blasfeo_amd_stream_t stream, const char *name)
{
	if (batch < 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: batch=%d < 0", name, batch);
	if (batch == 0)
		return BLASFEO_AMD_OK;
	if (!args)
		return set_error(BLASFEO_AMD_EINVAL, "%s: args is NULL", name);
	if (ps != 4 && ps != 8)
		return set_error(BLASFEO_AMD_EINVAL, "%s: unsupported panel size %d", name, ps);
	return launch_pack<DIR>(ps, batch, args, hint, (hipStream_t)stream, name);
}

} // namespace bfa

using namespace bfa;

extern "C" int blasfeo_amd_dpack_batched(int ps, int batch, const blasfeo_amd_dpack_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream)
{
	return pack_entry<PK_PACK>(ps, batch, args, hint, stream, "blasfeo_amd_dpack_batched");
}

extern "C" int blasfeo_amd_dpack_tran_batched(int ps, int batch, const blasfeo_amd_dpack_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream)
{
	return pack_entry<PK_TRAN>(ps, batch, args, hint, stream, "blasfeo_amd_dpack_tran_batched");
}

extern "C" int blasfeo_amd_pm_dgecp_batched(int batch, const blasfeo_amd_pm_dgecp_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream)
{
	const char *name = "blasfeo_amd_pm_dgecp_batched";
	if (batch < 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: batch=%d < 0", name, batch);
	if (batch == 0)
		return BLASFEO_AMD_OK;
	if (!args)
		return set_error(BLASFEO_AMD_EINVAL, "%s: args is NULL", name);
	const int nmax = hint && hint->max_n > 0 ? hint->max_n : PK_C;
	const int gy = max(1, min(16, (nmax + PK_C - 1) / PK_C));
	hipLaunchKernelGGL(k_pm_gecp, dim3(batch, gy), dim3(PK_NT), 0, (hipStream_t)stream, args);
	return check_launch(name);
}

extern "C" int blasfeo_amd_dunpack_batched(int ps, int batch, const blasfeo_amd_dpack_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream)
{
	return pack_entry<PK_UNPACK>(ps, batch, args, hint, stream, "blasfeo_amd_dunpack_batched");
}
