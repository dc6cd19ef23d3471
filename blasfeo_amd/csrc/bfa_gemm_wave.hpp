// bfa_gemm_wave.hpp -- gemm_nt / gemm_nn / syrk_ln(_mn) with one wavefront per
// 32x32 output tile, for batches of small and uneven instances.
//
// The wave's 2x2 block of 16x16 matrix-core accumulators takes its operands
// straight from global memory: lane l loads A(i0 + 16x + l%16, k0 + l/16) and
// the matching B element for each 4-column step of K, which in panel-major
// storage is four contiguous 128-byte runs per operand (one per 4-row panel
// for ps = 4), so the loads coalesce without staging through LDS.  With no
// LDS and no barriers, a wave never waits for another, a CU holds as many
// waves as registers allow (their load latencies overlap), and an instance of
// m = n = 4 costs one wave instead of a 256-thread workgroup with a 64x64
// tile.  Each wave keeps the operands of the next two K steps in flight.
// The role of the reference's small-size kernels (the vs/gen edge kernels of
// blasfeo_hp_pm/d_blas3_lib4.c:800-872,1137-1374).
#pragma once
#include "bfa_gemm_core.hpp"

namespace bfa {

constexpr int GW_T = 32;   // tile edge

// tile t of an instance -> (ti, tj); syrk walks the lower tiles column by
// column (tiles ti >= tj, tj < ntj)
template <int OP>
__device__ __forceinline__ void gw_tile(int t, int mt, int nt, int &ti, int &tj)
{
	if (OP == OP_SYRK || OP == OP_SYRK_MN)
	{
		int j = 0, rem = t;
		while (rem >= mt - j)
		{
			rem -= mt - j;
			j++;
		}
		ti = j + rem;
		tj = j;
	}
	else
	{
		ti = t % mt;
		tj = t / mt;
	}
}

template <int OP>
__device__ __forceinline__ int gw_ntiles(int mt, int nt)
{
	if (OP == OP_SYRK || OP == OP_SYRK_MN)
	{
		const int c = nt < mt ? nt : mt;
		return c * mt - c * (c - 1) / 2;
	}
	return mt * nt;
}

// operands of one 4-column K step (k0) for this lane
template <typename T, int PS, int OP>
__device__ __forceinline__ void gw_load(const PM<T, PS> &gA, int ai, int aj, const PM<T, PS> &gB, int bi, int bj,
	int i0, int j0, int m, int n, int k, int k0, T (&av)[2], T (&bv)[2])
{
	const int lane = threadIdx.x & (WAVE - 1);
	const int kq = k0 + (lane >> 4);
	const bool kok = kq < k;
#pragma unroll
	for (int x = 0; x < 2; x++)
	{
		const int i = i0 + 16 * x + (lane & 15);
		av[x] = (kok && i < m) ? gA.at(ai + i, aj + kq) : T(0);
	}
#pragma unroll
	for (int y = 0; y < 2; y++)
	{
		const int j = j0 + 16 * y + (lane & 15);
		if (OP == OP_NN)
			bv[y] = (kok && j < n) ? gB.at(bi + kq, bj + j) : T(0);
		else
			bv[y] = (kok && j < n) ? gB.at(bi + j, bj + kq) : T(0);
	}
}

// One wave per (instance, tile slot): `slots` slots per instance (the tile
// count of the hint's largest instance), a 1-D launch of batch * slots
// workgroups rounded up to a multiple of XCDS.  The slots of one instance
// map to consecutive logical indices on one XCD (xcd_local_index), so the
// tiles that re-read the same A row block or B column block run together
// on one L2; slot t >= the instance's tile count exits at once.  Registers
// are capped for five waves per SIMD (92 VGPRs, no spills), so that more
// waves' load latencies overlap: faster than four (124 VGPRs) on the
// 100000-call uneven batch; six and eight spill and are slower.
template <typename T, int PS, int OP>
__global__ __launch_bounds__(WAVE, 5) void k_gemm_wave(const typename GemmArg<T>::type *__restrict__ args, int batch,
	int slots)
{
	typedef typename Vec4<T>::type V4;
	const unsigned li = xcd_local_index(blockIdx.x, gridDim.x);
	const int inst = (int)(li / (unsigned)slots), slot = (int)(li % (unsigned)slots);
	if (inst >= batch)
		return;
	const typename GemmArg<T>::type &a = args[inst];
	const int m = a.m;
	const int n = (OP == OP_SYRK) ? a.m : a.n;
	const int k = a.k;
	if (m <= 0 || n <= 0)
		return;
	if (threadIdx.x == 0 && slot == 0)
		a.sD->use_dA = 0;
	const PM<T, PS> gA = view<T, PS>(a.sA), gB = view<T, PS>(a.sB), gC = view<T, PS>(a.sC);
	PM<T, PS> gD = view<T, PS>(a.sD);
	const T alpha = a.alpha, beta = a.beta;
	const int lane = threadIdx.x;
	const int mt = (m + GW_T - 1) / GW_T, nt = (n + GW_T - 1) / GW_T;
	const int ntiles = gw_ntiles<OP>(mt, nt);
	for (int t = slot; t < ntiles; t += slots)
	{
		int ti, tj;
		gw_tile<OP>(t, mt, nt, ti, tj);
		const int i0 = ti * GW_T, j0 = tj * GW_T;
		V4 acc[2][2];
#pragma unroll
		for (int x = 0; x < 2; x++)
#pragma unroll
			for (int y = 0; y < 2; y++)
				acc[x][y] = V4{0, 0, 0, 0};
		// two K steps of operands in flight
		T a0[2], b0[2], a1[2], b1[2];
		gw_load<T, PS, OP>(gA, a.ai, a.aj, gB, a.bi, a.bj, i0, j0, m, n, k, 0, a0, b0);
		gw_load<T, PS, OP>(gA, a.ai, a.aj, gB, a.bi, a.bj, i0, j0, m, n, k, 4, a1, b1);
		for (int k0 = 0; k0 < k; k0 += 8)
		{
#pragma unroll
			for (int x = 0; x < 2; x++)
#pragma unroll
				for (int y = 0; y < 2; y++)
					acc[x][y] = mfma16(a0[x], b0[y], acc[x][y]);
			if (k0 + 8 < k)
				gw_load<T, PS, OP>(gA, a.ai, a.aj, gB, a.bi, a.bj, i0, j0, m, n, k, k0 + 8, a0, b0);
			if (k0 + 4 < k)
			{
#pragma unroll
				for (int x = 0; x < 2; x++)
#pragma unroll
					for (int y = 0; y < 2; y++)
						acc[x][y] = mfma16(a1[x], b1[y], acc[x][y]);
				if (k0 + 12 < k)
					gw_load<T, PS, OP>(gA, a.ai, a.aj, gB, a.bi, a.bj, i0, j0, m, n, k, k0 + 12, a1, b1);
			}
		}
		// epilogue: D = beta*C + alpha*acc on the valid (for syrk: lower)
		// elements; each row half's C loads issued before its D stores
#pragma unroll
		for (int x = 0; x < 2; x++)
		{
			T cv[2][4];
#pragma unroll
			for (int y = 0; y < 2; y++)
#pragma unroll
				for (int r = 0; r < 4; r++)
				{
					const int gi = i0 + 16 * x + acc_row<T>(lane, r), gj = j0 + 16 * y + (lane & 15);
					const bool need = beta != T(0) && gi < m && gj < n &&
						!((OP == OP_SYRK || OP == OP_SYRK_MN) && gi < gj);
					cv[y][r] = need ? gC.at(a.ci + gi, a.cj + gj) : T(0);
				}
#pragma unroll
			for (int y = 0; y < 2; y++)
#pragma unroll
				for (int r = 0; r < 4; r++)
				{
					const int gi = i0 + 16 * x + acc_row<T>(lane, r), gj = j0 + 16 * y + (lane & 15);
					if (gi >= m || gj >= n || ((OP == OP_SYRK || OP == OP_SYRK_MN) && gi < gj))
						continue;
					T val = alpha * acc[x][y][r];
					if (beta != T(0))
						val = beta * cv[y][r] + val;
					gD.at(a.di + gi, a.dj + gj) = val;
				}
		}
	}
}

} // namespace bfa
