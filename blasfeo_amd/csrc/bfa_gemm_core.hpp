// bfa_gemm_core.hpp -- the 64x64-tile matrix-core gemm used by the level-3
// kernels and by the fused Riccati sweep.
//
// A workgroup is 4 waves; each output tile is 64x64 and each wave owns a 32x32
// quarter as 2x2 matrix-core 16x16 accumulators.  Operands go global -> LDS in
// K-chunks of 16 columns (the LDS-staged replacement of the reference's
// register-block kernels, kernel/avx2/kernel_dgemm_12x4_lib4.S and friends),
// with the next chunk prefetched into registers while the current one is
// consumed.
#pragma once
#include "bfa_common.hpp"

namespace bfa {

enum { OP_NT = 0, OP_NN = 1, OP_SYRK = 2, OP_SYRK_MN = 3, OP_TRMM_RLNN = 4 };

constexpr int NT3 = 256;  // threads per workgroup
constexpr int TM = 64;    // tile rows
constexpr int TN = 64;    // tile cols
constexpr int KC = 16;    // K chunk
constexpr int LDT = TM + 16;  // LDS leading dimension of a K-major tile

template <typename T> struct GemmArg;
template <> struct GemmArg<double> { typedef blasfeo_amd_dgemm_arg type; };
template <> struct GemmArg<float> { typedef blasfeo_amd_sgemm_arg type; };
template <typename T> struct TrsmArg;
template <> struct TrsmArg<double> { typedef blasfeo_amd_dtrsm_arg type; };
template <> struct TrsmArg<float> { typedef blasfeo_amd_strsm_arg type; };

// Register-staged prefetch of one K-chunk of an operand tile, walking the
// covering panels in memory order (see g2s).  MAXL bounds the elements a
// thread holds: (64 + PS) rows x 16 cols or (16 + PS) rows x 64 cols over 256 threads.
template <typename T, int PS>
struct TileLoader
{
	static constexpr int NPA = (TM + 2 * PS - 2) / PS;   // covering panels of a 64-row tile
	static constexpr int NPB = (KC + 2 * PS - 2) / PS;   // covering panels of a 16-row tile
	static constexpr int LA = (NPA * PS * KC + NT3 - 1) / NT3;
	static constexpr int LB = (NPB * PS * TN + NT3 - 1) / NT3;
	static constexpr int MAXL = LA > LB ? LA : LB;
	T v[MAXL];
	int dst[MAXL];

	// R x C block at absolute (r0, c0); valid extents Rv x Cv; TRANS stores
	// element (i,c) at i*LDT + c instead of c*LDT + i.  LOWERMASK zeroes
	// elements with (i + lo_i) < (c + lo_c), i.e. above the diagonal of the
	// operand in its own local coordinates.
	template <bool TRANS, bool LOWERMASK>
	__device__ __forceinline__ void load(const PM<T, PS> g, int r0, int c0, int R, int C, int Rv, int Cv, int lo_i, int lo_c)
	{
		const int rr = r0 & (PS - 1);
		const int np = (rr + R + PS - 1) / PS;
		const int run = PS * C;
		const int total = np * run;
		const BFA_GLOBAL T *base = (const BFA_GLOBAL T *)g.p + (size_t)(r0 - rr) * g.cn + (size_t)c0 * PS;
#pragma unroll
		for (int s = 0; s < MAXL; s++)
		{
			const int e = threadIdx.x + s * NT3;
			dst[s] = -1;
			v[s] = T(0);
			if (e < total)
			{
				const int p = e / run;
				const int rem = e - p * run;
				const int c = rem / PS;
				const int q = rem - c * PS;
				const int i = p * PS + q - rr;
				if (i >= 0 && i < R)
				{
					dst[s] = TRANS ? i * LDT + c : c * LDT + i;
					bool ok = i < Rv && c < Cv;
					if (LOWERMASK)
						ok = ok && (i + lo_i >= c + lo_c);
					if (ok)
						v[s] = base[(size_t)p * PS * g.cn + (size_t)c * PS + q];
				}
			}
		}
	}
	__device__ __forceinline__ void store(T *lds) const
	{
#pragma unroll
		for (int s = 0; s < MAXL; s++)
			if (dst[s] >= 0)
				lds[dst[s]] = v[s];
	}
};

// One gemm-family instance, tiles t0, t0 + tstep, ...
// OP_NT/OP_SYRK*: D = beta*C + alpha*A*B^T (B is n x k);  OP_NN: B is k x n;
// OP_TRMM_RLNN: D = alpha*B*A with A lower k x k (here A is the right operand).
template <typename T, int PS, int OP>
__device__ __forceinline__ void gemm_core(T *As, T *Bs, int m, int n, int k, T alpha, T beta,
	const PM<T, PS> gA, int ai, int aj, const PM<T, PS> gB, int bi, int bj,
	const PM<T, PS> gC, int ci, int cj, PM<T, PS> gD, int di, int dj, int t0, int tstep)
{
	typedef typename Vec4<T>::type V4;
	const int mt = (m + TM - 1) / TM;
	const int nt = (n + TN - 1) / TN;
	const int ntiles = mt * nt;
	const int lane = lane_id();
	const int w = wave_id();
	const int wr = (w >> 1) * 32;  // wave's row offset in the tile
	const int wc = (w & 1) * 32;   // wave's col offset in the tile

	for (int t = t0; t < ntiles; t += tstep)
	{
		const int ti = t % mt;
		const int tj = t / mt;
		const int i0 = ti * TM, j0 = tj * TN;
		if ((OP == OP_SYRK || OP == OP_SYRK_MN) && j0 > i0 + TM - 1)
			continue;
		const int mv = min(TM, m - i0);
		const int nv = min(TN, n - j0);

		V4 acc[2][2];
#pragma unroll
		for (int x = 0; x < 2; x++)
#pragma unroll
			for (int y = 0; y < 2; y++)
				acc[x][y] = V4{0, 0, 0, 0};

		// K range: trmm_rlnn only needs l >= j0 (A lower)
		const int kbeg = (OP == OP_TRMM_RLNN) ? (j0 / KC) * KC : 0;
		TileLoader<T, PS> la, lb;
		auto fetch = [&](int kk) {
			const int kv = min(KC, k - kk);
			if (OP == OP_TRMM_RLNN)
			{
				// left operand: B(i0.., kk..) of the call (sB), right: A(kk.., j0..) lower
				la.template load<false, false>(gB, bi + i0, bj + kk, TM, KC, mv, kv, 0, 0);
				lb.template load<true, true>(gA, ai + kk, aj + j0, KC, TN, kv, nv, kk, j0);
			}
			else
			{
				la.template load<false, false>(gA, ai + i0, aj + kk, TM, KC, mv, kv, 0, 0);
				if (OP == OP_NN)
					lb.template load<true, false>(gB, bi + kk, bj + j0, KC, TN, kv, nv, 0, 0);
				else
					lb.template load<false, false>(gB, bi + j0, bj + kk, TN, KC, nv, kv, 0, 0);
			}
		};
		if (kbeg < k)
			fetch(kbeg);
		for (int kk = kbeg; kk < k; kk += KC)
		{
			__syncthreads();
			la.store(As);
			lb.store(Bs);
			__syncthreads();
			if (kk + KC < k)
				fetch(kk + KC);
#pragma unroll
			for (int k4 = 0; k4 < KC; k4 += 4)
			{
				const int kr = (k4 + (lane >> 4)) * LDT + (lane & 15);
				T av[2], bv[2];
#pragma unroll
				for (int x = 0; x < 2; x++)
					av[x] = As[kr + wr + 16 * x];
#pragma unroll
				for (int y = 0; y < 2; y++)
					bv[y] = Bs[kr + wc + 16 * y];
#pragma unroll
				for (int x = 0; x < 2; x++)
#pragma unroll
					for (int y = 0; y < 2; y++)
						acc[x][y] = mfma16(av[x], bv[y], acc[x][y]);
			}
		}

		// epilogue: D = beta*C + alpha*acc (trmm: alpha*acc), element-wise,
		// a row half (x) of the wave's quarter at a time: its 8 elements of C
		// are all loaded before any of D is stored (C may be D: in place,
		// each element is read before it is written either way), so the
		// loads overlap instead of each waiting behind the previous store.
#pragma unroll
		for (int x = 0; x < 2; x++)
		{
			T cv[2][4];
#pragma unroll
			for (int y = 0; y < 2; y++)
			{
				const int jl = wc + 16 * y + (lane & 15);
#pragma unroll
				for (int r = 0; r < 4; r++)
				{
					const int il = wr + 16 * x + acc_row<T>(lane, r);
					const int gi = i0 + il, gj = j0 + jl;
					const bool need = OP != OP_TRMM_RLNN && beta != T(0) && il < mv && jl < nv &&
						!((OP == OP_SYRK || OP == OP_SYRK_MN) && gi < gj);
					cv[y][r] = need ? gC.at(ci + gi, cj + gj) : T(0);
				}
			}
#pragma unroll
			for (int y = 0; y < 2; y++)
			{
				const int jl = wc + 16 * y + (lane & 15);
#pragma unroll
				for (int r = 0; r < 4; r++)
				{
					const int il = wr + 16 * x + acc_row<T>(lane, r);
					if (il >= mv || jl >= nv)
						continue;
					const int gi = i0 + il, gj = j0 + jl;
					if ((OP == OP_SYRK || OP == OP_SYRK_MN) && gi < gj)
						continue;
					T val = alpha * acc[x][y][r];
					if (OP != OP_TRMM_RLNN && beta != T(0))
						val = beta * cv[y][r] + val;
					gD.at(di + gi, dj + gj) = val;
				}
			}
		}
	}
}

} // namespace bfa
