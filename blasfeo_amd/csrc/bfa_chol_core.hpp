// bfa_chol_core.hpp -- left-looking blocked Cholesky of one instance by one
// workgroup, shared by the potrf kernels and the fused Riccati sweep.
#pragma once
#include "bfa_common.hpp"

namespace bfa {

constexpr int NTL = 256;
constexpr int PW = 16;  // panel width

__device__ __forceinline__ double readlane(double v, int l)
{
	int2 x = __builtin_bit_cast(int2, v);
	x.x = __builtin_amdgcn_readlane(x.x, l);
	x.y = __builtin_amdgcn_readlane(x.y, l);
	return __builtin_bit_cast(double, x);
}
__device__ __forceinline__ float readlane(float v, int l)
{
	return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// Largest m a launch supports (LDS panel of (m+1) x 16 elements).
template <typename T> constexpr int fact_mmax() { return sizeof(T) == 8 ? 1024 : 2048; }

// ---------------------------------------------------------------------------
// Cholesky family: D(m x n lower trapezoid) = chol(C + A*B^T), n <= m.
// k == 0 gives potrf_l / potrf_l_mn.  Zero-pivot rule of
// blasfeo_ref/x_lapack_ref.c:84-92: pivot <= 0 (or NaN) stores L_jj = c*0 and dA = 0.
// ---------------------------------------------------------------------------
template <typename T, int PS>
__device__ void chol_instance(T *P, int LDP, T *L11, T *invs,
	int m, int n, int k,
	const PM<T, PS> gA, int ai, int aj, const PM<T, PS> gB, int bi, int bj,
	const PM<T, PS> gC, int ci, int cj, PM<T, PS> gD, int di, int dj, T *dD)
{
	typedef typename Vec4<T>::type V4;
	const int lane = lane_id();
	const int w = wave_id();
	for (int c0 = 0; c0 < n; c0 += PW)
	{
		const int cw = min(PW, n - c0);
		const int nr = m - c0;            // panel rows c0..m-1
		const int nrt = (nr + 15) / 16;   // 16-row tiles
		__syncthreads();
		g2s<T, PS, NTL, false>(gC, ci + c0, cj + c0, nrt * 16, PW, nr, cw, P, LDP);
		__syncthreads();
		// matrix-core updates of the panel, 16-row tile per wave
		if (k > 0 || c0 > 0)
		{
			for (int rt = w; rt < nrt; rt += NTL / WAVE)
			{
				V4 acc = V4{0, 0, 0, 0};
				const int ir = rt * 16 + (lane & 15);     // panel-local row of the A operand
				const bool iok = ir < nr;
				const bool jok = (lane & 15) < cw;
				// + A(c0+i, :) * B(c0+j, :)^T
#pragma unroll 1
				for (int kk = 0; kk < k; kk += 4)
				{
					const int kq = kk + (lane >> 4);
					const bool kok = kq < k;
					T av = (iok && kok) ? gA.at(ai + c0 + ir, aj + kq) : T(0);
					T bv = (jok && kok) ? gB.at(bi + c0 + (lane & 15), bj + kq) : T(0);
					acc = mfma16(av, bv, acc);
				}
				// - L(c0+i, 0:c0) * L(c0+j, 0:c0)^T  (c0 is a multiple of 16)
#pragma unroll 1
				for (int kk = 0; kk < c0; kk += 4)
				{
					const int kq = kk + (lane >> 4);
					T av = iok ? gD.at(di + c0 + ir, dj + kq) : T(0);
					T bv = jok ? gD.at(di + c0 + (lane & 15), dj + kq) : T(0);
					acc = mfma16(-av, bv, acc);
				}
#pragma unroll
				for (int r = 0; r < 4; r++)
				{
					const int il = rt * 16 + acc_row<T>(lane, r);
					P[(lane & 15) * LDP + il] += acc[r];
				}
			}
			__syncthreads();
		}
		// wave 0: rows 0..63 of the panel in registers (lane = row)
		if (w == 0)
		{
			T x[PW];
#pragma unroll
			for (int j = 0; j < PW; j++)
				x[j] = (lane < nr) ? P[j * LDP + lane] : T(0);
#pragma unroll
			for (int j = 0; j < PW; j++)
			{
				if (j < cw)
				{
					const T c = readlane(x[j], j);
					const T inv = (c > T(0)) ? rsqrt_ref<T>(c) : T(0);
					if (lane > j)
						x[j] *= inv;
					else if (lane == j)
						x[j] = c * inv;
					if (lane == 0)
						invs[j] = inv;
#pragma unroll
					for (int jj = j + 1; jj < PW; jj++)
					{
						if (jj < cw)
						{
							const T l = readlane(x[j], jj);
							if (lane >= jj)
								x[jj] -= x[j] * l;
						}
					}
				}
			}
#pragma unroll
			for (int j = 0; j < PW; j++)
			{
				if (lane < nr)
					P[j * LDP + lane] = x[j];
				if (lane < PW)
					L11[lane * (PW + 1) + j] = x[j];
			}
		}
		__syncthreads();
		// rows 64.. of the panel: solve against L11 (x_j = (x_j - sum_l<j x_l L_jl) * inv_j)
#pragma unroll 1
		for (int il = 64 + threadIdx.x; il < nr; il += NTL)
		{
			// keep the L11 reads inside the loop: hoisted, the 136 loop-invariant
			// values would pin ~270 registers for a loop of at most a few trips
			asm volatile("" ::: "memory");
			T x[PW];
#pragma unroll
			for (int j = 0; j < PW; j++)
				x[j] = P[j * LDP + il];
#pragma unroll
			for (int j = 0; j < PW; j++)
			{
				if (j < cw)
				{
					T c = x[j];
#pragma unroll
					for (int l = 0; l < j; l++)
						c -= x[l] * L11[j * (PW + 1) + l];
					x[j] = c * invs[j];
				}
			}
#pragma unroll
			for (int j = 0; j < PW; j++)
				P[j * LDP + il] = x[j];
		}
		__syncthreads();
		s2g<T, PS, NTL, true>(gD, di + c0, dj + c0, nr, cw, P, LDP, 0);
		for (int j = threadIdx.x; j < cw; j += NTL)
			dD[c0 + j] = invs[j];
	}
}

} // namespace bfa

namespace bfa {

// ---------------------------------------------------------------------------
// In-LDS right-looking blocked Cholesky of the m x n lower trapezoid (m >= n)
// of a matrix held in LDS and addressed through M(r, c) (column-major, or
// 16x16 tiles; m <= NMAX <= 128), in place: M <- L (potrf_l_mn semantics:
// L11 = chol(C11), L21 = C21 L11^{-T}),
// invd[j] = 1/sqrt(pivot) with the zero-pivot rule above (L_jj = c*invd[j]);
// invd must hold 16*ceil(n/16) entries (a full panel stores all 16).
// Rows m .. ceil16(m)-1 and columns n .. ceil16(n)-1 must hold zeros on entry
// (with PADCOLS the padding columns hold unspecified values on exit).
// Per 16-column block: wave 0 factors the whole panel (diagonal block and the
// rows below, two rows per lane, one when NMAX <= 64) in registers,
// broadcasting L(jj,j) with v_readlane, so the panel needs no workgroup
// barrier; the rank-16 trailing updates run on the matrix cores, a group of up
// to four tiles of one block column per task with independent accumulators
// (look-ahead order below).
// While wave 0 factors the first panel, the other waves call side(): work of
// the caller's that touches nothing the factorisation reads or writes (a
// no-op by default; with one wave, or no panel, every wave calls it once).
// Called by all NT threads of the workgroup; starts and ends with a barrier.
// ---------------------------------------------------------------------------
struct NoSide
{
	__device__ __forceinline__ void operator()() const {}
};

// Rank-16 update of NTL tiles (g..g+NTL-1, bj) of the trailing matrix by the
// panel at column c0: all operands (accumulators, the panel's A values of each
// tile, its B values of block column bj) are loaded before the first
// matrix-core step, so a wave waits once per task rather than once per step.
template <typename T, int NTL, typename ACC>
__device__ __forceinline__ void chol_update_tiles(const ACC M, int g, int bj, int c0)
{
	typedef typename Vec4<T>::type V4;
	const int lane = lane_id();
	const int lq = lane >> 4, lr = lane & 15;
	V4 acc[NTL];
	T a[NTL][4], b[4];
#pragma unroll
	for (int t = 0; t < NTL; t++)
#pragma unroll
		for (int r = 0; r < 4; r++)
			acc[t][r] = M((g + t) * 16 + acc_row<T>(lane, r), bj * 16 + lr);
#pragma unroll
	for (int s = 0; s < 4; s++)
		b[s] = M(bj * 16 + lr, c0 + 4 * s + lq);
#pragma unroll
	for (int t = 0; t < NTL; t++)
#pragma unroll
		for (int s = 0; s < 4; s++)
			a[t][s] = M((g + t) * 16 + lr, c0 + 4 * s + lq);
#pragma unroll
	for (int s = 0; s < 4; s++)
#pragma unroll
		for (int t = 0; t < NTL; t++)
			acc[t] = mfma16(-a[t][s], b[s], acc[t]);
#pragma unroll
	for (int t = 0; t < NTL; t++)
#pragma unroll
		for (int r = 0; r < 4; r++)
			M((g + t) * 16 + acc_row<T>(lane, r), bj * 16 + lr) = acc[t][r];
}

template <typename T, int NT, int NMAX, typename ACC, bool PADCOLS = true, typename SIDE = NoSide>
__device__ __forceinline__ void chol_acc(const ACC M, int m, int n, T *invd, const SIDE &side = SIDE{})
{
	m = __builtin_amdgcn_readfirstlane(m);
	n = __builtin_amdgcn_readfirstlane(n);
	typedef typename Vec4<T>::type V4;
	constexpr int NW = NT / WAVE;
	const int lane = lane_id();
	const int w = wave_id();
	const int nt = (n + 15) / 16;   // panels
	const int mt = (m + 15) / 16;   // tile rows
	__syncthreads();
	// no first panel (n = 0): every wave calls side() here; with one wave it
	// is called after the factorisation
	if (nt == 0)
		side();
	// tiles (g..g+3, bj) (fewer at the bottom) updated by the panel at column cp
	auto upd = [&](int g, int bj, int cp) {
		const int ntl = min(4, mt - g);
		if (ntl == 4)
			chol_update_tiles<T, 4>(M, g, bj, cp);
		else if (ntl == 3)
			chol_update_tiles<T, 3>(M, g, bj, cp);
		else if (ntl == 2)
			chol_update_tiles<T, 2>(M, g, bj, cp);
		else
			chol_update_tiles<T, 1>(M, g, bj, cp);
	};
	// the block columns bj0..nt-1 updated by the panel at column cp: tasks
	// (block column, group of up to 4 tiles) dealt to waves w0, w0+ws, ...
	auto upd_rest = [&](int bj0, int cp, int w0, int ws) {
		int ntask = 0;
		for (int bj = bj0; bj < nt; bj++)
			ntask += (mt - bj + 3) / 4;
		for (int task = w0; task < ntask; task += ws)
		{
			int bj = bj0, rem = task;
			while (rem >= (mt - bj + 3) / 4)
			{
				rem -= (mt - bj + 3) / 4;
				bj++;
			}
			upd(bj + 4 * rem, bj, cp);
		}
	};
	// Look-ahead: in step jb, wave 0 applies the previous panel's update to
	// block column jb alone and factors panel jb, while the other waves apply
	// the previous panel's update to the block columns after jb (side() in
	// step 0), so the matrix-core updates run beside the panel's serial chain
	// and a step needs one barrier.
	for (int jb = 0; jb < nt; jb++)
	{
		const int c0 = jb * 16;
		const int cw = min(16, n - c0);
		const int nr = m - c0;
		if (w == 0)
		{
			if (jb > 0)
			{
				for (int g = jb; g < mt; g += 4)
					upd(g, jb, c0 - 16);
				if (NW == 1)
					upd_rest(jb + 1, c0 - 16, 0, 1);
				// the panel below reads what other lanes of this wave stored
				wave_sync();
			}
			T x0[16], x1[16];
			const bool r0ok = lane < nr, r1ok = NMAX > 64 && lane + 64 < nr;
#pragma unroll
			for (int j = 0; j < 16; j++)
				x0[j] = x1[j] = T(0);
			// one guarded block per row set (per-element guards become one
			// exec-mask branch per load)
			if (r0ok)
#pragma unroll
				for (int j = 0; j < 16; j++)
					x0[j] = M(c0 + lane, c0 + j);
			if (r1ok)
#pragma unroll
				for (int j = 0; j < 16; j++)
					x1[j] = M(c0 + 64 + lane, c0 + j);
			// guarded form: lane j keeps 1/sqrt(pivot j), stored after the panel
			T myinv = T(0);
			// PADCOLS: the padding columns of a partial panel are factored too
			// (unguarded); otherwise a partial panel takes the guarded form.
			if (PADCOLS || cw == 16)
			{
				// No guards: the rows above the diagonal (lanes < j of the diagonal
				// tile) are scaled and updated too, but nothing below the diagonal
				// reads them and they are never stored to D.  Columns cw..15 of a
				// partial panel are factored as well; they are padding columns,
				// which no real column reads (column jj is updated only from
				// columns j < jj) and no store returns; invd is stored for
				// columns < cw only.
				// Lane j holds the pivot c itself, so scaling every lane gives
				// it c * inv without a per-column lane select, and every lane
				// stores the (uniform) reciprocal: per-column lane masks held
				// 32 SGPRs across the panel, and the compiler spilled the
				// broadcast values to VGPR lanes and reloaded them.
				if (!(NMAX > 64 && nr > 64))
				{
#pragma unroll
					for (int j = 0; j < 16; j++)
					{
						const T c = readlane(x0[j], j);
						const T inv = (c > T(0)) ? rsqrt_fast(c) : T(0);
						x0[j] *= inv;
						invd[c0 + j] = inv;
#pragma unroll
						for (int jj = j + 1; jj < 16; jj++)
						{
							const T l = readlane(x0[j], jj);
							x0[jj] -= x0[j] * l;
						}
					}
				}
				else
				{
					// the same with a second row per lane (rows 64..127)
#pragma unroll
					for (int j = 0; j < 16; j++)
					{
						const T c = readlane(x0[j], j);
						const T inv = (c > T(0)) ? rsqrt_fast(c) : T(0);
						x0[j] *= inv;
						x1[j] *= inv;
						invd[c0 + j] = inv;
#pragma unroll
						for (int jj = j + 1; jj < 16; jj++)
						{
							const T l = readlane(x0[j], jj);
							x0[jj] -= x0[j] * l;
							x1[jj] -= x1[j] * l;
						}
					}
				}
			}
			else
			{
				// partial panel, per-column and per-lane guards (fewer
				// registers than the unguarded form)
#pragma unroll
				for (int j = 0; j < 16; j++)
				{
					if (j < cw)
					{
						const T c = readlane(x0[j], j);
						const T inv = (c > T(0)) ? rsqrt_fast(c) : T(0);
						if (lane > j)
							x0[j] *= inv;
						else if (lane == j)
							x0[j] = c * inv;
						x1[j] *= inv;
						myinv = (lane == j) ? inv : myinv;
#pragma unroll
						for (int jj = j + 1; jj < 16; jj++)
						{
							if (jj < cw)
							{
								const T l = readlane(x0[j], jj);
								if (lane >= jj)
									x0[jj] -= x0[j] * l;
								x1[jj] -= x1[j] * l;
							}
						}
					}
				}
			}
			if (!(PADCOLS || cw == 16) && lane < cw)
				invd[c0 + lane] = myinv;
#pragma unroll
			for (int j = 0; j < 16; j++)
			{
				if (r0ok)
					M(c0 + lane, c0 + j) = x0[j];
				if (r1ok)
					M(c0 + 64 + lane, c0 + j) = x1[j];
			}
		}
		else if (jb == 0)
			side();
		else if constexpr (NW > 1)
			upd_rest(jb + 1, c0 - 16, w - 1, NW - 1);
		__syncthreads();
	}
	if (NW == 1 && nt > 0)
		side();
	if (nt == 0 || NW == 1)
		__syncthreads();
}


// Column-major LDS matrix: element (r, c) at p[c*ld + r].
template <typename T>
struct ColMajor
{
	T *p;
	int ld;
	__device__ __forceinline__ T &operator()(int r, int c) const { return p[c * ld + r]; }
};

// Packed lower tiles in LDS: 16x16 tiles with a leading dimension of 17 inside
// the tile (a row of a tile then spans distinct banks), tile (I,J), I >= J,
// at I(I+1)/2 + J.  The Riccati sweep keeps its stage matrix so, and the
// size-bucketed LDS Cholesky its trapezoid: mt(mt+1)/2 tiles for mt tile
// rows, 78 KB at 128 rows against 133 KB for the square column-major form.
constexpr int RT_LD = 17;
constexpr int RT_SZ = 16 * RT_LD;

__host__ __device__ __forceinline__ int ric_tix(int I, int J) { return I * (I + 1) / 2 + J; }

// T: the arithmetic type (double; float for the mixed-precision sweep)
template <typename T>
struct TileLowerT
{
	T *p;
	__device__ __forceinline__ T &operator()(int r, int c) const
	{
		return p[ric_tix(r >> 4, c >> 4) * RT_SZ + (c & 15) * RT_LD + (r & 15)];
	}
};

template <typename T, int NT = NTL, int NMAX = 128>
__device__ __forceinline__ void chol_lds(T *M, int ld, int n, T *invd)
{
	chol_acc<T, NT, NMAX>(ColMajor<T>{M, ld}, n, n, invd);
}

} // namespace bfa
