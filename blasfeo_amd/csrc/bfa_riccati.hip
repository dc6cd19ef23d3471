// bfa_riccati.hip -- fused backward Riccati factorisation sweep
// (d_back_ric_trf_libstr, examples/example_d_riccati_recursion.c:115-140).
//
// One instance's whole sweep runs in one workgroup: L[N] = chol(RSQ[N]), then
// for k = N-1 .. 0
//     W    = BAbt[k] * L[k+1](nu[k+1]:, nu[k+1]:)      (trmm_rlnn)
//     L[k] = chol(RSQ[k] + W * W^T)                    (syrk_dpotrf_ln)
//
// Launch model (size-bucketed work queue):
//   1. a plan kernel gives every instance an LDS footprint class from its
//      largest stage and appends its index to that class's list;
//   2. one persistent launch per class, largest first, the launches spread
//      over the caller's stream and helper streams so that the classes run
//      side by side; the workgroups of a launch pull instances from the
//      class's atomic work queue.  Two or four waves per instance by default
//      (one wave when the tuning knobs ask for it), a few instances per CU so
//      that their latency chains overlap;
//   3. instances too large for LDS (nx > 64 or footprint > 160 KB) take the
//      general sweep, which keeps only a 16-column panel in LDS.
#include <mutex>
#include <map>

#include "bfa_gemm_core.hpp"
#include "bfa_chol_core.hpp"
#include "bfa_internal.hpp"

namespace bfa {

constexpr int RIC_LDS_MAX = 163840;   // LDS bytes a workgroup may use on gfx950
constexpr int RIC_NCLASS = 8;         // LDS-resident footprint classes
constexpr int RIC_GENERAL = RIC_NCLASS;
constexpr int RIC_HEAD = 32;          // offset of the work-queue heads in the plan workspace
static_assert(RIC_HEAD + RIC_NCLASS <= 64, "plan workspace header holds 64 ints");
// class capacities in bytes: multiples of 1 KiB that divide the CU's LDS
// into 20, 10, 6, 5, 4, 3, 2 and 1 resident instances
__constant__ int c_ric_cap[RIC_NCLASS] = {8192, 16384, 24576, 32768, 40960, 54272, 81920, 163840};
static const int h_ric_cap[RIC_NCLASS] = {8192, 16384, 24576, 32768, 40960, 54272, 81920, 163840};

// LDS storage of the resident form: the stage matrix M as packed lower tiles
// (TileLowerT, bfa_chol_core.hpp) and W as a rectangle of the same 16x17
// tiles.  Against a square column-major M this keeps only the lower tiles,
// so the largest stage (nu+nx = 80, nx = 64) needs 76,800 B and two
// instances share a CU instead of one.
typedef TileLowerT<double> TileLower;

template <typename T>
struct TileRectT
{
	T *p;
	int nct;  // tile columns per tile row
	__device__ __forceinline__ T &operator()(int r, int c) const
	{
		return p[((r >> 4) * nct + (c >> 4)) * RT_SZ + (c & 15) * RT_LD + (r & 15)];
	}
};
typedef TileRectT<double> TileRect;

// The first RIC_QSLOT bytes of a class kernel's dynamic LDS hold the work-queue
// slot (bytes 0..3) and a 0.0 of the sweep's type (just before the sweep's M);
// the class kernels declare no static LDS, so a launch asks for exactly its
// class cap (at most the 160 KB of a CU).
constexpr int RIC_QSLOT = 16;

// LDS bytes of the resident form for an instance whose largest stage is
// nkmax = max(nu+nx) and whose largest state is nxmax (queue slot included),
// with esz-byte elements.
__host__ __device__ __forceinline__ int ric_footprint(int nkmax, int nxmax, int esz = 8)
{
	const int nkt = (nkmax + 15) / 16;
	const int nxt = (nxmax + 15) / 16;
	return ((nkt * (nkt + 1) / 2 + nkt * nxt) * RT_SZ + nkt * 16) * esz + RIC_QSLOT;
}

// ---------------------------------------------------------------------------
// the stage sweep with the working set in LDS (NT threads per instance)
//   M (packed lower tiles): L[k+1] at the start of stage k, then
//     RSQ[k] + W W^T, then L[k]
//   W (tile rectangle): BAbt[k], then W = BAbt[k] * L[k+1](nu1:, nu1:)
// The two matrix-core stages of a Riccati step are inlined into the kernels
// of two and four waves per instance.  The one-wave kernel calls them out of
// line (inlined, their address arithmetic is hoisted across the whole sweep
// and it needs 248 VGPRs + 64 AGPRs, one wave per SIMD), with the LDS base
// passed opaquely (see lds_opaque).
// An LDS address rebuilt from its 32-bit offset behind an empty asm: the
// callee then keeps the base in a register instead of resolving the kernel's
// dynamic-LDS variable through the offset table at every access.
__device__ __forceinline__ double *lds_opaque(double *p)
{
	unsigned off = (unsigned)(size_t)(__attribute__((address_space(3))) double *)p;
	asm volatile("" : "+v"(off));
	off = __builtin_amdgcn_readfirstlane(off);
	return (double *)(__attribute__((address_space(3))) double *)(size_t)off;
}

// ---------------------------------------------------------------------------
// A value every lane holds, moved to a scalar register: loops and branches
// that it controls are then uniform (scalar branches, no exec masking), so
// the operand loads of a block of matrix-core steps can all be issued before
// the first of them waits.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// W <- W * Lxx (trmm_rlnn), Lxx = L[k+1](u1:u1+n1, u1:u1+n1), held in M.
// 16-row strips per wave, the strip's 16 x 64 result in four accumulators;
// only this wave reads its strip, so it is overwritten in place.  Per 16-wide
// block C of the inner dimension the block's operands are loaded first (the
// strip's four A columns, and Lxx's B values for the column tiles ct <= C, the
// others being zero above the diagonal), then its matrix-core steps run.  The
// B loads are unconditional (a conditional load would be waited for on its
// own): row u1 + min(lk, n1 - 1), column u1 + min(j, n1 - 1), inside M's
// packed tiles, and for an element above Lxx's diagonal the address of the
// zero kept just before M (ric_zero_slot).  Rows lk >= n1 need no zeroing:
// W's columns from n1 on (the A operand) are zero.
template <int NT, typename T = double>
__device__ __forceinline__ void ric_trmm_w(const TileLowerT<T> M, const TileRectT<T> W, int nkt, int n1, int n1t, int u1)
{
	typedef typename Vec4<T>::type V4;
	constexpr int NW = NT / WAVE;
	const int lane = lane_id();
	const int w = wave_id();
	const int lq = lane >> 4, lr = lane & 15;
	int hcol[4], jv[4];
#pragma unroll
	for (int ct = 0; ct < 4; ct++)
	{
		jv[ct] = ct * 16 + lr;
		const int cc = u1 + min(jv[ct], n1 - 1);
		hcol[ct] = (cc >> 4) * RT_SZ + (cc & 15) * RT_LD;
	}
	for (int rt = w; rt < nkt; rt += NW)
	{
		V4 acc[4];
#pragma unroll
		for (int ct = 0; ct < 4; ct++)
			acc[ct] = V4{0, 0, 0, 0};
		const T *pa = W.p + (size_t)rt * W.nct * RT_SZ + lq * RT_LD + lr;
		constexpr int ZOFF = -1;   // M.p[-1] holds 0.0
#pragma unroll
		for (int C = 0; C < 4; C++)
		{
			if (C < n1t)
			{
				T a[4], b[4][4];
#pragma unroll
				for (int s = 0; s < 4; s++)
					a[s] = pa[C * RT_SZ + s * 4 * RT_LD];
#pragma unroll
				for (int s = 0; s < 4; s++)
				{
					const int lk = C * 16 + s * 4 + lq;
					const int rr = u1 + min(lk, n1 - 1);
					const int R = rr >> 4;
					const int roff = (R * (R + 1) / 2) * RT_SZ + (rr & 15);
#pragma unroll
					for (int ct = 0; ct <= C; ct++)
						b[ct][s] = M.p[lk >= jv[ct] ? roff + hcol[ct] : ZOFF];
				}
#pragma unroll
				for (int s = 0; s < 4; s++)
#pragma unroll
					for (int ct = 0; ct <= C; ct++)
						acc[ct] = mfma16(a[s], b[ct][s], acc[ct]);
			}
		}
		wave_sync();
#pragma unroll
		for (int ct = 0; ct < 4; ct++)
			if (ct < n1t)
#pragma unroll
				for (int r = 0; r < 4; r++)
					W(rt * 16 + acc_row<T>(lane, r), ct * 16 + lr) = acc[ct][r];
	}
}

// acc[t] += W(strip I[t]) W(strip J[t])^T over the inner tiles 0..kt-1 of
// W, for NTL tiles; per inner tile the 8*NTL operands are loaded before its
// matrix-core steps.
template <int NTL, typename T>
__device__ __forceinline__ void ric_syrk_tiles(const TileRectT<T> W, const int (&I)[4], const int (&J)[4], int kt,
	typename Vec4<T>::type (&acc)[4])
{
	const int lane = lane_id();
	const int lo = (lane >> 4) * RT_LD + (lane & 15);
	const int rs = W.nct * RT_SZ;   // one strip
	for (int C = 0; C < kt; C++)
	{
		T a[NTL][4], b[NTL][4];
#pragma unroll
		for (int t = 0; t < NTL; t++)
#pragma unroll
			for (int s = 0; s < 4; s++)
			{
				a[t][s] = W.p[I[t] * rs + C * RT_SZ + s * 4 * RT_LD + lo];
				b[t][s] = W.p[J[t] * rs + C * RT_SZ + s * 4 * RT_LD + lo];
			}
#pragma unroll
		for (int s = 0; s < 4; s++)
#pragma unroll
			for (int t = 0; t < NTL; t++)
				acc[t] = mfma16(a[t][s], b[t][s], acc[t]);
	}
}

// One wave's syrk tiles taken four at a time: chunk c0 of the wave's list
// (tiles w + NW*(c0..c0+3)), and with FROMG the chunk's RSQ values.
template <typename T>
struct RicChunk
{
	int I[4], J[4], cnt;
	typename Vec4<T>::type rq[4];
};

template <int NT, int PS, bool FROMG, bool AUG>
__device__ __forceinline__ int ric_syrk_mine(int nkt, int nk)
{
	constexpr int NW = NT / WAVE;
	const int w = uni(wave_id());
	const int mrt = AUG ? (nk + 1 + 15) / 16 : nkt;   // row tiles
	int ntile = 0;
	for (int bj = 0; bj < nkt; bj++)
		ntile += mrt - bj;
	return (ntile - w + NW - 1) / NW;   // tiles w, w + NW, ...
}

template <int NT, int PS, bool FROMG, bool AUG, typename T>
__device__ __forceinline__ void ric_syrk_chunk(RicChunk<T> &ch, int c0, int mine, int nkt, int nk,
	const PM<double, PS> gR)
{
	constexpr int NW = NT / WAVE;
	const int lane = lane_id();
	const int w = uni(wave_id());
	const int mrt = AUG ? (nk + 1 + 15) / 16 : nkt;
	const int nrows = AUG ? nk + 1 : nk;
	ch.cnt = min(4, mine - c0);
#pragma unroll
	for (int t = 0; t < 4; t++)
	{
		int q = w + NW * (c0 + min(t, ch.cnt - 1)), bj = 0;
		while (q >= mrt - bj)
		{
			q -= mrt - bj;
			bj++;
		}
		ch.I[t] = bj + q;
		ch.J[t] = bj;
	}
#pragma unroll
	for (int t = 0; t < 4; t++)
	{
		ch.rq[t] = typename Vec4<T>::type{0, 0, 0, 0};
		if (FROMG && t < ch.cnt)
		{
			const int col = ch.J[t] * 16 + (lane & 15);
#pragma unroll
			for (int r = 0; r < 4; r++)
			{
				const int row = ch.I[t] * 16 + acc_row<T>(lane, r);
				if (row < nrows && col < nk && row >= col)
					ch.rq[t][r] = (T)gR.at(row, col);
			}
		}
	}
}

// M += W W^T on the lower tiles (syrk_ln).  The lower tiles (I, J), I >= J,
// numbered down the block columns, are dealt round-robin to the waves, which
// take theirs four at a time, so the waves' shares differ by one tile at
// most.  FROMG: M = lower(gR) + W W^T instead: the nk x nk lower triangle of
// gR (global memory, zero elsewhere) is loaded into registers of its own
// before the matrix-core steps and added after them, so that its latency
// overlaps them and M needs no separate staging pass.
// AUG: the (nk+1) x nk trapezoid of the fused sweep (one more row of W and
// M, the linear-term row): row tiles of nk+1 rows, column tiles of nk
// columns, and the columns from nk on written as zero (the Cholesky needs its
// padding columns zero; W's row nk is not padding).
// W's columns from n1 on are zero (the trmm writes them so), so the inner
// dimension runs over whole 16-column tiles.
template <int NT, int PS = 4, bool FROMG = false, bool AUG = false, typename T = double>
__device__ __forceinline__ void ric_syrk_m(const TileLowerT<T> M, const TileRectT<T> W, int nkt, int n1,
	const PM<double, PS> gR = PM<double, PS>{}, int nk = 0)
{
	typedef typename Vec4<T>::type V4;
	const int lane = lane_id();
	const int kt = (n1 + 15) / 16;
	const int mine = ric_syrk_mine<NT, PS, FROMG, AUG>(nkt, nk);
	for (int c0 = 0; c0 < mine; c0 += 4)
	{
		RicChunk<T> ch;
		ric_syrk_chunk<NT, PS, FROMG, AUG, T>(ch, c0, mine, nkt, nk, gR);
		const int cnt = ch.cnt;
		V4 acc[4];
#pragma unroll
		for (int t = 0; t < 4; t++)
		{
			acc[t] = V4{0, 0, 0, 0};
			if (!FROMG && t < cnt)
#pragma unroll
				for (int r = 0; r < 4; r++)
					acc[t][r] = M(ch.I[t] * 16 + acc_row<T>(lane, r), ch.J[t] * 16 + (lane & 15));
		}
		if (cnt == 4)
			ric_syrk_tiles<4, T>(W, ch.I, ch.J, kt, acc);
		else if (cnt == 3)
			ric_syrk_tiles<3, T>(W, ch.I, ch.J, kt, acc);
		else if (cnt == 2)
			ric_syrk_tiles<2, T>(W, ch.I, ch.J, kt, acc);
		else
			ric_syrk_tiles<1, T>(W, ch.I, ch.J, kt, acc);
#pragma unroll
		for (int t = 0; t < 4; t++)
		{
			if (t < cnt)
			{
				const int col = ch.J[t] * 16 + (lane & 15);
#pragma unroll
				for (int r = 0; r < 4; r++)
				{
					const T v = FROMG ? acc[t][r] + ch.rq[t][r] : acc[t][r];
					M(ch.I[t] * 16 + acc_row<T>(lane, r), col) = (AUG && col >= nk) ? T(0) : v;
				}
			}
		}
	}
}

// The one-wave kernels call the in-LDS Cholesky out of line: inlined it
// would raise them above the 168 registers that three waves per SIMD allow.
template <int NT>
__device__ __attribute__((noinline)) void ric_trmm_call(const TileLower M, const TileRect W, int nkt, int n1, int n1t, int u1)
{
	ric_trmm_w<NT>(TileLower{lds_opaque(M.p)}, TileRect{lds_opaque(W.p), W.nct}, nkt, n1, n1t, u1);
}

template <int NT>
__device__ __attribute__((noinline)) void ric_syrk_call(const TileLower M, const TileRect W, int nkt, int n1)
{
	ric_syrk_m<NT>(TileLower{lds_opaque(M.p)}, TileRect{lds_opaque(W.p), W.nct}, nkt, n1);
}

template <int NT, int NMAXC>
__device__ __attribute__((noinline)) void ric_chol_call(const TileLower M, int nk, double *invd)
{
	chol_acc<double, NT, NMAXC, TileLower, false>(TileLower{lds_opaque(M.p)}, nk, nk, lds_opaque(invd));
}

// the work matrix of a record: trf's `work`, the fused sweep's `work_mat`
__device__ __forceinline__ blasfeo_dmat *ric_work(const blasfeo_amd_riccati_arg &a) { return a.work; }
__device__ __forceinline__ blasfeo_dmat *ric_work(const blasfeo_amd_riccati_sv_arg &a) { return a.work_mat; }

// AUG = true: the backward part of the fused sweep (d_back_ric_sv_libstr,
// examples/example_d_riccati_recursion.c:58-79): every stage matrix has one
// more row, the linear terms (BAbt[k]'s row nk = b^T, RSQrq[k]'s row nk =
// [r; q]^T), carried through the trmm, the syrk and the trapezoid Cholesky
// (L[k] is (nk+1) x nk); before the syrk the row of the previous factor is
// added to W's last row (the reference's blasfeo_dgead).
// T = float: the mixed-precision form -- the stage matrices and factors are
// held and computed in FP32 (inputs converted on staging, factors written to
// the FP64 outputs), multi-wave kernels only.
template <int PS, int NT, bool AUG = false, typename ArgT = blasfeo_amd_riccati_arg, typename T = double>
__device__ void ric_sweep_lds(const ArgT &a, T *smem)
{
	static_assert(sizeof(T) == 8 || NT != WAVE, "the FP32 sweep runs on two or four waves per instance");
	constexpr int NMAXC = (NT == WAVE) ? 64 : 128;
	constexpr int XR = AUG ? 1 : 0;   // extra row of the augmented form
	const int N = uni(a.N);
	const int *nx = a.nx;
	const int *nu = a.nu;
	blasfeo_dmat *L = a.L;
	blasfeo_dmat *RSQ = a.RSQrq;
	blasfeo_dmat *BAbt = a.BAbt;
	int nkmax = nx[N], nxmax = nx[N];
	for (int k = 0; k < N; k++)
	{
		nkmax = max(nkmax, nu[k] + nx[k]);
		nxmax = max(nxmax, nx[k]);
	}
	nkmax = uni(nkmax);
	nxmax = uni(nxmax);
	const int nkt_max = (nkmax + XR + 15) / 16;
	const int nxt_max = (nxmax + 15) / 16;
	const int nlow = nkt_max * (nkt_max + 1) / 2;
	const TileLowerT<T> M{smem};
	const TileRectT<T> W{smem + (size_t)nlow * RT_SZ, nxt_max};
	T *invd = smem + (size_t)(nlow + nkt_max * nxt_max) * RT_SZ;

	// s = -1 is the last stage, L[N] = chol(Q_N) (potrf_l; potrf_l_mn of the
	// (nx+1) x nx trapezoid when AUG); s >= 0 is stage k = N-1-s.  One call
	// site of the in-LDS Cholesky.  Inlined, it addresses the kernel's dynamic
	// LDS directly; out of line, every LDS access of the callee first loads
	// the kernel's LDS offset from a table and waits for it.
	for (int s = -1; s < N; s++)
	{
		const bool last = s < 0;
		const int k = last ? N : N - 1 - s;
		const int nk = uni(last ? nx[N] : nu[k] + nx[k]);
		const int nkt = (nk + 15) / 16;         // column tiles
		const int mrt = (nk + XR + 15) / 16;    // row tiles
		if (!last)
		{
			const int n1 = uni(nx[k + 1]);
			const int u1 = uni(nu[k + 1]);
			const int n1t = (n1 + 15) / 16;
			// stage BAbt[k] (nk(+1) x n1) into W; M still holds L[k+1].  The
			// multi-wave kernels loaded it during the previous stage's
			// Cholesky (its first panel, on the waves other than 0), so W and
			// M are complete at the Cholesky's barriers; without a barrier
			// here the previous stage's stores of L (which only read M)
			// overlap the trmm (which only reads M and writes W)
			if constexpr (NT == WAVE)
			{
				g2s_acc<double, PS, NT, false>(view<double, PS>(&BAbt[k]), 0, 0, mrt * 16, n1t * 16, nk + XR, n1, W);
				__syncthreads();
			}
			if constexpr (NT == WAVE)
				ric_trmm_call<NT>(M, W, mrt, n1, n1t, u1);
			else
				ric_trmm_w<NT, T>(M, W, mrt, n1, n1t, u1);
			__syncthreads();
			if constexpr (AUG)
			{
				// W(nk, :) += L[k+1](nu1 + nx1, nu1:) (blasfeo_dgead): the linear
				// term's row of the previous factor, still in M
				for (int j = threadIdx.x; j < n1; j += NT)
					W(nk, j) += M(u1 + n1, u1 + j);
				__syncthreads();
			}
		}
		// M <- lower(RSQ[k]) + W W^T (L[k+1] is no longer needed); the
		// multi-wave kernels start the syrk accumulators from RSQ[k] itself
		if (last || NT == WAVE)
			g2s_acc<double, PS, NT, true>(view<double, PS>(&RSQ[k]), 0, 0, mrt * 16, nkt * 16, nk + XR, nk, M);
		if (!last)
		{
			if constexpr (NT == WAVE)
			{
				__syncthreads();
				ric_syrk_call<NT>(M, W, nkt, uni(nx[k + 1]));
			}
			else
				ric_syrk_m<NT, PS, true, AUG, T>(M, W, nkt, uni(nx[k + 1]), view<double, PS>(&RSQ[k]), nk);
		}
		// L[k] = chol(M); W is free (the syrk has read it), so the waves
		// other than 0 load the next stage's BAbt into it meanwhile
		if constexpr (NT == WAVE)
			ric_chol_call<NT, NMAXC>(M, nk, invd);
		else
		{
			const int kn = k - 1;   // next stage (k = N for the last stage: kn = N - 1)
			auto side = [&]() {
				if (kn >= 0)
				{
					const int nkn = uni(nu[kn] + nx[kn]), n1n = uni(nx[kn + 1]);
					g2s_acc<double, PS, NT, false, TileRectT<T>, 1>(view<double, PS>(&BAbt[kn]), 0, 0,
						(nkn + XR + 15) / 16 * 16, (n1n + 15) / 16 * 16, nkn + XR, n1n, W);
				}
			};
			chol_acc<T, NT, NMAXC, TileLowerT<T>, true, decltype(side)>(M, nk + XR, nk, invd, side);
		}
		s2g_acc<double, PS, NT, true>(view<double, PS>(&L[k]), 0, 0, nk + XR, nk, M, 0);
		for (int j = threadIdx.x; j < nk; j += NT)
			L[k].dA[j] = invd[j];
		if (threadIdx.x == 0)
		{
			// potrf_l marks the whole diagonal cached; syrk_dpotrf_ln and the
			// trapezoid forms (potrf_l_mn, syrk_dpotrf_ln_mn) one entry
			L[k].use_dA = (last && !AUG) ? nk : 1;
			if (!last)
				ric_work(a)->use_dA = 0;
		}
	}
	__syncthreads();
}

// ---------------------------------------------------------------------------
// the general sweep: the chain through global memory, a 16-column panel in LDS
// ---------------------------------------------------------------------------
template <int PS>
__device__ void ric_sweep_general(const blasfeo_amd_riccati_arg &a, double *P, int LDP, int mmax,
	double *As, double *Bs, double *L11, double *invs)
{
	const int N = a.N;
	const int *nx = a.nx;
	const int *nu = a.nu;
	blasfeo_dmat *L = a.L;
	blasfeo_dmat *RSQ = a.RSQrq;
	blasfeo_dmat *BAbt = a.BAbt;
	blasfeo_dmat *W = a.work;
	{
		const int m = nx[N];
		PM<double, PS> gL = view<double, PS>(&L[N]);
		chol_instance<double, PS>(P, LDP, L11, invs, m, m, 0, gL, 0, 0, gL, 0, 0,
			view<double, PS>(&RSQ[N]), 0, 0, gL, 0, 0, L[N].dA);
		if (threadIdx.x == 0)
			L[N].use_dA = m;
	}
	for (int s = 0; s < N; s++)
	{
		const int k = N - 1 - s;
		const int nk = nu[k] + nx[k];
		const int nx1 = nx[k + 1];
		const int nu1 = nu[k + 1];
		__syncthreads();
		if (nk > 0 && nx1 > 0)
		{
			PM<double, PS> gL1 = view<double, PS>(&L[k + 1]);
			gemm_core<double, PS, OP_TRMM_RLNN>(As, Bs, nk, nx1, nx1, 1.0, 0.0,
				gL1, nu1, nu1, view<double, PS>(&BAbt[k]), 0, 0,
				gL1, 0, 0, view<double, PS>(W), 0, 0, 0, 1);
		}
		__syncthreads();
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
	__syncthreads();
}

// ---------------------------------------------------------------------------
// plan and launches
// ---------------------------------------------------------------------------
// counts[0..RIC_NCLASS] (class sizes) and counts[RIC_HEAD + c] (the work-queue
// head of class c)
__global__ void k_ric_zero(int *counts)
{
	if (threadIdx.x < 64)
		counts[threadIdx.x] = 0;
}

// The argument record of the sweep: trf's, or the fused sweep's (AUG).
template <bool AUG> struct RicArg { typedef blasfeo_amd_riccati_arg type; };
template <> struct RicArg<true> { typedef blasfeo_amd_riccati_sv_arg type; };

// The plan: one thread per instance takes the size class of its largest
// stage (ric_class, -1 for none) and the workgroup appends to the class
// lists.  Instances with a negative horizon or size are dropped (the
// reference has no defined behaviour for them); those larger than mmax are
// dropped from the general class, which cannot hold them.
// AUG: stage matrices have one more row; no global-memory sweep serves them,
// so instances that fit no LDS class are dropped (the entry refuses size
// bounds that allow them).  T = float: the classes of the FP32 sweep (half
// the bytes per element); it has no global-memory form either.
template <bool AUG = false, typename T = double>
__device__ __forceinline__ int ric_class(const typename RicArg<AUG>::type &a, int mmax)
{
	constexpr int XR = AUG ? 1 : 0;
	constexpr bool NOGEN = AUG || sizeof(T) != 8;
	const int N = a.N;
	if (N < 0)
		return -1;
	int nkmax = a.nx[N], nxmax = a.nx[N];
	bool bad = a.nx[N] < 0;
	for (int k = 0; k < N; k++)
	{
		const int x = a.nx[k], u = a.nu[k];
		bad = bad || x < 0 || u < 0;
		nkmax = max(nkmax, u + x);
		nxmax = max(nxmax, x);
	}
	if (bad)
		return -1;
	int cls = RIC_GENERAL;
	if (nxmax <= 64 && nkmax + XR <= 128)
	{
		// one-wave classes (cap <= 40960) factor with one panel row per lane: nkmax <= 64
		const int f = ric_footprint(nkmax + XR, nxmax, (int)sizeof(T));
		for (int c = RIC_NCLASS - 1; c >= 0; c--)
			if (f <= c_ric_cap[c] && (c_ric_cap[c] > 40960 || nkmax + XR <= 64))
				cls = c;
	}
	if (cls == RIC_GENERAL && (NOGEN || nkmax > mmax))
		return -1;
	return cls;
}

template <bool AUG = false, typename T = double>
__global__ void k_ric_plan(const typename RicArg<AUG>::type *__restrict__ args, int batch, int mmax,
	int *counts, int *lists)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	plan_bucket<RIC_GENERAL + 1>(i, i < batch ? ric_class<AUG, T>(args[i], mmax) : -1, batch, counts, lists);
}

// Next instance of the class for this workgroup from the class's work-queue
// head (one atomic per instance), so that a workgroup that drew short
// instances takes more of them.  Every workgroup leaves once the head passes
// the class size.  Several waves share the value through the LDS slot, which
// is rewritten only after the sweep's last barrier, which every thread passes
// after reading it.
template <int NT>
__device__ __forceinline__ int ric_next(int *head, int *slot)
{
	if constexpr (NT == WAVE)
	{
		int t = 0;
		if (threadIdx.x == 0)
			t = atomicAdd(head, 1);
		return __shfl(t, 0, WAVE);
	}
	else
	{
		if (threadIdx.x == 0)
			*slot = atomicAdd(head, 1);
		__syncthreads();
		return *slot;
	}
}

template <int PS, int NT, bool AUG = false, typename T = double>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void k_ric_class(const typename RicArg<AUG>::type *__restrict__ args,
	const int *__restrict__ count, const int *__restrict__ list, int *head)
{
	static_assert(!(AUG && NT == WAVE), "the fused sweep runs on two or four waves per instance");
	extern __shared__ double smem_d[];
	int *slot = reinterpret_cast<int *>(smem_d);
	T *work = reinterpret_cast<T *>(smem_d) + RIC_QSLOT / sizeof(T);
	if (threadIdx.x == 0)
		work[-1] = T(0);   // the trmm's zero operand (ric_trmm_w)
	const int cnt = *count;
	for (int t = ric_next<NT>(head, slot); t < cnt; t = ric_next<NT>(head, slot))
		ric_sweep_lds<PS, NT, AUG, typename RicArg<AUG>::type, T>(args[list[t]], work);
}

// one wave per instance (selected by the tuning knobs, not by default):
// registers capped at 168 so that three instances share each SIMD.  Only
// classes of at most 40960 B can use it, whose stages have nu+nx <= 64 (the
// plan's rule), so the Cholesky keeps one panel row per lane.
template <int PS>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(3))) void k_ric_class1(
	const blasfeo_amd_riccati_arg *__restrict__ args, const int *__restrict__ count, const int *__restrict__ list,
	int *head)
{
	extern __shared__ double smem_d[];
	double *work = smem_d + RIC_QSLOT / sizeof(double);
	if (threadIdx.x == 0)
		work[-1] = 0.0;   // the trmm's zero operand (ric_trmm_w)
	const int cnt = *count;
	for (int t = ric_next<WAVE>(head, nullptr); t < cnt; t = ric_next<WAVE>(head, nullptr))
		ric_sweep_lds<PS, WAVE>(args[list[t]], work);
}

template <int PS>
__global__ __launch_bounds__(NT3) void k_ric_general(const blasfeo_amd_riccati_arg *__restrict__ args,
	const int *__restrict__ count, const int *__restrict__ list, int mmax)
{
	extern __shared__ unsigned char smem_raw[];
	__shared__ double As[KC * LDT];
	__shared__ double Bs[KC * LDT];
	__shared__ double L11[PW * (PW + 1)];
	__shared__ double invs[PW];
	double *P = reinterpret_cast<double *>(smem_raw);
	const int LDP = (mmax + 15) / 16 * 16 + 1;
	const int cnt = *count;
	for (int t = blockIdx.x; t < cnt; t += gridDim.x)
		ric_sweep_general<PS>(args[list[t]], P, LDP, mmax, As, Bs, L11, invs);
}

template <int PS, int NT, bool AUG = false, typename T = double>
static int launch_class(int cls, int grid, const typename RicArg<AUG>::type *args, int *counts, const int *lists,
	int batch, hipStream_t st, const char *name)
{
	if constexpr (AUG || sizeof(T) != 8)
	{
		static_assert(NT != WAVE, "the fused and FP32 sweeps run on two or four waves per instance");
		if (const int rc = lds_check(reinterpret_cast<const void *>(&k_ric_class<PS, NT, AUG, T>), h_ric_cap[cls], name))
			return rc;
		hipLaunchKernelGGL((k_ric_class<PS, NT, AUG, T>), dim3(grid), dim3(NT), h_ric_cap[cls], st, args,
			counts + cls, lists + (size_t)cls * batch, counts + RIC_HEAD + cls);
		return check_launch(name);
	}
	else
	{
		const void *fn = NT == WAVE ? reinterpret_cast<const void *>(&k_ric_class1<PS>)
		                            : reinterpret_cast<const void *>(&k_ric_class<PS, NT>);
		if (const int rc = lds_check(fn, h_ric_cap[cls], name))
			return rc;
		if (NT == WAVE)
			hipLaunchKernelGGL((k_ric_class1<PS>), dim3(grid), dim3(WAVE), h_ric_cap[cls], st, args,
				counts + cls, lists + (size_t)cls * batch, counts + RIC_HEAD + cls);
		else
			hipLaunchKernelGGL((k_ric_class<PS, NT>), dim3(grid), dim3(NT), h_ric_cap[cls], st, args,
				counts + cls, lists + (size_t)cls * batch, counts + RIC_HEAD + cls);
		return check_launch(name);
	}
}

// Helper streams (default priority) on which the class launches run side by
// side with the caller's stream: forked from and joined back to it with
// events, so the call stays ordered on the caller's stream, and a capture of
// that stream captures the helpers' launches as parallel branches.  Created
// per device on the first call (outside any capture), kept for the process.
constexpr int RIC_AUX = 3;
struct RicFork
{
	hipStream_t aux[RIC_AUX];
	hipEvent_t fork;
	hipEvent_t join[RIC_AUX];
	std::mutex mu;
};

static RicFork *ric_fork_state()
{
	static std::mutex gmu;
	static std::map<int, RicFork *> per_dev;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess)
		return nullptr;
	std::lock_guard<std::mutex> lock(gmu);
	auto it = per_dev.find(dev);
	if (it != per_dev.end())
		return it->second;
	RicFork *f = new RicFork;
	bool ok = hipEventCreateWithFlags(&f->fork, hipEventDisableTiming) == hipSuccess;
	for (int i = 0; i < RIC_AUX && ok; i++)
		ok = hipStreamCreateWithFlags(&f->aux[i], hipStreamNonBlocking) == hipSuccess &&
			hipEventCreateWithFlags(&f->join[i], hipEventDisableTiming) == hipSuccess;
	if (!ok)
		return nullptr;
	per_dev[dev] = f;
	return f;
}

// ws: the plan workspace, (RIC_NCLASS + 1) * batch + 64 ints.  fork: run the
// class launches side by side on helper streams (else all on st, in order).
template <int PS, bool AUG = false, typename T = double>
static int launch_ric(int batch, const typename RicArg<AUG>::type *args, int mmax, int maxfoot, int *ws, bool fork,
	hipStream_t st, const char *name)
{
	int *counts = ws;
	int *lists = ws + 64;
	hipLaunchKernelGGL(k_ric_zero, dim3(1), dim3(64), 0, st, counts);
	hipLaunchKernelGGL((k_ric_plan<AUG, T>), dim3((batch + 255) / 256), dim3(256), 0, st, args, batch, mmax, counts, lists);
	int rc = check_launch(name);
	if (rc)
		return rc;
	const int cus = num_cus() > 0 ? num_cus() : 256;
	RicFork *fk = nullptr;
	std::unique_lock<std::mutex> lock;
	hipStream_t ss[1 + RIC_AUX] = {st, st, st, st};
	if (fork)
	{
		fk = ric_fork_state();
		if (!fk)
			return set_error(BLASFEO_AMD_ELAUNCH, "%s: cannot create the helper streams", name);
		lock = std::unique_lock<std::mutex>(fk->mu);
		if (hipEventRecord(fk->fork, st) != hipSuccess)
			return check_launch(name);
		for (int i = 0; i < RIC_AUX; i++)
		{
			ss[1 + i] = fk->aux[i];
			if (hipStreamWaitEvent(fk->aux[i], fk->fork, 0) != hipSuccess)
				return check_launch(name);
		}
	}
	// Launch schedule: (class, stream) in launch order; stream 0 is the
	// caller's, 1..3 the helpers.  The general launch and the 40, 16 and
	// 32 KB classes run one after another on the caller's stream; the 160,
	// 80 and 54 KB classes start at once on the helpers, the 24 and 8 KB
	// classes after the 80 KB one.  Measured on the 8192 x N=20 batch (nx ~
	// U{8..64}, nu ~ U{2..16}) against other orders on MI355X: 5.52 ms
	// replayed in a graph and 5.49 ms eager, against 5.85 and 5.96 ms for
	// largest-first dealt round-robin over the four streams (the 16 and
	// 32 KB classes started only after the 80 and 54 KB ones and ran alone
	// at the end; an empty class launch holds its stream while its
	// workgroups wait for LDS).
	static const int sched[RIC_NCLASS + 1][2] = {
		{RIC_GENERAL, 0}, {7, 1}, {6, 2}, {5, 3}, {4, 0}, {1, 0}, {3, 0}, {2, 2}, {0, 2}};
	static_assert(RIC_NCLASS == 8 && RIC_AUX == 3, "the launch schedule lists 8 classes on 4 streams");
	(void)maxfoot;
	for (int e = 0; e <= RIC_NCLASS && !rc; e++)
	{
		const int cls = sched[e][0];
		hipStream_t sc = ss[sched[e][1]];
		if (cls == RIC_GENERAL)
		{
			if constexpr (!AUG && sizeof(T) == 8)
			{
				const size_t shm = sizeof(double) * PW * ((mmax + 15) / 16 * 16 + 1);
				const int grid = min(batch, cus);
				rc = lds_check(reinterpret_cast<const void *>(&k_ric_general<PS>), shm, name);
				if (!rc)
				{
					hipLaunchKernelGGL((k_ric_general<PS>), dim3(grid), dim3(NT3), shm, sc, args,
						counts + RIC_GENERAL, lists + (size_t)RIC_GENERAL * batch, mmax);
					rc = check_launch(name);
				}
			}
			continue;
		}
		// built-in launch shape per class (threads per instance, resident
		// instances per CU), from autotuner runs on the 8192 x N=20 batch
		// (nx ~ U{8..64}, nu ~ U{2..16}) on MI355X: two waves per instance up
		// to 24 KB (4 per CU up to 8 KB, 8 above: fewer resident instances
		// spread a small class over more CUs), four waves and 3 per CU at
		// 32 KB, four waves and 2 per CU above.  The tuning knobs override
		// both.
		int per_cu, nt;
		if (h_ric_cap[cls] <= 24576)
		{
			nt = 128;
			per_cu = h_ric_cap[cls] <= 8192 ? 4 : min(RIC_LDS_MAX / h_ric_cap[cls], 8);
		}
		else if (h_ric_cap[cls] == 32768)
		{
			nt = 256;
			per_cu = 3;
		}
		else
		{
			nt = 256;
			per_cu = max(1, min(2, RIC_LDS_MAX / h_ric_cap[cls]));
		}
		if constexpr (sizeof(T) != 8)
		{
			// FP32 sweep: as many resident instances as LDS and three waves
			// per SIMD allow
			per_cu = min(nt == 128 ? 6 : 3, RIC_LDS_MAX / h_ric_cap[cls]);
		}
		if (const int v = tune_value(TUNE_RIC_PER_CU, cls))
			per_cu = v;
		if (const int v = tune_value(TUNE_RIC_THREADS, cls))
			nt = v;
		const int grid = min(batch, cus * per_cu);
		constexpr bool ONEWAVE = !AUG && sizeof(T) == 8;
		if (nt == 64 && ONEWAVE)
			rc = launch_class<PS, ONEWAVE ? 64 : 128, AUG, T>(cls, grid, args, counts, lists, batch, sc, name);
		else if (nt <= 128)
			rc = launch_class<PS, 128, AUG, T>(cls, grid, args, counts, lists, batch, sc, name);
		else
			rc = launch_class<PS, 256, AUG, T>(cls, grid, args, counts, lists, batch, sc, name);
		if (rc)
			break;
	}
	// join back whatever was launched (also after a launch error, so that
	// the caller's stream never runs ahead of the helpers)
	for (int i = 0; fork && i < RIC_AUX; i++)
		if (hipEventRecord(fk->join[i], fk->aux[i]) != hipSuccess ||
			hipStreamWaitEvent(st, fk->join[i], 0) != hipSuccess)
			return check_launch(name);
	return rc ? rc : check_launch(name);
}

// ---------------------------------------------------------------------------
// The allocation-free entry: the plan's counts, queue heads and class lists
// live in a file-scope device array instead of the per-stream workspace, and
// the class launches run in order on the caller's stream (no helper
// streams), so that a call creates nothing and allocates nothing.
// ---------------------------------------------------------------------------
constexpr int RIC_NOALLOC_MAX = 65536;   // most instances per call
static __device__ __align__(256) int scratch[64 + (RIC_NCLASS + 1) * RIC_NOALLOC_MAX];

// Device address of this device's copy of `scratch` (each GPU holds its own),
// looked up once per device with that device current.
static int *noalloc_scratch()
{
	constexpr int MAXDEV = 64;
	static std::mutex mu;
	static int *ptr[MAXDEV];
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAXDEV)
		return nullptr;
	std::lock_guard<std::mutex> lock(mu);
	if (!ptr[dev])
	{
		void *p = nullptr;
		if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess)
		{
			(void)hipGetLastError();
			return nullptr;
		}
		ptr[dev] = static_cast<int *>(p);
	}
	return ptr[dev];
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
	const int nxmax = hint->max_n > 0 ? hint->max_n : mmax;
	const int maxfoot = ric_footprint(mmax, min(nxmax, mmax));
	hipStream_t st = (hipStream_t)stream;
	const size_t wsbytes = sizeof(int) * ((size_t)(RIC_NCLASS + 1) * batch + 64);
	int *ws = (int *)workspace(wsbytes, st);
	if (!ws)
		return set_error(BLASFEO_AMD_ENOMEM, "%s: cannot allocate the %zu-byte plan workspace", name, wsbytes);
	return ps == 4 ? launch_ric<4>(batch, args, mmax, maxfoot, ws, true, st, name)
	               : launch_ric<8>(batch, args, mmax, maxfoot, ws, true, st, name);
}

extern "C" int blasfeo_amd_d_back_ric_trf_batched_noalloc(int ps, int batch, const blasfeo_amd_riccati_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream)
{
	const char *name = "blasfeo_amd_d_back_ric_trf_batched_noalloc";
	if (batch < 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: batch=%d < 0", name, batch);
	if (batch == 0)
		return BLASFEO_AMD_OK;
	if (batch > RIC_NOALLOC_MAX)
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: at most %d instances per call (its fixed scratch), got %d",
			name, RIC_NOALLOC_MAX, batch);
	if (!args)
		return set_error(BLASFEO_AMD_EINVAL, "%s: args is NULL", name);
	if (ps != 4 && ps != 8)
		return set_error(BLASFEO_AMD_EINVAL, "%s: unsupported panel size %d", name, ps);
	if (!hint || hint->max_m <= 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: hint->max_m (largest nu[k]+nx[k]) is required", name);
	const int mmax = hint->max_m;
	if (mmax > fact_mmax<double>())
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: nu+nx up to %d supported, got %d", name, fact_mmax<double>(), mmax);
	const int nxmax = hint->max_n > 0 ? hint->max_n : mmax;
	const int maxfoot = ric_footprint(mmax, min(nxmax, mmax));
	int *ws = noalloc_scratch();
	if (!ws)
		return set_error(BLASFEO_AMD_ELAUNCH, "%s: no device address for the scratch array", name);
	hipStream_t st = (hipStream_t)stream;
	return ps == 4 ? launch_ric<4>(batch, args, mmax, maxfoot, ws, false, st, name)
	               : launch_ric<8>(batch, args, mmax, maxfoot, ws, false, st, name);
}

// ---------------------------------------------------------------------------
// The fused sweep's forward substitution (d_back_ric_sv_libstr,
// examples/example_d_riccati_recursion.c:81-107), one wavefront per instance,
// k = 0 .. N-1 (nk = nu[k] + nx[k], ns = nk on stage 0 and nu[k] after it):
//   ux[k](0:ns) = -L[k](nk, 0:ns)                         (the factor's last row)
//   ux[k] = trsv_ltn_mn(nk, ns) of L[k] on ux[k]            (ux[k](ns:) from stage k-1)
//   ux[k+1](nu1:) = BAbt[k](nk, :)^T + BAbt[k]^T ux[k]      (gemv_t; the row nk is b)
//   pi[k] = Lxx(k+1) (Lxx(k+1)^T ux[k+1](nu1:) + L[k+1](nk1, nu1:)^T)
// The substitution follows the Riccati solve's (bfa_ric_trs.hip): lanes own
// rows, the column sums are wave reductions, one column's coefficients and
// reciprocal loaded ahead of the chain; rows up to 128 (two per lane).
// ---------------------------------------------------------------------------
namespace bfa {

constexpr int SV_MAX = 2 * WAVE;

template <int PS>
__global__ __launch_bounds__(WAVE) void k_ric_sv_fwd(const blasfeo_amd_riccati_sv_arg *__restrict__ args, int batch)
{
	__shared__ double s_x[SV_MAX];
	__shared__ double s_t[SV_MAX];
	const int inst = blockIdx.x;
	const int lane = threadIdx.x;
	if (inst >= batch)
		return;
	const blasfeo_amd_riccati_sv_arg &a = args[inst];
	const int N = a.N;
	for (int k = 0; k < N; k++)
	{
		const int nu0 = a.nu[k], n0 = nu0 + a.nx[k];
		const int nu1 = a.nu[k + 1], nx1 = a.nx[k + 1], nk1 = nu1 + nx1;
		const blasfeo_dmat &Lk = a.L[k];
		const blasfeo_dmat &Lk1 = a.L[k + 1];
		const blasfeo_dmat &Bk = a.BAbt[k];
		const PM<double, PS> L0{Lk.pA, Lk.cn}, L1{Lk1.pA, Lk1.cn}, B{Bk.pA, Bk.cn};
		double *ux0 = a.ux[k].pa;
		double *ux1 = a.ux[k + 1].pa;
		double *pi = a.pi[k].pa;
		const int ns = k == 0 ? n0 : nu0;
		double z[2];
#pragma unroll
		for (int q = 0; q < 2; q++)
		{
			const int i = lane + q * WAVE;
			z[q] = i < ns ? -L0.at(n0, i) : (i < n0 ? ux0[i] : 0.0);
		}
		// z(j) = (z(j) - sum_{i > j} L(i, j) z(i)) / L(j, j), j = ns-1 .. 0
		const bool cached = Lk.use_dA == 1;
		const double *dA = Lk.dA;
		auto coef = [&](int j, double (&e)[2]) {
#pragma unroll
			for (int q = 0; q < 2; q++)
			{
				const int i = lane + q * WAVE;
				e[q] = (i > j && i < n0) ? L0.at(i, j) : 0.0;
			}
		};
		auto recip = [&](int j) { return cached ? dA[j] : 1.0 / L0.at(j, j); };
		double e[2] = {0.0, 0.0}, d = 0.0;
		if (ns > 0)
		{
			coef(ns - 1, e);
			d = recip(ns - 1);
		}
		for (int j = ns - 1; j >= 0; j--)
		{
			double en[2] = {0.0, 0.0}, dn = 0.0;
			const bool more = j > 0;
			if (more)
			{
				coef(j - 1, en);
				dn = recip(j - 1);
			}
			const int i0 = lane, i1 = lane + WAVE;
			const double part = (i0 > j && i0 < n0 ? e[0] * z[0] : 0.0) + (i1 > j && i1 < n0 ? e[1] * z[1] : 0.0);
			const double sum = readlane(wave_sum_d(part), WAVE - 1);
			if (!cached && lane == 0)
				Lk.dA[j] = d;
#pragma unroll
			for (int q = 0; q < 2; q++)
				if (lane + q * WAVE == j)
					z[q] = (z[q] - sum) * d;
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
		// x(k+1) = b + BAbt^T z (gemv_t), one column per lane; pi starts as it
		for (int j = lane; j < nx1; j += WAVE)
		{
			double acc = 0.0;
#pragma unroll 8
			for (int i = 0; i < n0; i++)
				acc += B.at(i, j) * s_x[i];
			const double y = B.at(n0, j) + acc;
			ux1[nu1 + j] = y;
			s_t[j] = y;
		}
		__syncthreads();
		// t = Lxx^T x + L[k+1](nk1, nu1:)^T (trmv_ltn, axpy), one column per lane
		double tv[2];
#pragma unroll
		for (int q = 0; q < 2; q++)
		{
			const int j = lane + q * WAVE;
			double acc = 0.0;
			if (j < nx1)
			{
				for (int i = j; i < nx1; i++)
					acc += L1.at(nu1 + i, nu1 + j) * s_t[i];
				acc += L1.at(nk1, nu1 + j);
			}
			tv[q] = acc;
		}
		__syncthreads();
#pragma unroll
		for (int q = 0; q < 2; q++)
		{
			const int j = lane + q * WAVE;
			if (j < nx1)
				s_x[j] = tv[q];
		}
		__syncthreads();
		// pi = Lxx t (trmv_lnn), one row per lane
#pragma unroll
		for (int q = 0; q < 2; q++)
		{
			const int i = lane + q * WAVE;
			if (i < nx1)
			{
				double acc = 0.0;
				for (int j = 0; j <= i; j++)
					acc += L1.at(nu1 + i, nu1 + j) * s_x[j];
				pi[i] = acc;
			}
		}
		__syncthreads();
	}
}

} // namespace bfa

extern "C" int blasfeo_amd_d_back_ric_sv_batched(int ps, int batch, const blasfeo_amd_riccati_sv_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream)
{
	using namespace bfa;
	const char *name = "blasfeo_amd_d_back_ric_sv_batched";
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
	const int nxmax = hint->max_n > 0 ? hint->max_n : mmax;
	if (mmax + 1 > 128 || nxmax > 64)
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: nu+nx up to 127 and nx up to 64 supported, got %d and %d",
			name, mmax, nxmax);
	hipStream_t st = (hipStream_t)stream;
	const size_t wsbytes = sizeof(int) * ((size_t)(RIC_NCLASS + 1) * batch + 64);
	int *ws = (int *)workspace(wsbytes, st);
	if (!ws)
		return set_error(BLASFEO_AMD_ENOMEM, "%s: cannot allocate the %zu-byte plan workspace", name, wsbytes);
	const int maxfoot = ric_footprint(mmax + 1, min(nxmax, mmax));
	int rc = ps == 4 ? launch_ric<4, true>(batch, args, mmax, maxfoot, ws, true, st, name)
	                 : launch_ric<8, true>(batch, args, mmax, maxfoot, ws, true, st, name);
	if (rc)
		return rc;
	if (ps == 4)
		hipLaunchKernelGGL(k_ric_sv_fwd<4>, dim3(batch), dim3(WAVE), 0, st, args, batch);
	else
		hipLaunchKernelGGL(k_ric_sv_fwd<8>, dim3(batch), dim3(WAVE), 0, st, args, batch);
	return check_launch(name);
}

// ---------------------------------------------------------------------------
// Mixed-precision Riccati solve: FP32 factorisation, solve on those factors,
// refinement with FP64 residuals of the KKT equations, FP64 fallback.
// ---------------------------------------------------------------------------
namespace bfa {

// the argument records of the steps, from the caller's records
__global__ void k_mix_setup(const blasfeo_amd_riccati_mixed_arg *__restrict__ ma, int batch,
	blasfeo_amd_riccati_arg *ra, blasfeo_amd_riccati_trs_arg *tm, blasfeo_amd_riccati_trs_arg *tc)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= batch)
		return;
	const blasfeo_amd_riccati_mixed_arg a = ma[i];
	ra[i] = blasfeo_amd_riccati_arg{a.N, a.nx, a.nu, a.BAbt, a.RSQrq, a.L, a.work_mat};
	tm[i] = blasfeo_amd_riccati_trs_arg{a.N, a.nx, a.nu, a.BAbt, a.b, a.rq, a.L, a.Pb, a.ux, a.pi, a.work_vec};
	tc[i] = blasfeo_amd_riccati_trs_arg{a.N, a.nx, a.nu, a.BAbt, a.res_b, a.res_rq, a.L, a.d_Pb, a.d_ux, a.d_pi,
		a.work_vec};
	if (a.info)
		*a.info = blasfeo_amd_solve_info{0, 0, 0.0};
}

// max over the wave of v >= 0, NaN counted as +inf, in every lane
__device__ __forceinline__ double mix_wmax(double v)
{
	v = (v == v) ? v : __builtin_huge_val();
	return readlane(wave_max_d(v), WAVE - 1);
}

// One wave per instance still open (tc[i].N >= 0): mode 1 first applies the
// correction (ux += d_ux, pi += d_pi); then the KKT residuals of ux, pi are
// written to res_rq / res_b and the normwise backward error to info->berr;
// an instance at or below tol is closed (tc[i].N = -1), so the next
// correction solve and pass skip it.  mode 2 (after the FP64 fallback)
// only records the backward error.
template <int PS>
__global__ __launch_bounds__(WAVE) void k_mix_residual(const blasfeo_amd_riccati_mixed_arg *__restrict__ ma,
	blasfeo_amd_riccati_trs_arg *tc, int batch, double tol, int mode)
{
	const int i = blockIdx.x;
	if (i >= batch || tc[i].N < 0)
		return;
	const blasfeo_amd_riccati_mixed_arg &a = ma[i];
	const int lane = lane_id();
	const int N = uni(a.N);
	if (mode == 1)
	{
		for (int k = 0; k <= N; k++)
		{
			const int nk = uni(a.nu[k] + a.nx[k]);
			const int nx1 = k < N ? uni(a.nx[k + 1]) : 0;
			for (int j = lane; j < nk; j += WAVE)
				a.ux[k].pa[j] += a.d_ux[k].pa[j];
			for (int j = lane; j < nx1; j += WAVE)
				a.pi[k].pa[j] += a.d_pi[k].pa[j];
		}
		if (lane == 0 && a.info)
			a.info->iters++;
	}
	double berr = 0.0;
	for (int k = 0; k <= N; k++)
	{
		const int nuk = uni(a.nu[k]), nk = uni(a.nu[k] + a.nx[k]);
		const int nx1 = k < N ? uni(a.nx[k + 1]) : 0;
		const double *z = a.ux[k].pa;
		const double *g = a.rq[k].pa;
		const double *p1 = k < N ? a.pi[k].pa : nullptr;
		const double *p0 = k > 0 ? a.pi[k - 1].pa : nullptr;
		const PM<double, PS> H = view<double, PS>(&a.RSQrq[k]);
		double rmax = 0.0, hmax = 0.0, cmax = 0.0;
		for (int r = lane; r < nk; r += WAVE)
		{
			double acc = g[r], hs = 0.0, cs = 0.0;
			for (int j = 0; j < nk; j++)
			{
				const double h = j <= r ? H.at(r, j) : H.at(j, r);
				acc += h * z[j];
				hs += fabs(h);
			}
			if (k < N)
			{
				const PM<double, PS> C = view<double, PS>(&a.BAbt[k]);
				for (int j = 0; j < nx1; j++)
				{
					const double c = C.at(r, j);
					acc += c * p1[j];
					cs += fabs(c);
				}
			}
			if (k > 0 && r >= nuk)
				acc -= p0[r - nuk];
			a.res_rq[k].pa[r] = acc;
			rmax = fmax(rmax, (acc == acc) ? fabs(acc) : __builtin_huge_val());
			hmax = fmax(hmax, hs);
			cmax = fmax(cmax, cs);
		}
		double zmax = 0.0, gmax = 0.0, p1max = 0.0, p0max = 0.0;
		for (int j = lane; j < nk; j += WAVE)
		{
			zmax = fmax(zmax, fabs(z[j]));
			gmax = fmax(gmax, fabs(g[j]));
		}
		for (int j = lane; j < nx1; j += WAVE)
			p1max = fmax(p1max, fabs(p1[j]));
		if (k > 0)
			for (int j = lane; j < nk - nuk; j += WAVE)
				p0max = fmax(p0max, fabs(p0[j]));
		rmax = mix_wmax(rmax);
		const double nz = mix_wmax(zmax);
		double den = mix_wmax(hmax) * nz + mix_wmax(gmax) + mix_wmax(cmax) * mix_wmax(p1max) + mix_wmax(p0max);
		berr = fmax(berr, rmax > 0.0 ? rmax / den : 0.0);
		if (k < N)
		{
			// dynamics: s = BAbt^T ux_k + b_k - x_{k+1}
			const PM<double, PS> C = view<double, PS>(&a.BAbt[k]);
			const double *b = a.b[k].pa;
			const int nu1 = uni(a.nu[k + 1]);
			const double *x1 = a.ux[k + 1].pa + nu1;
			double smax = 0.0, ctmax = 0.0, bmax = 0.0, xmax = 0.0;
			for (int c = lane; c < nx1; c += WAVE)
			{
				double acc = b[c] - x1[c], cs = 0.0;
				for (int r = 0; r < nk; r++)
				{
					const double v = C.at(r, c);
					acc += v * z[r];
					cs += fabs(v);
				}
				a.res_b[k].pa[c] = acc;
				smax = fmax(smax, (acc == acc) ? fabs(acc) : __builtin_huge_val());
				ctmax = fmax(ctmax, cs);
				bmax = fmax(bmax, fabs(b[c]));
				xmax = fmax(xmax, fabs(x1[c]));
			}
			smax = mix_wmax(smax);
			den = mix_wmax(ctmax) * nz + mix_wmax(bmax) + mix_wmax(xmax);
			berr = fmax(berr, smax > 0.0 ? smax / den : 0.0);
		}
	}
	if (lane == 0)
	{
		if (a.info)
			a.info->berr = berr;
		if (mode == 2 || berr <= tol)
			tc[i].N = -1;
	}
}

// the FP64 fallback: instances still open keep their N in ra / tm (and are
// marked path 1), the others get N = -1 so the FP64 sweep and solve skip them
__global__ void k_mix_fallback(const blasfeo_amd_riccati_mixed_arg *__restrict__ ma, int batch,
	blasfeo_amd_riccati_arg *ra, blasfeo_amd_riccati_trs_arg *tm, blasfeo_amd_riccati_trs_arg *tc)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= batch)
		return;
	const bool open = tc[i].N >= 0;
	ra[i].N = open ? ma[i].N : -1;
	tm[i].N = open ? ma[i].N : -1;
	if (open && ma[i].info)
		ma[i].info->path = 1;
}

} // namespace bfa

extern "C" int blasfeo_amd_d_ric_solve_mixed_batched(int ps, int batch, const blasfeo_amd_riccati_mixed_arg *args,
	double tol, int maxit, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream)
{
	using namespace bfa;
	const char *name = "blasfeo_amd_d_ric_solve_mixed_batched";
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
	if (maxit < 0 || !(tol >= 0.0))
		return set_error(BLASFEO_AMD_EINVAL, "%s: maxit=%d, tol=%g", name, maxit, tol);
	const int mmax = hint->max_m;
	const int nxmax = hint->max_n > 0 ? hint->max_n : mmax;
	if (mmax > 128 || nxmax > 64)
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: nu+nx up to 128 and nx up to 64 supported, got %d and %d",
			name, mmax, nxmax);
	hipStream_t st = (hipStream_t)stream;
	// plan workspace, then the three argument arrays
	const size_t plan = sizeof(int) * ((size_t)(RIC_NCLASS + 1) * batch + 64);
	const size_t off_ra = (plan + 255) / 256 * 256;
	const size_t off_tm = off_ra + (sizeof(blasfeo_amd_riccati_arg) * batch + 255) / 256 * 256;
	const size_t off_tc = off_tm + (sizeof(blasfeo_amd_riccati_trs_arg) * batch + 255) / 256 * 256;
	const size_t wsbytes = off_tc + sizeof(blasfeo_amd_riccati_trs_arg) * batch;
	char *ws = (char *)workspace(wsbytes, st);
	if (!ws)
		return set_error(BLASFEO_AMD_ENOMEM, "%s: cannot allocate the %zu-byte workspace", name, wsbytes);
	int *pws = (int *)ws;
	auto *ra = (blasfeo_amd_riccati_arg *)(ws + off_ra);
	auto *tm = (blasfeo_amd_riccati_trs_arg *)(ws + off_tm);
	auto *tc = (blasfeo_amd_riccati_trs_arg *)(ws + off_tc);
	const dim3 g256((batch + 255) / 256);
	hipLaunchKernelGGL(k_mix_setup, g256, dim3(256), 0, st, args, batch, ra, tm, tc);
	int rc = check_launch(name);
	// FP32 factorisation, solve on it
	const int foot32 = ric_footprint(mmax, min(nxmax, mmax), 4);
	if (!rc)
		rc = ps == 4 ? launch_ric<4, false, float>(batch, ra, mmax, foot32, pws, true, st, name)
		             : launch_ric<8, false, float>(batch, ra, mmax, foot32, pws, true, st, name);
	blasfeo_amd_hint th{mmax, nxmax, 0};
	if (!rc)
		rc = blasfeo_amd_d_back_ric_trs_batched(ps, batch, tm, &th, stream);
	auto residual = [&](int mode) {
		if (ps == 4)
			hipLaunchKernelGGL(k_mix_residual<4>, dim3(batch), dim3(WAVE), 0, st, args, tc, batch, tol, mode);
		else
			hipLaunchKernelGGL(k_mix_residual<8>, dim3(batch), dim3(WAVE), 0, st, args, tc, batch, tol, mode);
		return check_launch(name);
	};
	if (!rc)
		rc = residual(0);
	// corrections on the FP32 factors
	for (int it = 0; it < maxit && !rc; it++)
	{
		rc = blasfeo_amd_d_back_ric_trs_batched(ps, batch, tc, &th, stream);
		if (!rc)
			rc = residual(1);
	}
	// FP64 for whatever is still open
	if (!rc)
	{
		hipLaunchKernelGGL(k_mix_fallback, g256, dim3(256), 0, st, args, batch, ra, tm, tc);
		rc = check_launch(name);
	}
	const int foot64 = ric_footprint(mmax, min(nxmax, mmax));
	if (!rc)
		rc = ps == 4 ? launch_ric<4>(batch, ra, mmax, foot64, pws, true, st, name)
		             : launch_ric<8>(batch, ra, mmax, foot64, pws, true, st, name);
	if (!rc)
		rc = blasfeo_amd_d_back_ric_trs_batched(ps, batch, tm, &th, stream);
	if (!rc)
		rc = residual(2);
	return rc;
}
