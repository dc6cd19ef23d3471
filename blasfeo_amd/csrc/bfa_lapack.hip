// bfa_lapack.hip -- batched factorisations for gfx950: potrf_l, potrf_l_mn,
// syrk_dpotrf_ln(_mn) and getrf_rp on panel-major matrices.
//
// One workgroup (4 waves) per instance.  Both factorisations are left-looking
// over 16-column panels held in LDS at full height (the role of the
// reference's 12x4 / 8x8 register blocks, blasfeo_hp_pm/d_lapack_lib4.c:954-985
// and :2443-2634):
//   potrf: panel <- C(+A*B^T) - L(:,0:c0) L(c0:c0+16,0:c0)^T on the matrix
//          cores, then one wave factors the top 64 rows in registers (the
//          diagonal 16x16 Cholesky, broadcasting L(jj,j) with v_readlane) and
//          all waves solve the remaining rows against it;
//   getrf: panel <- rows gathered through the current row permutation, the
//          previous unit-lower panels applied block by block (unit solve of
//          the 16-row block, matrix-core update of the rows below), then one
//          wave factors the panel column by column with a wave-wide argmax
//          for the pivot (first index of the largest |x|, as idamax and
//          blasfeo_ref/x_lapack_ref.c:1275-1302), and the panel's swaps are
//          applied to the finished columns on the left.
#include "bfa_chol_core.hpp"
#include "bfa_internal.hpp"

namespace bfa {

template <typename T> struct FactArg;
template <> struct FactArg<double> { typedef blasfeo_amd_dfact_arg type; };
template <> struct FactArg<float> { typedef blasfeo_amd_sfact_arg type; };

template <typename T> struct SPArg;
template <> struct SPArg<double> { typedef blasfeo_amd_dsyrk_dpotrf_arg type; };

// potrf_l for n > 256: left-looking over 16-column LDS panels at full height,
// the earlier panels read back from D (chol_instance)
template <typename T, int PS>
__global__ __launch_bounds__(NTL) void k_potrf(const typename FactArg<T>::type *__restrict__ args, int mmax)
{
	extern __shared__ unsigned char smem_raw[];
	__shared__ T L11[PW * (PW + 1)];
	__shared__ T invs[PW];
	T *P = reinterpret_cast<T *>(smem_raw);
	const typename FactArg<T>::type &a = args[blockIdx.x];
	const int m = a.m;
	if (m <= 0 || m > mmax)
		return;
	const int LDP = (mmax + 15) / 16 * 16 + 1;
	PM<T, PS> gD = view<T, PS>(a.sD);
	chol_instance<T, PS>(P, LDP, L11, invs, m, m, 0, gD, 0, 0, gD, 0, 0,
		view<T, PS>(a.sC), a.ci, a.cj, gD, a.di, a.dj, a.sD->dA);
	if (threadIdx.x == 0)
		a.sD->use_dA = (a.di == 0 && a.dj == 0) ? m : 0;
}

// The trapezoid's LDS form: square column-major with leading dimension np + 1
// (the one-instance-per-workgroup kernels), or packed lower 16x17 tiles (the
// bucketed launches: 78 KB instead of 133 KB at 128 rows, two instances per
// CU instead of one).
template <bool PACKED, typename T>
__device__ __forceinline__ auto lds_acc(T *M, int ld)
{
	if constexpr (PACKED)
		return TileLowerT<T>{M};
	else
		return ColMajor<T>{M, ld};
}
// LDS elements of the trapezoid plus the reciprocal diagonal for np rows
__host__ __device__ __forceinline__ size_t lds_elems(int np, bool packed)
{
	const size_t t = (size_t)(np / 16);
	return (packed ? t * (t + 1) / 2 * RT_SZ : (size_t)np * (np + 1)) + np;
}

// potrf_l (MN = false) and potrf_l_mn (MN = true) with the whole m x n lower
// trapezoid in LDS (m <= 128): stage it from C, factor in place (chol_acc),
// stream L and dA out.  NT = 64 (one wave) for m <= 32 so that several
// instances share a CU.
template <typename T, int PS, int NT, bool MN, bool PACKED>
__device__ __forceinline__ void potrf_lds_instance(const typename FactArg<T>::type &a, int np, T *M, T *invd)
{
	const int LDM = np + 1;
	const int m = a.m;
	int n = MN ? a.n : a.m;
	if (m <= 0 || n <= 0 || m > np)
		return;
	if (n > m)
		n = m;
	const int mt = (m + 15) / 16, nt = (n + 15) / 16;
	const auto Ml = lds_acc<PACKED>(M, LDM);
	if constexpr (PACKED)
		g2s_acc<T, PS, NT, true>(view<T, PS>(a.sC), a.ci, a.cj, mt * 16, nt * 16, m, n, Ml);
	else
		g2s<T, PS, NT, false, true>(view<T, PS>(a.sC), a.ci, a.cj, mt * 16, nt * 16, m, n, M, LDM);
	// the one-wave form serves m <= 32: one panel row per lane, no second row set
	chol_acc<T, NT, NT == WAVE ? 64 : 128>(Ml, m, n, invd);
	if constexpr (PACKED)
		s2g_acc<T, PS, NT, true>(view<T, PS>(a.sD), a.di, a.dj, m, n, Ml, 0);
	else
		s2g<T, PS, NT, true>(view<T, PS>(a.sD), a.di, a.dj, m, n, M, LDM, 0);
	T *dD = a.sD->dA;
	for (int j = threadIdx.x; j < n; j += NT)
		dD[j] = invd[j];
	if (threadIdx.x == 0)
		a.sD->use_dA = (a.di == 0 && a.dj == 0) ? (MN ? 1 : m) : 0;
}

// The LDS-resident factorisations run either one instance per workgroup
// (k_potrf_lds, k_syrk_potrf_lds: workgroup b takes instance b) or as a
// persistent launch over one size bucket (the _q forms), each workgroup
// taking the next instance of the bucket's list from its atomic queue head.
// Two kernels rather than one with both modes: the loop costs registers that
// the one-instance form (several instances per CU at 32 rows) cannot spare.
// The queued forms are compiled for three waves per SIMD (at most 168
// registers; the syrk form then fits in 149 VGPRs without spills, against
// 145 VGPRs + 40 AGPRs and two waves per SIMD unconstrained), so that three
// four-wave workgroups share a CU where the LDS holds them (up to 96 rows).  The leading barrier makes every
// thread finish the previous instance (and its read of the slot) before the
// slot is overwritten and the LDS staged again.  Returns -1 once the bucket
// is drained.
__device__ __forceinline__ int queue_next(const int *count, const int *list, int *head, int *qslot)
{
	__syncthreads();
	if (threadIdx.x == 0)
		*qslot = atomicAdd(head, 1);
	__syncthreads();
	const int t = *qslot;
	return t < *count ? list[t] : -1;
}

template <typename T, int PS, int NT, bool MN>
__global__ __launch_bounds__(NT) void k_potrf_lds(const typename FactArg<T>::type *__restrict__ args, int np)
{
	extern __shared__ unsigned char smem_raw[];
	T *M = reinterpret_cast<T *>(smem_raw);
	potrf_lds_instance<T, PS, NT, MN, false>(args[blockIdx.x], np, M, M + lds_elems(np, false) - np);
}
template <typename T, int PS, int NT, bool MN>
__global__ __launch_bounds__(NT, 3) void k_potrf_lds_q(const typename FactArg<T>::type *__restrict__ args, int np,
	const int *__restrict__ count, const int *__restrict__ list, int *head)
{
	extern __shared__ unsigned char smem_raw[];
	__shared__ int qslot;
	T *M = reinterpret_cast<T *>(smem_raw);
	for (int i; (i = queue_next(count, list, head, &qslot)) >= 0;)
		potrf_lds_instance<T, PS, NT, MN, true>(args[i], np, M, M + lds_elems(np, true) - np);
}

// syrk_dpotrf_ln(_mn), m <= 128: D = chol(C + A B^T) on the m x n lower
// trapezoid (blasfeo_hp_pm/d_lapack_lib4.c:1848-2077).  C's trapezoid is staged
// in LDS, A B^T is added tile by tile on the matrix cores with A and B read
// from global memory (a group of up to four row tiles shares B's operand), and
// the sum is factored in place (chol_acc).
template <typename T, int PS, int NT, bool PACKED>
__device__ __forceinline__ void syrk_potrf_lds_instance(const typename SPArg<T>::type &a, int np, T *M, T *invd)
{
	typedef typename Vec4<T>::type V4;
	constexpr int NW = NT / WAVE;
	const int LDM = np + 1;
	const int m = a.m, k = a.k;
	int n = a.n;
	if (m <= 0 || n <= 0 || m > np || k < 0)
		return;
	if (n > m)
		n = m;
	const int lane = lane_id(), w = wave_id();
	const int mt = (m + 15) / 16, nt = (n + 15) / 16;
	const auto Ml = lds_acc<PACKED>(M, LDM);
	if constexpr (PACKED)
		g2s_acc<T, PS, NT, true>(view<T, PS>(a.sC), a.ci, a.cj, mt * 16, nt * 16, m, n, Ml);
	else
		g2s<T, PS, NT, false, true>(view<T, PS>(a.sC), a.ci, a.cj, mt * 16, nt * 16, m, n, M, LDM);
	__syncthreads();
	if (k > 0)
	{
		const PM<T, PS> gA = view<T, PS>(a.sA), gB = view<T, PS>(a.sB);
		int ntask = 0;
		for (int bj = 0; bj < nt; bj++)
			ntask += (mt - bj + 3) / 4;
		for (int task = w; task < ntask; task += NW)
		{
			int bj = 0, rem = task;
			while (rem >= (mt - bj + 3) / 4)
			{
				rem -= (mt - bj + 3) / 4;
				bj++;
			}
			const int g = bj + 4 * rem;
			V4 acc[4];
#pragma unroll
			for (int t = 0; t < 4; t++)
				acc[t] = V4{0, 0, 0, 0};
			const int jr = bj * 16 + (lane & 15);     // row of B for this lane's operand
			// this lane's rows of A (one per tile) and of B at column 0 of k; a
			// panel-major row advances by PS elements per column
			const BFA_GLOBAL T *rb = jr < n ? &gB.at(a.bi + jr, a.bj) : nullptr;
			const BFA_GLOBAL T *ra[4];
#pragma unroll
			for (int t = 0; t < 4; t++)
			{
				const int ir = (g + t) * 16 + (lane & 15);
				ra[t] = (g + t < mt && ir < m) ? &gA.at(a.ai + ir, a.aj) : nullptr;
			}
			// KU matrix-core steps per round, all their global loads issued
			// before the first step: one memory latency per 4 KU columns of k
			// rather than per 4 (KU = 2 costs no registers at four waves; 3
			// and 4 raise the kernel from 140 to 218 and 250 VGPRs)
			constexpr int KU = 2;
#pragma unroll 1
			for (int kk = 0; kk < k; kk += 4 * KU)
			{
				T bv[KU], av[4][KU];
#pragma unroll
				for (int u = 0; u < KU; u++)
				{
					const int kq = kk + 4 * u + (lane >> 4);
					const bool kok = kq < k;
					bv[u] = (kok && rb) ? rb[(size_t)kq * PS] : T(0);
#pragma unroll
					for (int t = 0; t < 4; t++)
						av[t][u] = (kok && ra[t]) ? ra[t][(size_t)kq * PS] : T(0);
				}
#pragma unroll
				for (int u = 0; u < KU; u++)
					if (kk + 4 * u < k)
#pragma unroll
						for (int t = 0; t < 4; t++)
							if (g + t < mt)
								acc[t] = mfma16(av[t][u], bv[u], acc[t]);
			}
#pragma unroll
			for (int t = 0; t < 4; t++)
				if (g + t < mt)
#pragma unroll
					for (int r = 0; r < 4; r++)
					{
						const int ir = (g + t) * 16 + acc_row<T>(lane, r), jc = bj * 16 + (lane & 15);
						if (ir < m && jc < n && ir >= jc)
							Ml(ir, jc) += acc[t][r];
					}
		}
	}
	chol_acc<T, NT, 128>(Ml, m, n, invd);
	if constexpr (PACKED)
		s2g_acc<T, PS, NT, true>(view<T, PS>(a.sD), a.di, a.dj, m, n, Ml, 0);
	else
		s2g<T, PS, NT, true>(view<T, PS>(a.sD), a.di, a.dj, m, n, M, LDM, 0);
	T *dD = a.sD->dA;
	for (int j = threadIdx.x; j < n; j += NT)
		dD[j] = invd[j];
	if (threadIdx.x == 0)
		a.sD->use_dA = (a.di == 0 && a.dj == 0) ? 1 : 0;
}

template <typename T, int PS, int NT>
__global__ __launch_bounds__(NT) void k_syrk_potrf_lds(const typename SPArg<T>::type *__restrict__ args, int np)
{
	extern __shared__ unsigned char smem_raw[];
	T *M = reinterpret_cast<T *>(smem_raw);
	syrk_potrf_lds_instance<T, PS, NT, false>(args[blockIdx.x], np, M, M + lds_elems(np, false) - np);
}
template <typename T, int PS, int NT>
__global__ __launch_bounds__(NT, 3) void k_syrk_potrf_lds_q(const typename SPArg<T>::type *__restrict__ args, int np,
	const int *__restrict__ count, const int *__restrict__ list, int *head)
{
	extern __shared__ unsigned char smem_raw[];
	__shared__ int qslot;
	T *M = reinterpret_cast<T *>(smem_raw);
	for (int i; (i = queue_next(count, list, head, &qslot)) >= 0;)
		syrk_potrf_lds_instance<T, PS, NT, true>(args[i], np, M, M + lds_elems(np, true) - np);
}

// ---------------------------------------------------------------------------
// potrf_l for 128 < n <= 256: left-looking over block panels of 16 BW columns
// held in LDS as 16x16 tiles (col-major inside a tile, tile (r, jj) of the
// panel at (BW r + jj) * 256), one workgroup of NTH threads per instance: BW =
// 4 (128 KB, one instance per CU, 8 waves), 2 (64 KB, two per CU, 4 waves
// each) or 1 (32 KB, four per CU): more resident instances overlap their
// serial chains, at the cost of reading the earlier columns more often.
//   1. the panel's lower part is staged from C (identity beyond n);
//   2. the earlier block columns are applied on the matrix cores,
//      P(r, jj) -= L(rows r, 0:c0) L(rows jj, 0:c0)^T, with L read from D
//      (one wave per tile row, the A operand shared by the row's BW tiles);
//   3. the panel is factored right-looking in LDS: wave 0 factors the
//      diagonal tile in registers (as chol_lds), every thread then solves one
//      row below it against L_jj with the reciprocal diagonal, and the
//      panel's remaining tile columns get a rank-16 matrix-core update;
//   4. L and the reciprocal diagonal are written to D.
// Compared with the 16-column left-looking form this reads the earlier
// columns once per 16 BW columns instead of once per 16.
// ---------------------------------------------------------------------------
constexpr int PT_NMAX = 256;

template <typename T, int BW>
__host__ __device__ inline size_t potrf_tiles_lds_bytes()
{
	return sizeof(T) * (16 * BW * 256 + 16 * 17 + 16);
}

template <typename T, int PS, int BW, int NTH>
__global__ __launch_bounds__(NTH) void k_potrf_tiles(const typename FactArg<T>::type *__restrict__ args, int nmax)
{
	constexpr int PT_NT = NTH;
	constexpr int PT_NW = NTH / WAVE;
	typedef typename Vec4<T>::type V4;
	extern __shared__ unsigned char smem_raw[];
	T *P = reinterpret_cast<T *>(smem_raw);
	T *L11 = P + 16 * BW * 256;
	T *invs = L11 + 16 * 17;
	const typename FactArg<T>::type &a = args[blockIdx.x];
	const int n = a.m;
	if (n <= 0 || n > nmax)
		return;
	const int lane = lane_id();
	const int w = wave_id();
	const int nb = (n + 15) / 16;
	const PM<T, PS> gC = view<T, PS>(a.sC);
	PM<T, PS> gD = view<T, PS>(a.sD);
	const int ci = a.ci, cj = a.cj, di = a.di, dj = a.dj;
	T *dD = a.sD->dA;

	for (int J0 = 0; J0 < nb; J0 += BW)
	{
		const int R = nb - J0;           // tile rows of the panel
		const int C = min(BW, R);        // tile columns of the panel
		const int c0 = 16 * J0;          // first column of the panel
		__syncthreads();
		// 1. stage the panel (lower part; identity beyond n), four tiles per
		// wave at a time: their 16 loads per lane are issued before any store
		for (int t0 = w; t0 < R * BW; t0 += 4 * PT_NW)
		{
			T v[4][4];
#pragma unroll
			for (int u = 0; u < 4; u++)
			{
				const int t = t0 + u * PT_NW;
				const int r = t / BW, jj = t % BW;
				const int i = lane & 15;
				const int gr = c0 + 16 * r + i;
#pragma unroll
				for (int k = 0; k < 4; k++)
				{
					const int j = 4 * k + (lane >> 4);
					const int gc = c0 + 16 * jj + j;
					const bool live = t < R * BW && jj < C && r >= jj;
					v[u][k] = (live && gr < n && gc < n && gr >= gc) ? gC.at(ci + gr, cj + gc)
					                                                  : ((gr == gc) ? T(1) : T(0));
				}
			}
#pragma unroll
			for (int u = 0; u < 4; u++)
			{
				const int t = t0 + u * PT_NW;
				const int r = t / BW, jj = t % BW;
				if (t >= R * BW || jj >= C || r < jj)
					continue;
				T *PT = P + t * 256;
				const int i = lane & 15;
#pragma unroll
				for (int k = 0; k < 4; k++)
					PT[(4 * k + (lane >> 4)) * 16 + i] = v[u][k];
			}
		}
		__syncthreads();
		// 2. apply the earlier block columns (L from D)
		if (J0 > 0)
		{
			for (int r = w; r < R; r += PT_NW)
			{
				V4 acc[BW];
#pragma unroll
				for (int jj = 0; jj < BW; jj++)
				{
					acc[jj] = V4{0, 0, 0, 0};
					if (jj < C && jj <= r)
#pragma unroll
						for (int q = 0; q < 4; q++)
							acc[jj][q] = P[(r * BW + jj) * 256 + (lane & 15) * 16 + acc_row<T>(lane, q)];
				}
				const int ar = c0 + 16 * r + (lane & 15);
				const bool aok = ar < n;

				// 16 columns of L per step (c0 is a multiple of 16 BW): all of a
				// step's loads are issued before its products, so a wave waits
				// for memory once per 16 columns instead of once per 4
				for (int kk = 0; kk < c0; kk += 16)
				{
					T av[4], bv[4][BW];
#pragma unroll
					for (int u = 0; u < 4; u++)
					{
						const int kq = kk + 4 * u + (lane >> 4);
						av[u] = aok ? gD.at(di + ar, dj + kq) : T(0);
#pragma unroll
						for (int jj = 0; jj < BW; jj++)
						{
							const int br = c0 + 16 * jj + (lane & 15);
							bv[u][jj] = (jj < C && jj <= r && br < n) ? gD.at(di + br, dj + kq) : T(0);
						}
					}
#pragma unroll
					for (int u = 0; u < 4; u++)
#pragma unroll
						for (int jj = 0; jj < BW; jj++)
							if (jj < C && jj <= r)
								acc[jj] = mfma16(-av[u], bv[u][jj], acc[jj]);
				}
#pragma unroll
				for (int jj = 0; jj < BW; jj++)
					if (jj < C && jj <= r)
#pragma unroll
						for (int q = 0; q < 4; q++)
							P[(r * BW + jj) * 256 + (lane & 15) * 16 + acc_row<T>(lane, q)] = acc[jj][q];
			}
			__syncthreads();
		}
		// 3. factor the panel
		for (int jj = 0; jj < C; jj++)
		{
			T *D = P + (jj * BW + jj) * 256;
			if (w == 0)
			{
				T x[16];
#pragma unroll
				for (int j = 0; j < 16; j++)
					x[j] = lane < 16 ? D[j * 16 + lane] : T(0);
				// no per-lane guards: the upper triangle of the tile (lanes < j)
				// is updated too, but only the lower triangle is read (the
				// row solve, L11 below its diagonal) or written to D.  Lane j
				// holds the pivot c, so scaling every lane gives it c * inv;
				// every lane stores the uniform reciprocal (no per-column lane
				// masks, which held SGPRs across the tile)
#pragma unroll
				for (int j = 0; j < 16; j++)
				{
					const T c = readlane(x[j], j);
					const T inv = (c > T(0)) ? rsqrt_fast(c) : T(0);
					x[j] *= inv;
					invs[j] = inv;
#pragma unroll
					for (int jn = j + 1; jn < 16; jn++)
					{
						const T l = readlane(x[j], jn);
						x[jn] -= x[j] * l;
					}
				}
				if (lane < 16)
#pragma unroll
					for (int j = 0; j < 16; j++)
					{
						D[j * 16 + lane] = x[j];
						L11[lane * 17 + j] = x[j];
					}
			}
			__syncthreads();
			// rows below the diagonal tile: x_j = (x_j - sum_{l<j} x_l L_jl) * inv_j
			{
				const int nrow = (R - jj - 1) * 16;
				for (int q = threadIdx.x; q < nrow; q += PT_NT)
				{
					asm volatile("" ::: "memory");
					const int r = jj + 1 + (q >> 4), i = q & 15;
					T *row = P + (r * BW + jj) * 256 + i;
					T x[16];
#pragma unroll
					for (int j = 0; j < 16; j++)
						x[j] = row[j * 16];
#pragma unroll
					for (int j = 0; j < 16; j++)
					{
						T c = x[j];
#pragma unroll
						for (int l = 0; l < j; l++)
							c -= x[l] * L11[j * 17 + l];
						x[j] = c * invs[j];
					}
#pragma unroll
					for (int j = 0; j < 16; j++)
						row[j * 16] = x[j];
				}
			}
			__syncthreads();
			// rank-16 update of the panel's later tile columns
			{
				const int ntask = (C - jj - 1) * R;
				for (int t = w; t < ntask; t += PT_NW)
				{
					const int j2 = jj + 1 + t / R, r = t % R;
					if (r < j2)
						continue;
					V4 acc;
#pragma unroll
					for (int q = 0; q < 4; q++)
						acc[q] = P[(r * BW + j2) * 256 + (lane & 15) * 16 + acc_row<T>(lane, q)];
#pragma unroll
					for (int kk = 0; kk < 16; kk += 4)
					{
						const int o = (kk + (lane >> 4)) * 16 + (lane & 15);
						acc = mfma16(-P[(r * BW + jj) * 256 + o], P[(j2 * BW + jj) * 256 + o], acc);
					}
#pragma unroll
					for (int q = 0; q < 4; q++)
						P[(r * BW + j2) * 256 + (lane & 15) * 16 + acc_row<T>(lane, q)] = acc[q];
				}
			}
			if (threadIdx.x < 16 && c0 + 16 * jj + threadIdx.x < n)
				dD[c0 + 16 * jj + threadIdx.x] = invs[threadIdx.x];
			__syncthreads();
		}
		// 4. write L (lower part, rows and columns < n)
		for (int t = w; t < R * BW; t += PT_NW)
		{
			const int r = t / BW, jj = t % BW;
			if (jj >= C || r < jj)
				continue;
			const T *PT = P + t * 256;
			const int i = lane & 15;
			const int gr = c0 + 16 * r + i;
#pragma unroll
			for (int k = 0; k < 4; k++)
			{
				const int j = 4 * k + (lane >> 4);
				const int gc = c0 + 16 * jj + j;
				if (gr < n && gc < n && gr >= gc)
					gD.at(di + gr, dj + gc) = PT[j * 16 + i];
			}
		}
	}
	if (threadIdx.x == 0)
		a.sD->use_dA = (a.di == 0 && a.dj == 0) ? n : 0;
}


// ---------------------------------------------------------------------------
// LU with partial row pivoting (blasfeo_ref/x_lapack_ref.c:1156-1536 semantics);
// PIV = false: LU without pivoting, getrf_np (x_lapack_ref.c:880-1150): no
// pivot search, no swaps, ipiv not touched, and dA = 1/U(j,j) with no guard
// for a zero pivot (the reference divides unconditionally)
// ---------------------------------------------------------------------------
// Launch order: the plan kernel sorts the instances into buckets of 16 rows
// of max(m, n), largest first, and workgroup b of the launch takes the b-th
// instance of that order (counts/lists == nullptr: the input order), so the
// largest instances start first and the launch does not end on one of them
// running alone.
constexpr int GETRF_NB = 16;
template <typename T>
__global__ void k_getrf_plan(const typename FactArg<T>::type *__restrict__ args, int batch, int *counts, int *lists)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	int b = -1;
	if (i < batch)
	{
		const int s = max(args[i].m, args[i].n);
		b = s <= 0 ? GETRF_NB - 1 : max(0, GETRF_NB - 1 - (s - 1) / 16);
	}
	plan_bucket<GETRF_NB>(i, b, batch, counts, lists);
}

__global__ void k_zero_ints(int *p, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = 0;
}

__device__ __forceinline__ int getrf_order(int t, int batch, const int *counts, const int *lists)
{
	if (!counts)
		return t;
	for (int b = 0; b < GETRF_NB; b++)
	{
		const int c = counts[b];
		if (t < c)
			return lists[(size_t)b * batch + t];
		t -= c;
	}
	return -1;
}

template <typename T, int PS, bool PIV = true>
__global__ __launch_bounds__(NTL) __attribute__((amdgpu_waves_per_eu(4))) void k_getrf(const typename FactArg<T>::type *__restrict__ args, int mmax,
	int batch = 0, const int *__restrict__ counts = nullptr, const int *__restrict__ lists = nullptr)
{
	typedef typename Vec4<T>::type V4;
	extern __shared__ unsigned char smem_raw[];
	__shared__ T invs[PW];
	__shared__ int pivs[PW];
	__shared__ T Lb[16 * 17];   // the unit-lower block L(r0:r0+16, r0:r0+16) of the solve
	__shared__ T wbest[NTL / WAVE];   // per-wave pivot candidates
	__shared__ int widx[NTL / WAVE];
	const int LDP = (mmax + 15) / 16 * 16 + 1;
	T *P = reinterpret_cast<T *>(smem_raw);                              // [PW][LDP]
	int *perm = reinterpret_cast<int *>(smem_raw + sizeof(T) * PW * LDP);  // [mmax]

	const int inst = getrf_order(blockIdx.x, batch, counts, lists);
	if (inst < 0)
		return;
	const typename FactArg<T>::type &a = args[inst];
	const int m = a.m, n = a.n;
	if (threadIdx.x == 0)
		a.sD->use_dA = (a.di == 0 && a.dj == 0) ? 1 : 0;
	if (m <= 0 || n <= 0 || m > mmax)
		return;
	const PM<T, PS> gC = view<T, PS>(a.sC);
	PM<T, PS> gD = view<T, PS>(a.sD);
	const int ci = a.ci, cj = a.cj, di = a.di, dj = a.dj;
	T *dD = a.sD->dA;
	int *ipiv = a.ipiv;
	const int p = min(m, n);
	const int lane = lane_id();
	const int w = wave_id();
	const int mt = (m + 15) / 16;

	for (int i = threadIdx.x; i < m; i += NTL)
		perm[i] = i;

	for (int c0 = 0; c0 < n; c0 += PW)
	{
		const int cw = min(PW, n - c0);
		__syncthreads();
		// gather the panel through the current row permutation, eight
		// elements per thread at a time (all loads issued before the stores)
		for (int e0 = threadIdx.x; e0 < PW * mt * 16; e0 += 8 * NTL)
		{
			T v[8];
#pragma unroll
			for (int u = 0; u < 8; u++)
			{
				const int e = e0 + u * NTL;
				const int c = e / (mt * 16);
				const int i = e - c * (mt * 16);
				v[u] = (e < PW * mt * 16 && i < m && c < cw) ? gC.at(ci + perm[i], cj + c0 + c) : T(0);
			}
#pragma unroll
			for (int u = 0; u < 8; u++)
			{
				const int e = e0 + u * NTL;
				const int c = e / (mt * 16);
				const int i = e - c * (mt * 16);
				if (e < PW * mt * 16)
					P[c * LDP + i] = v[u];
			}
		}
		// apply the finished unit-lower panels, 16 rows at a time
		const int rbmax = min(c0, p);
		for (int r0 = 0; r0 < rbmax; r0 += 16)
		{
			const int rh = min(16, rbmax - r0);
			// stage the strictly lower part of L(r0:r0+rh, r0:r0+rh), one
			// element per thread, so that the solve reads it from LDS
			if (threadIdx.x < 256)
			{
				const int r = threadIdx.x >> 4, l = threadIdx.x & 15;
				if (l < r && r < rh)
					Lb[r * 17 + l] = gD.at(di + r0 + r, dj + r0 + l);
			}
			__syncthreads();
			// unit-lower solve of rows r0..r0+rh: one lane per panel column
			if (w == 0 && lane < cw)
			{
				T x[16];
#pragma unroll
				for (int r = 0; r < 16; r++)
					x[r] = P[lane * LDP + r0 + r];
#pragma unroll
				for (int r = 1; r < 16; r++)
				{
					if (r < rh)
					{
						T c = x[r];
#pragma unroll
						for (int l = 0; l < r; l++)
							c -= Lb[r * 17 + l] * x[l];
						x[r] = c;
					}
				}
#pragma unroll
				for (int r = 0; r < 16; r++)
					P[lane * LDP + r0 + r] = x[r];
			}
			__syncthreads();
			// rows below: P(r0+16.., :) -= L(r0+16.., r0:r0+16) * P(r0:r0+16, :)
			const int rs = r0 + 16;
			const int ntile = (m - rs + 15) / 16;
			// the right operand (the solved rows r0..r0+16) is the same for
			// every tile: read once
			T bv[4];
#pragma unroll
			for (int u = 0; u < 4; u++)
			{
				const int kq = 4 * u + (lane >> 4);
				bv[u] = kq < rh ? P[(lane & 15) * LDP + r0 + kq] : T(0);
			}
			// up to four of the wave's row tiles at a time: all their loads
			// are issued before any product
			for (int rt0 = w; rt0 < ntile; rt0 += 4 * (NTL / WAVE))
			{
				T av[4][4];
#pragma unroll
				for (int t = 0; t < 4; t++)
				{
					const int rt = rt0 + t * (NTL / WAVE);
					const int ir = rs + rt * 16 + (lane & 15);
					const bool iok = rt < ntile && ir < m;
#pragma unroll
					for (int u = 0; u < 4; u++)
					{
						const int kq = 4 * u + (lane >> 4);
						av[t][u] = (iok && kq < rh) ? gD.at(di + ir, dj + r0 + kq) : T(0);
					}
				}
#pragma unroll
				for (int t = 0; t < 4; t++)
				{
					const int rt = rt0 + t * (NTL / WAVE);
					if (rt >= ntile)
						break;
					V4 acc = V4{0, 0, 0, 0};
#pragma unroll
					for (int u = 0; u < 4; u++)
						acc = mfma16(av[t][u], bv[u], acc);
#pragma unroll
					for (int r = 0; r < 4; r++)
					{
						const int il = rs + rt * 16 + acc_row<T>(lane, r);
						P[(lane & 15) * LDP + il] -= acc[r];
					}
				}
			}
		}
		__syncthreads();
		// factor the panel rows c0..m-1 (only columns that have a pivot row)
		const int cf = min(cw, p - c0);
		if (cf > 0 && m - c0 > WAVE)
		{
			// all waves, rows striped over the threads: per column a wave
			// argmax, the waves' winners combined through LDS (the same
			// total order, so the same pivot), the swap, then the rank-1
			// update of the rows below; three barriers per column
			for (int j = 0; j < cf; j++)
			{
				const int jr = c0 + j;
				int ip = jr;
				if (PIV)
				{
				T best = T(0);
				int bi = jr;
				for (int i = jr + threadIdx.x; i < m; i += NTL)
				{
					const T v = P[j * LDP + i];
					const T av = v > T(0) ? v : -v;
					if (av > best)
					{
						best = av;
						bi = i;
					}
				}
				wave_argmax_abs_to_last<T>(best, bi);
				if (lane == WAVE - 1)
				{
					wbest[w] = best;
					widx[w] = bi;
				}
				__syncthreads();
				T b = wbest[0];
				ip = widx[0];
				for (int ww = 1; ww < NTL / WAVE; ww++)
					if (wbest[ww] > b || (wbest[ww] == b && widx[ww] < ip))
					{
						b = wbest[ww];
						ip = widx[ww];
					}
				if (threadIdx.x == 0)
					pivs[j] = ip;
				}
				if (PIV && ip != jr)
				{
					if (threadIdx.x < PW)
					{
						const T t0 = P[threadIdx.x * LDP + jr];
						P[threadIdx.x * LDP + jr] = P[threadIdx.x * LDP + ip];
						P[threadIdx.x * LDP + ip] = t0;
					}
					if (threadIdx.x == 0)
					{
						const int t1 = perm[jr];
						perm[jr] = perm[ip];
						perm[ip] = t1;
					}
				}
				__syncthreads();
				const T piv = P[j * LDP + jr];
				const T inv = PIV ? (piv != T(0) ? T(1) / piv : T(0)) : T(1) / piv;
				if (threadIdx.x == 0)
					invs[j] = inv;
				for (int i = jr + 1 + threadIdx.x; i < m; i += NTL)
				{
					const T l = P[j * LDP + i] * inv;
					P[j * LDP + i] = l;
					for (int jj = j + 1; jj < cw; jj++)
						P[jj * LDP + i] -= l * P[jj * LDP + jr];
				}
				__syncthreads();
			}
		}
		else if (w == 0 && cf > 0)
		{
			for (int j = 0; j < cf; j++)
			{
				const int jr = c0 + j;
				int ip = jr;
				if (PIV)
				{
				// pivot: first row of largest |x| among jr..m-1, from (0, jr)
				T best = T(0);
				int bi = jr;
				for (int i = jr + lane; i < m; i += WAVE)
				{
					const T v = P[j * LDP + i];
					const T av = v > T(0) ? v : -v;
					if (av > best)
					{
						best = av;
						bi = i;
					}
				}
				ip = wave_argmax_abs<T>(best, bi);
				if (lane == 0)
					pivs[j] = ip;
				}
				if (PIV && ip != jr)
				{
					if (lane < PW)
					{
						const T t0 = P[lane * LDP + jr];
						P[lane * LDP + jr] = P[lane * LDP + ip];
						P[lane * LDP + ip] = t0;
					}
					if (lane == 0)
					{
						const int t1 = perm[jr];
						perm[jr] = perm[ip];
						perm[ip] = t1;
					}
				}
				wave_sync();
				const T piv = P[j * LDP + jr];
				const T inv = PIV ? (piv != T(0) ? T(1) / piv : T(0)) : T(1) / piv;
				if (lane == 0)
					invs[j] = inv;
				for (int i = jr + 1 + lane; i < m; i += WAVE)
				{
					const T l = P[j * LDP + i] * inv;
					P[j * LDP + i] = l;
					for (int jj = j + 1; jj < cw; jj++)
						P[jj * LDP + i] -= l * P[jj * LDP + jr];
				}
				wave_sync();
			}
		}
		__syncthreads();
		// write the panel (all rows, current order) and its pivot data
		s2g<T, PS, NTL, false>(gD, di, dj + c0, m, cw, P, LDP, 0);
		for (int j = threadIdx.x; j < cf; j += NTL)
		{
			if (PIV)
				ipiv[c0 + j] = pivs[j];
			dD[c0 + j] = invs[j];
		}
		__syncthreads();
		// this panel's swaps on the finished columns 0..c0-1 (one thread per column)
		for (int c = threadIdx.x; PIV && c < c0; c += NTL)
		{
			for (int j = 0; j < cf; j++)
			{
				const int r1 = c0 + j, r2 = pivs[j];
				if (r1 != r2)
				{
					auto &u = gD.at(di + r1, dj + c);
					auto &v = gD.at(di + r2, dj + c);
					const T t = u;
					u = v;
					v = t;
				}
			}
		}
	}
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
template <typename T>
static int fact_mmax_arg(const blasfeo_amd_hint *hint, const char *name, int *mmax)
{
	int mm = (hint && hint->max_m > 0) ? hint->max_m : 256;
	if (mm > fact_mmax<T>())
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: m up to %d supported, got %d", name, fact_mmax<T>(), mm);
	*mmax = mm;
	return BLASFEO_AMD_OK;
}

static int fact_check(int ps, int batch, const void *args, bool is_float, const char *name)
{
	if (batch < 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: batch=%d < 0", name, batch);
	if (batch > 0 && !args)
		return set_error(BLASFEO_AMD_EINVAL, "%s: args is NULL", name);
	const bool ok = is_float ? (ps == 4 || ps == 8 || ps == 16) : (ps == 4 || ps == 8);
	if (!ok)
		return set_error(BLASFEO_AMD_EINVAL, "%s: unsupported panel size %d", name, ps);
	return BLASFEO_AMD_OK;
}

// ---------------------------------------------------------------------------
// potrf_l_mn and syrk_dpotrf_ln(_mn) for m > 128: the m x n lower trapezoid
// as the n x n Cholesky of its top block (potrf_l, whose kernels store the
// reciprocal diagonal in D's dA) followed by the rows below,
// L21 = C21 L11^{-T}, solved with those reciprocals (trsm_rltn reading dA);
// syrk_dpotrf first forms C + A B^T in D (syrk_ln_mn) and then factors D in
// place.  The per-instance records of the steps are built on the device into
// the stream's workspace, so the call stays asynchronous.  This is the
// reference's own decomposition of the trapezoid
// (blasfeo_ref/x_lapack_ref.c:185-326: column j of the rows below is the
// pivot column's update times 1/L_jj).
// ---------------------------------------------------------------------------
constexpr int MN_NMAX = 256;

__global__ void k_mn_split(const blasfeo_amd_dfact_arg *__restrict__ args, int batch,
	blasfeo_amd_dfact_arg *top, blasfeo_amd_dtrsm_arg *rows)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= batch)
		return;
	const blasfeo_amd_dfact_arg a = args[i];
	const int m = a.m;
	const int n = a.n < m ? a.n : m;
	blasfeo_amd_dfact_arg t = a;
	blasfeo_amd_dtrsm_arg r;
	const bool live = m > 0 && n > 0;
	t.m = t.n = live ? n : 0;
	t.ipiv = nullptr;
	r.m = live ? m - n : 0;
	r.n = n;
	r.ai = a.di; r.aj = a.dj;
	r.bi = a.ci + n; r.bj = a.cj;
	r.di = a.di + n; r.dj = a.dj;
	r.alpha = 1.0;
	r.sA = a.sD; r.sB = a.sC; r.sD = a.sD;
	top[i] = t;
	rows[i] = r;
}

// use_dA of the trapezoid factorisations (blasfeo_hp_pm/d_lapack_lib4.c:1185,1545,1882)
__global__ void k_mn_finish(const blasfeo_amd_dfact_arg *__restrict__ args, int batch)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= batch)
		return;
	const blasfeo_amd_dfact_arg &a = args[i];
	if (a.m > 0 && a.n > 0)
		a.sD->use_dA = (a.di == 0 && a.dj == 0) ? 1 : 0;
}

// syrk_dpotrf records -> syrk_ln_mn records (D = C + A B^T on the trapezoid)
// and trapezoid Cholesky records (in place on D)
__global__ void k_sp_split(const blasfeo_amd_dsyrk_dpotrf_arg *__restrict__ args, int batch,
	blasfeo_amd_dgemm_arg *upd, blasfeo_amd_dfact_arg *fact)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= batch)
		return;
	const blasfeo_amd_dsyrk_dpotrf_arg a = args[i];
	const int m = a.m;
	const int n = a.n < m ? a.n : m;
	const bool live = m > 0 && n > 0 && a.k >= 0;
	blasfeo_amd_dgemm_arg g;
	g.m = live ? m : 0; g.n = n; g.k = a.k;
	g.ai = a.ai; g.aj = a.aj; g.bi = a.bi; g.bj = a.bj;
	g.ci = a.ci; g.cj = a.cj; g.di = a.di; g.dj = a.dj;
	g.alpha = 1.0; g.beta = 1.0;
	g.sA = a.sA; g.sB = a.sB; g.sC = a.sC; g.sD = a.sD;
	blasfeo_amd_dfact_arg f;
	f.m = live ? m : 0; f.n = n;
	f.ci = a.di; f.cj = a.dj; f.di = a.di; f.dj = a.dj;
	f.sC = a.sD; f.sD = a.sD;
	f.ipiv = nullptr;
	upd[i] = g;
	fact[i] = f;
}

static int mn_nmax(const blasfeo_amd_hint *hint, int mmax)
{
	return (hint && hint->max_n > 0 && hint->max_n < mmax) ? hint->max_n : mmax;
}

static int launch_potrf_mn_split(int ps, int batch, const blasfeo_amd_dfact_arg *args, const blasfeo_amd_hint *hint, int mmax, hipStream_t st, const char *name)
{
	const int nmax = mn_nmax(hint, mmax);
	if (nmax > MN_NMAX)
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: n up to %d supported when m > 128, got a bound of %d", name, MN_NMAX, nmax);
	const size_t tb = ((size_t)batch * sizeof(blasfeo_amd_dfact_arg) + 255) / 256 * 256;
	unsigned char *ws = (unsigned char *)workspace(tb + (size_t)batch * sizeof(blasfeo_amd_dtrsm_arg), st);
	if (!ws)
		return set_error(BLASFEO_AMD_ENOMEM, "%s: workspace allocation failed", name);
	blasfeo_amd_dfact_arg *top = (blasfeo_amd_dfact_arg *)ws;
	blasfeo_amd_dtrsm_arg *rows = (blasfeo_amd_dtrsm_arg *)(ws + tb);
	const int nb = (batch + 255) / 256;
	hipLaunchKernelGGL(k_mn_split, dim3(nb), dim3(256), 0, st, args, batch, top, rows);
	int rc = check_launch(name);
	if (rc)
		return rc;
	blasfeo_amd_hint ht = {nmax, nmax, 0};
	if ((rc = blasfeo_amd_dpotrf_l_batched(ps, batch, top, &ht, st)))
		return rc;
	blasfeo_amd_hint hr = {mmax, nmax, 0};
	if ((rc = trsm_rltn_given_dA(ps, batch, rows, &hr, st, name)))
		return rc;
	hipLaunchKernelGGL(k_mn_finish, dim3(nb), dim3(256), 0, st, args, batch);
	return check_launch(name);
}

// The m <= 128 LDS-resident launches hold the whole trapezoid in LDS, sized
// for np rows.  With a bound above 32 rows a plan kernel sorts the instances
// into buckets of 16 rows (instance i to bucket ceil(m/16) - 1) and each
// bucket gets a persistent launch whose LDS is sized for that bucket, largest
// first, so that an instance is charged its own footprint rather than the
// bound's, and holds it as packed lower tiles (79 KB at 128 rows: two
// workgroups per CU; the square form, 133 KB, allowed one).  Up to 32 rows,
// one workgroup per instance with the square form (at most 8.7 KB).
// counts[0..LDS_NB-1] and the queue heads counts[16..] start at 0.
constexpr int LDS_NB = 8;
template <typename R>
__global__ void k_lds_plan(const R *__restrict__ args, int batch, int *counts, int *lists)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	plan_bucket<LDS_NB>(i, i < batch ? min(LDS_NB - 1, max(0, (args[i].m - 1) / 16)) : -1, batch, counts, lists);
}

// Eight waves rather than four for a bucket's queued launch (tuning knob
// lds_threads[bucket] = 512): the update tasks of a panel step and of the
// syrk are shared by twice the waves, but the registers of the eight-wave
// form leave one workgroup per CU.  Four waves by default: with the packed
// tiles two 128-row instances share a CU, and 65536 x dsyrk_dpotrf_ln of
// m, k ~ U{8..128} take 8.38 ms on MI355X against 9.97 ms with eight waves
// from 97 rows (with the square LDS form, one instance per CU from 97 rows,
// eight waves there had won: 10.17 against 10.72 ms).
static bool lds_wide(int np, const int *queued)
{
	return queued && np > 32 && tune_value(TUNE_LDS_THREADS, (np - 1) / 16) == 512;
}

// go(np, grid, shm, count, list, head) launches the kernel form for np rows
// (the one-wave form for np <= 32)
template <typename R, typename Go>
static int launch_lds_buckets(const R *args, int batch, int npad, size_t elem, hipStream_t st, const char *name, Go go)
{
	if (npad <= 32)
		return go(npad, batch, elem * lds_elems(npad, false), (const int *)nullptr, (const int *)nullptr, (int *)nullptr);
	const size_t wsbytes = sizeof(int) * (64 + (size_t)LDS_NB * batch);
	int *ws = (int *)plan_workspace(wsbytes, st);
	if (!ws)
		return set_error(BLASFEO_AMD_ENOMEM, "%s: cannot allocate the %zu-byte plan workspace", name, wsbytes);
	int *counts = ws, *heads = ws + 16, *lists = ws + 64;
	hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, ws, 64);
	hipLaunchKernelGGL((k_lds_plan<R>), dim3((batch + 255) / 256), dim3(256), 0, st, args, batch, counts, lists);
	int rc = check_launch(name);
	if (rc)
		return rc;
	const int cus = num_cus() > 0 ? num_cus() : 256;
	for (int b = LDS_NB - 1; b >= 0; b--)
	{
		if (16 * b >= npad)
			continue;   // m beyond the bound: no instance runs (as in the one-launch form)
		const int np = min(16 * (b + 1), npad);
		const size_t shm = elem * lds_elems(np, true);
		// as many workgroups as the LDS holds, at most 16 one-wave or 8 four-wave ones per CU
		const int per_cu = max(1, min(np <= 32 ? 16 : 8, (int)(163840 / shm)));
		if ((rc = go(np, min(batch, per_cu * cus), shm, counts + b, lists + (size_t)b * batch, heads + b)))
			return rc;
	}
	return 0;
}

template <typename T, int MODE>
static int launch_potrf(int ps, int batch, const typename FactArg<T>::type *args, const blasfeo_amd_hint *hint, hipStream_t st, const char *name)
{
	int rc = fact_check(ps, batch, args, sizeof(T) == 4, name);
	if (rc || batch == 0)
		return rc;
	int mmax;
	if ((rc = fact_mmax_arg<T>(hint, name, &mmax)))
		return rc;
	if (mmax <= 128)
	{
		auto go = [&](int np, int grid, size_t shm, const int *cnt, const int *lst, int *hd) {
#define BFA_POTRF_LDS(PS_, NT_)                                                                            \
	do                                                                                                     \
	{                                                                                                      \
		if (cnt)                                                                                           \
			hipLaunchKernelGGL((k_potrf_lds_q<T, PS_, NT_, MODE == 1>), dim3(grid), dim3(NT_), shm, st, args, np, cnt, lst, hd); \
		else                                                                                               \
			hipLaunchKernelGGL((k_potrf_lds<T, PS_, NT_, MODE == 1>), dim3(grid), dim3(NT_), shm, st, args, np); \
	} while (0)
			if (np <= 32)
			{
				if (ps == 4)
					BFA_POTRF_LDS(4, 64);
				else if (ps == 8)
					BFA_POTRF_LDS(8, 64);
				else
					BFA_POTRF_LDS(16, 64);
			}
			else if (lds_wide(np, cnt))
			{
				if (ps == 4)
					hipLaunchKernelGGL((k_potrf_lds_q<T, 4, 2 * NTL, MODE == 1>), dim3(grid), dim3(2 * NTL), shm, st, args, np, cnt, lst, hd);
				else if (ps == 8)
					hipLaunchKernelGGL((k_potrf_lds_q<T, 8, 2 * NTL, MODE == 1>), dim3(grid), dim3(2 * NTL), shm, st, args, np, cnt, lst, hd);
				else
					hipLaunchKernelGGL((k_potrf_lds_q<T, 16, 2 * NTL, MODE == 1>), dim3(grid), dim3(2 * NTL), shm, st, args, np, cnt, lst, hd);
			}
			else
			{
				if (ps == 4)
					BFA_POTRF_LDS(4, NTL);
				else if (ps == 8)
					BFA_POTRF_LDS(8, NTL);
				else
					BFA_POTRF_LDS(16, NTL);
			}
#undef BFA_POTRF_LDS
			return check_launch(name);
		};
		return launch_lds_buckets(args, batch, (mmax + 15) / 16 * 16, sizeof(T), st, name, go);
	}
	if (MODE == 1)
		return launch_potrf_mn_split(ps, batch, (const blasfeo_amd_dfact_arg *)args, hint, mmax, st, name);
	if (MODE == 0 && mmax <= PT_NMAX)
	{
		// block panels of 16 columns (four 4-wave instances per CU, whose
		// chains overlap), 32 (two per CU) or 64 (one 8-wave instance per
		// CU): the potrf + trsm solve sequence for 2048 x n = 256 takes 1.31,
		// 1.65 and 1.90 ms on MI355X.  16 by default, the tuning knob
		// potrf_bw overrides
		const int knob = tune_value(TUNE_POTRF_BW, 0);
		const int bw = knob == 4 ? 4 : (knob == 2 ? 2 : 1);
		if (bw == 1)
		{
			const size_t shm3 = potrf_tiles_lds_bytes<T, 1>();
			if (ps == 4)
				hipLaunchKernelGGL((k_potrf_tiles<T, 4, 1, 256>), dim3(batch), dim3(256), shm3, st, args, mmax);
			else if (ps == 8)
				hipLaunchKernelGGL((k_potrf_tiles<T, 8, 1, 256>), dim3(batch), dim3(256), shm3, st, args, mmax);
			else
				hipLaunchKernelGGL((k_potrf_tiles<T, 16, 1, 256>), dim3(batch), dim3(256), shm3, st, args, mmax);
		}
		else if (bw == 4)
		{
			const size_t shm3 = potrf_tiles_lds_bytes<T, 4>();
			if (ps == 4)
				hipLaunchKernelGGL((k_potrf_tiles<T, 4, 4, 512>), dim3(batch), dim3(512), shm3, st, args, mmax);
			else if (ps == 8)
				hipLaunchKernelGGL((k_potrf_tiles<T, 8, 4, 512>), dim3(batch), dim3(512), shm3, st, args, mmax);
			else
				hipLaunchKernelGGL((k_potrf_tiles<T, 16, 4, 512>), dim3(batch), dim3(512), shm3, st, args, mmax);
		}
		else
		{
			const size_t shm3 = potrf_tiles_lds_bytes<T, 2>();
			if (ps == 4)
				hipLaunchKernelGGL((k_potrf_tiles<T, 4, 2, 256>), dim3(batch), dim3(256), shm3, st, args, mmax);
			else if (ps == 8)
				hipLaunchKernelGGL((k_potrf_tiles<T, 8, 2, 256>), dim3(batch), dim3(256), shm3, st, args, mmax);
			else
				hipLaunchKernelGGL((k_potrf_tiles<T, 16, 2, 256>), dim3(batch), dim3(256), shm3, st, args, mmax);
		}
		return check_launch(name);
	}
	const size_t shm = sizeof(T) * PW * ((mmax + 15) / 16 * 16 + 1);
	if (ps == 4)
		hipLaunchKernelGGL((k_potrf<T, 4>), dim3(batch), dim3(NTL), shm, st, args, mmax);
	else if (ps == 8)
		hipLaunchKernelGGL((k_potrf<T, 8>), dim3(batch), dim3(NTL), shm, st, args, mmax);
	else
		hipLaunchKernelGGL((k_potrf<T, 16>), dim3(batch), dim3(NTL), shm, st, args, mmax);
	return check_launch(name);
}

template <typename T, bool PIV = true>
static int launch_getrf(int ps, int batch, const typename FactArg<T>::type *args, const blasfeo_amd_hint *hint, hipStream_t st, const char *name)
{
	int rc = fact_check(ps, batch, args, sizeof(T) == 4, name);
	if (rc || batch == 0)
		return rc;
	int mmax;
	if ((rc = fact_mmax_arg<T>(hint, name, &mmax)))
		return rc;
	const size_t shm = sizeof(T) * PW * ((mmax + 15) / 16 * 16 + 1) + sizeof(int) * mmax;
	// the size order (k_getrf_plan): counts (64 ints) and one list per bucket
	int *ws = (int *)workspace(sizeof(int) * (64 + (size_t)GETRF_NB * batch), st);
	if (!ws)
		return set_error(BLASFEO_AMD_ENOMEM, "%s: workspace allocation failed", name);
	int *counts = ws, *lists = ws + 64;
	hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, counts, 64);
	hipLaunchKernelGGL((k_getrf_plan<T>), dim3((batch + 255) / 256), dim3(256), 0, st, args, batch, counts, lists);
	if (ps == 4)
		hipLaunchKernelGGL((k_getrf<T, 4, PIV>), dim3(batch), dim3(NTL), shm, st, args, mmax, batch, counts, lists);
	else if (ps == 8)
		hipLaunchKernelGGL((k_getrf<T, 8, PIV>), dim3(batch), dim3(NTL), shm, st, args, mmax, batch, counts, lists);
	else
		hipLaunchKernelGGL((k_getrf<T, 16, PIV>), dim3(batch), dim3(NTL), shm, st, args, mmax, batch, counts, lists);
	return check_launch(name);
}

static int launch_syrk_potrf(int ps, int batch, const blasfeo_amd_dsyrk_dpotrf_arg *args, const blasfeo_amd_hint *hint, hipStream_t st, const char *name)
{
	int rc = fact_check(ps, batch, args, false, name);
	if (rc || batch == 0)
		return rc;
	int mmax;
	if ((rc = fact_mmax_arg<double>(hint, name, &mmax)))
		return rc;
	if (mmax > 128)
	{
		const int nmax = mn_nmax(hint, mmax);
		if (nmax > MN_NMAX)
			return set_error(BLASFEO_AMD_ETOOBIG, "%s: n up to %d supported when m > 128, got a bound of %d", name, MN_NMAX, nmax);
		const size_t gb = ((size_t)batch * sizeof(blasfeo_amd_dgemm_arg) + 255) / 256 * 256;
		const size_t fb = ((size_t)batch * sizeof(blasfeo_amd_dfact_arg) + 255) / 256 * 256;
		const size_t tb = fb;
		// layout: [gemm records][fact records][potrf_l_mn split: top records][rows]
		unsigned char *ws = (unsigned char *)workspace(gb + fb + tb + (size_t)batch * sizeof(blasfeo_amd_dtrsm_arg), st);
		if (!ws)
			return set_error(BLASFEO_AMD_ENOMEM, "%s: workspace allocation failed", name);
		blasfeo_amd_dgemm_arg *upd = (blasfeo_amd_dgemm_arg *)ws;
		blasfeo_amd_dfact_arg *fact = (blasfeo_amd_dfact_arg *)(ws + gb);
		blasfeo_amd_dfact_arg *top = (blasfeo_amd_dfact_arg *)(ws + gb + fb);
		blasfeo_amd_dtrsm_arg *rows = (blasfeo_amd_dtrsm_arg *)(ws + gb + fb + tb);
		const int nb = (batch + 255) / 256;
		hipLaunchKernelGGL(k_sp_split, dim3(nb), dim3(256), 0, st, args, batch, upd, fact);
		if ((rc = check_launch(name)))
			return rc;
		blasfeo_amd_hint hg = {mmax, nmax, hint ? hint->max_k : 0};
		if ((rc = blasfeo_amd_dsyrk_ln_mn_batched(ps, batch, upd, &hg, st)))
			return rc;
		hipLaunchKernelGGL(k_mn_split, dim3(nb), dim3(256), 0, st, fact, batch, top, rows);
		if ((rc = check_launch(name)))
			return rc;
		blasfeo_amd_hint ht = {nmax, nmax, 0};
		if ((rc = blasfeo_amd_dpotrf_l_batched(ps, batch, top, &ht, st)))
			return rc;
		blasfeo_amd_hint hr = {mmax, nmax, 0};
		if ((rc = trsm_rltn_given_dA(ps, batch, rows, &hr, st, name)))
			return rc;
		hipLaunchKernelGGL(k_mn_finish, dim3(nb), dim3(256), 0, st, fact, batch);
		return check_launch(name);
	}
	auto go = [&](int np, int grid, size_t shm, const int *cnt, const int *lst, int *hd) {
#define BFA_SP_LDS(PS_, NT_)                                                                               \
	do                                                                                                     \
	{                                                                                                      \
		if (cnt)                                                                                           \
			hipLaunchKernelGGL((k_syrk_potrf_lds_q<double, PS_, NT_>), dim3(grid), dim3(NT_), shm, st, args, np, cnt, lst, hd); \
		else                                                                                               \
			hipLaunchKernelGGL((k_syrk_potrf_lds<double, PS_, NT_>), dim3(grid), dim3(NT_), shm, st, args, np); \
	} while (0)
		if (np <= 32)
		{
			if (ps == 4)
				BFA_SP_LDS(4, 64);
			else
				BFA_SP_LDS(8, 64);
		}
		else if (lds_wide(np, cnt))
		{
			if (ps == 4)
				hipLaunchKernelGGL((k_syrk_potrf_lds_q<double, 4, 2 * NTL>), dim3(grid), dim3(2 * NTL), shm, st, args, np, cnt, lst, hd);
			else
				hipLaunchKernelGGL((k_syrk_potrf_lds_q<double, 8, 2 * NTL>), dim3(grid), dim3(2 * NTL), shm, st, args, np, cnt, lst, hd);
		}
		else
		{
			if (ps == 4)
				BFA_SP_LDS(4, NTL);
			else
				BFA_SP_LDS(8, NTL);
		}
#undef BFA_SP_LDS
		return check_launch(name);
	};
	return launch_lds_buckets(args, batch, (mmax + 15) / 16 * 16, sizeof(double), st, name, go);
}

} // namespace bfa

using namespace bfa;

extern "C" {

int blasfeo_amd_dpotrf_l_batched(int ps, int batch, const blasfeo_amd_dfact_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_potrf<double, 0>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dpotrf_l_batched"); }
// m <= 128: the LDS-resident forms; larger m: the trapezoid as potrf_l of
// the top block plus the solve of the rows below (see launch_potrf_mn_split).
int blasfeo_amd_dpotrf_l_mn_batched(int ps, int batch, const blasfeo_amd_dfact_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_potrf<double, 1>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dpotrf_l_mn_batched"); }
int blasfeo_amd_dsyrk_dpotrf_ln_batched(int ps, int batch, const blasfeo_amd_dsyrk_dpotrf_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_syrk_potrf(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dsyrk_dpotrf_ln_batched"); }
int blasfeo_amd_dgetrf_rp_batched(int ps, int batch, const blasfeo_amd_dfact_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_getrf<double>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dgetrf_rp_batched"); }
int blasfeo_amd_dgetrf_np_batched(int ps, int batch, const blasfeo_amd_dfact_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_getrf<double, false>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dgetrf_np_batched"); }
int blasfeo_amd_spotrf_l_batched(int ps, int batch, const blasfeo_amd_sfact_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_potrf<float, 0>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_spotrf_l_batched"); }
int blasfeo_amd_sgetrf_rp_batched(int ps, int batch, const blasfeo_amd_sfact_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_getrf<float>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_sgetrf_rp_batched"); }

} // extern "C"
