// bfa_blas3.hip -- batched level-3 kernels for gfx950: gemm_nt, gemm_nn,
// syrk_ln(_mn), trmm_rlnn and trsm_rltn on panel-major matrices.
//
// Launch model: one workgroup per (instance, tile slot), `slots` slots per
// instance (tile t of an instance goes to slot t mod slots), so a batch of
// small instances uses one workgroup each and a batch of large ones spreads
// its 64x64 tiles over several (tile engine: bfa_gemm_core.hpp).  The gemm
// and trmm launches are 1-D (batch_slot): in XCD-aware order when an
// instance has more than four slots (the slots of one instance run together
// on one XCD and share its L2: 4000 gemm_nt of U{64..512} 7.68 -> 6.41 ms,
// 20000 of U{4..300} 7.62 -> 6.34 ms on MI355X), instance-fastest otherwise
// (100000 of U{4..128} on this engine: 5.35 against 8.07 ms in XCD order).
// trsm keeps a 2-D grid (x the instance, y the slot).
#include "bfa_gemm_core.hpp"
#include "bfa_gemm_wave.hpp"
#include "bfa_chol_core.hpp"
#include "bfa_internal.hpp"

namespace bfa {

template <typename T, int PS, int OP>
__global__ __launch_bounds__(NT3) __attribute__((amdgpu_waves_per_eu(3))) void k_gemm(const typename GemmArg<T>::type *__restrict__ args,
	int batch, int slots)
{
	__shared__ T As[KC * LDT];
	__shared__ T Bs[KC * LDT];
	int inst, slot;
	batch_slot(slots > 4, batch, slots, inst, slot);
	if (inst >= batch)
		return;
	const typename GemmArg<T>::type &a = args[inst];
	const int m = a.m;
	const int n = (OP == OP_SYRK) ? a.m : a.n;
	if (m <= 0 || n <= 0)
		return;
	if (threadIdx.x == 0 && slot == 0)
		a.sD->use_dA = 0;
	gemm_core<T, PS, OP>(As, Bs, m, n, a.k, a.alpha, a.beta,
		view<T, PS>(a.sA), a.ai, a.aj, view<T, PS>(a.sB), a.bi, a.bj,
		view<T, PS>(a.sC), a.ci, a.cj, view<T, PS>(a.sD), a.di, a.dj, slot, slots);
}

// D = alpha * B * A, A lower n x n (blasfeo_ref/x_blas3_ref.c:3218-3330)
template <typename T, int PS>
__global__ __launch_bounds__(NT3) __attribute__((amdgpu_waves_per_eu(3))) void k_trmm_rlnn(const typename TrsmArg<T>::type *__restrict__ args, int batch, int slots)
{
	__shared__ T As[KC * LDT];
	__shared__ T Bs[KC * LDT];
	int inst, slot;
	batch_slot(slots > 4, batch, slots, inst, slot);
	if (inst >= batch)
		return;
	const typename TrsmArg<T>::type &a = args[inst];
	if (a.m <= 0 || a.n <= 0)
		return;
	if (threadIdx.x == 0 && slot == 0)
		a.sD->use_dA = 0;
	const PM<T, PS> gA = view<T, PS>(a.sA);
	gemm_core<T, PS, OP_TRMM_RLNN>(As, Bs, a.m, a.n, a.n, a.alpha, T(0),
		gA, a.ai, a.aj, view<T, PS>(a.sB), a.bi, a.bj,
		gA, 0, 0, view<T, PS>(a.sD), a.di, a.dj, slot, slots);
}

// ---------------------------------------------------------------------------
// trsm_rltn: D = alpha * B * A^{-T}, A lower n x n (blasfeo_ref/x_blas3_ref.c:1642-1752)
// A workgroup solves a 64-row block of B: D's row block lives in LDS while
// the columns are produced 16 at a time: the already-solved columns are
// applied with matrix-core steps (X -= D(:,0:j) A(j,0:j)^T), then one wave
// solves the 16x16 diagonal triangle row by row with the reciprocal diagonal.
// ---------------------------------------------------------------------------
constexpr int TRSM_NMAX = 256;
constexpr int LDX = TM + 1;

// One instance of trsm_rltn, row blocks t0, t0 + tstep, ... (all threads of
// the workgroup call it together).  DA_GIVEN: the reciprocal diagonal is
// read from sA->dA whatever use_dA says (the trapezoid Cholesky, whose
// diagonal block was just factored into the same descriptor with its
// reciprocals stored there).
template <typename T, int PS, bool DA_GIVEN>
__device__ __forceinline__ void trsm_rltn_instance(const typename TrsmArg<T>::type &a, int nmax, int t0, int tstep,
	T *Xs, T *inv_s, T *Ad)
{
	typedef typename Vec4<T>::type V4;
	const int m = a.m, n = a.n;
	if (m <= 0 || n <= 0 || n > nmax)
		return;
	const PM<T, PS> gA = view<T, PS>(a.sA);
	const PM<T, PS> gB = view<T, PS>(a.sB);
	PM<T, PS> gD = view<T, PS>(a.sD);
	const T alpha = a.alpha;
	const int lane = lane_id();
	const int w = wave_id();

	// reciprocal diagonal: cached copy when valid at launch, else 1/A_jj
	// (an output aliasing A invalidates the cache first, as in the reference)
	const bool cached = DA_GIVEN || (a.ai == 0 && a.aj == 0 && a.sD != a.sA && a.sA->use_dA >= n);
	for (int j = threadIdx.x; j < n; j += NT3)
		inv_s[j] = cached ? a.sA->dA[j] : T(1) / gA.at(a.ai + j, a.aj + j);

	const int mt = (m + TM - 1) / TM;
	const int nb = (n + 15) / 16;
	for (int t = t0; t < mt; t += tstep)
	{
		const int i0 = t * TM;
		const int mv = min(TM, m - i0);
		__syncthreads();
		// X <- alpha * B(row block, :)
		g2s<T, PS, NT3, false>(gB, a.bi + i0, a.bj, TM, nb * 16, mv, n, Xs, LDX);
		__syncthreads();
		if (alpha != T(1))   // (x * 1 == x: nothing to do for alpha = 1)
			for (int e = threadIdx.x; e < nb * 16 * LDX; e += NT3)
				Xs[e] *= alpha;
		for (int cb = 0; cb < nb; cb++)
		{
			const int c0 = cb * 16;
			const int cv = min(16, n - c0);
			// diagonal block of A (zero outside the valid triangle): loaded
			// here, stored to LDS after the update below, so that its load
			// latency overlaps the update (NT3 = 256 threads, one element each)
			__syncthreads();
			T adv;
			{
				const int r = threadIdx.x & 15, c = threadIdx.x >> 4;
				adv = (r < cv && c < cv && r > c) ? gA.at(a.ai + c0 + r, a.aj + c0 + c) : T(0);
			}
			// X(:, c0:c0+16) -= X(:, 0:c0) * A(c0:c0+16, 0:c0)^T ; wave w owns rows 16w..16w+15.
			// A is read from global memory four K steps (16 columns) at a time
			// with the next four in flight; c0 is a multiple of 16, so no
			// step runs past it.  The products are accumulated in the same
			// order as one step at a time.
			if (c0 > 0)
			{
				V4 acc = V4{0, 0, 0, 0};
				const int jrow = a.ai + c0 + (lane & 15);
				const bool jok = (lane & 15) < cv;
				T bn[4];
#pragma unroll
				for (int u = 0; u < 4; u++)
					bn[u] = jok ? gA.at(jrow, a.aj + 4 * u + (lane >> 4)) : T(0);
				for (int kk = 0; kk < c0; kk += 16)
				{
					T bv[4], av[4];
#pragma unroll
					for (int u = 0; u < 4; u++)
						bv[u] = bn[u];
					if (kk + 16 < c0)
#pragma unroll
						for (int u = 0; u < 4; u++)
							bn[u] = jok ? gA.at(jrow, a.aj + kk + 16 + 4 * u + (lane >> 4)) : T(0);
#pragma unroll
					for (int u = 0; u < 4; u++)
						av[u] = Xs[(kk + 4 * u + (lane >> 4)) * LDX + 16 * w + (lane & 15)];
#pragma unroll
					for (int u = 0; u < 4; u++)
						acc = mfma16(av[u], bv[u], acc);
				}
#pragma unroll
				for (int r = 0; r < 4; r++)
				{
					const int il = 16 * w + acc_row<T>(lane, r);
					const int jl = c0 + (lane & 15);
					Xs[jl * LDX + il] -= acc[r];
				}
			}
			Ad[(threadIdx.x & 15) * 17 + (threadIdx.x >> 4)] = adv;
			__syncthreads();
			// triangular solve of the 16 columns, one row per thread of wave 0
			if (w == 0)
			{
				const int il = lane;
				T x[16];
#pragma unroll
				for (int j = 0; j < 16; j++)
					x[j] = Xs[(c0 + j) * LDX + il];
#pragma unroll
				for (int j = 0; j < 16; j++)
				{
					if (j < cv)
					{
						T c = x[j];
#pragma unroll
						for (int l = 0; l < j; l++)
							c -= x[l] * Ad[j * 17 + l];
						x[j] = c * inv_s[c0 + j];
					}
				}
#pragma unroll
				for (int j = 0; j < 16; j++)
					Xs[(c0 + j) * LDX + il] = x[j];
			}
		}
		__syncthreads();
		s2g<T, PS, NT3, false>(gD, a.di + i0, a.dj, mv, n, Xs, LDX, 0);
	}
}

// With a queue (count != nullptr) the launch is persistent: each workgroup
// takes instances from the list (one atomic per instance on *head) and
// solves all their row blocks; otherwise instance blockIdx.x, row blocks
// blockIdx.y, blockIdx.y + gridDim.y, ...
template <typename T, int PS, bool DA_GIVEN = false>
__global__ __launch_bounds__(NT3) void k_trsm_rltn(const typename TrsmArg<T>::type *__restrict__ args, int nmax,
	const int *__restrict__ count, const int *__restrict__ list, int *head)
{
	extern __shared__ unsigned char smem_raw[];
	T *Xs = reinterpret_cast<T *>(smem_raw);      // [npad][LDX] column-major row block
	__shared__ T inv_s[TRSM_NMAX];
	__shared__ T Ad[16 * 17];                       // diagonal 16x16 block of A
	__shared__ int qslot;
	if (!count)
	{
		trsm_rltn_instance<T, PS, DA_GIVEN>(args[blockIdx.x], nmax, blockIdx.y, gridDim.y, Xs, inv_s, Ad);
		return;
	}
	const int cnt = *count;
	for (;;)
	{
		__syncthreads();   // every thread has read the previous slot
		if (threadIdx.x == 0)
			qslot = atomicAdd(head, 1);
		__syncthreads();
		const int t = qslot;
		if (t >= cnt)
			break;
		trsm_rltn_instance<T, PS, DA_GIVEN>(args[list[t]], nmax, 0, 1, Xs, inv_s, Ad);
	}
}

// ---------------------------------------------------------------------------
// trsm_llnn / trsm_lltn / trsm_llnu / trsm_lunn: D = alpha * A^{-1} * B,
// alpha * A^{-T} * B with A lower, alpha * A^{-1} * B with A unit lower, or
// alpha * U^{-1} * B with U the upper triangle of A; m x m, B m x n
// (blasfeo_ref/x_blas3_ref.c:445-673, 674-900, 572-672, 920-1100).  lunn is
// lltn on the transposed view of A (U(i,k) = A^T(i,k)), so it shares lltn's
// bottom-up sweep; llnu is llnn without the reciprocal diagonal.
// A workgroup keeps a block of up to LNC columns of X = alpha*B for all m rows
// in LDS (column-major, odd leading dimension) and sweeps the 16-row blocks of
// A, top-down for llnn and bottom-up for lltn: wave 0 solves the 16x16
// diagonal triangle of the block (one column of X per lane, the 16 rows in
// registers, multiplying by the reciprocal diagonal as the reference does),
// then every thread applies the solved rows to one remaining row of X (below
// for llnn, above for lltn), reading that row's 16 elements of A once.  The
// next diagonal block is staged into the other half of a double buffer during
// the update, so each block costs two workgroup barriers.
// ---------------------------------------------------------------------------
constexpr int LNC = 16;
template <typename T> constexpr int trsml_mmax() { return sizeof(T) == 8 ? 1024 : 2048; }

enum TrsmLeftMode { TL_LLNN = 0, TL_LLTN = 1, TL_LLNU = 2, TL_LUNN = 3 };
// the sweep runs bottom-up for the transposed-lower and the upper solves
constexpr bool tl_back(int mode) { return mode == TL_LLTN || mode == TL_LUNN; }

// element (r, c) of the lower triangle the sweep works on: A itself, or A^T
// for the upper solve (its upper triangle read as a lower one)
template <typename T, int PS, int MODE>
__device__ __forceinline__ T tl_at(const PM<T, PS> gA, int ai, int aj, int r, int c)
{
	return MODE == TL_LUNN ? gA.at(ai + c, aj + r) : gA.at(ai + r, aj + c);
}

template <typename T, int PS, int MODE>
__device__ __forceinline__ void trsml_stage_diag(const PM<T, PS> gA, int ai, int aj, int r0, int rv, T *Ad)
{
	for (int e = threadIdx.x; e < 256; e += NT3)
	{
		const int r = e & 15, c = e >> 4;
		T v = T(0);
		if (r < rv && c < rv && r > c)
			v = tl_at<T, PS, MODE>(gA, ai, aj, r0 + r, r0 + c);
		Ad[r * 17 + c] = v;
	}
}

template <typename T, int PS, int MODE>
__global__ __launch_bounds__(NT3) void k_trsm_left(const typename TrsmArg<T>::type *__restrict__ args, int mmax)
{
	constexpr bool TRANS = tl_back(MODE);
	constexpr bool UNIT = MODE == TL_LLNU;
	extern __shared__ unsigned char smem_raw[];
	__shared__ T Ad[2][16 * 17];
	const int mpad = (mmax + 15) / 16 * 16;
	const int LDM = mpad + 1;
	T *Xs = reinterpret_cast<T *>(smem_raw);   // [LNC][LDM]
	T *inv_s = Xs + (size_t)LNC * LDM;           // [mpad]

	const typename TrsmArg<T>::type &a = args[blockIdx.x];
	const int m = a.m, n = a.n;
	if (m <= 0 || n <= 0 || m > mmax)
		return;
	const PM<T, PS> gA = view<T, PS>(a.sA);
	const PM<T, PS> gB = view<T, PS>(a.sB);
	PM<T, PS> gD = view<T, PS>(a.sD);
	const int ai = a.ai, aj = a.aj;
	const T alpha = a.alpha;
	const int lane = lane_id();
	const int w = wave_id();
	const int mb = (m + 15) / 16;

	if (!UNIT)
	{
		const bool cached = (ai == 0 && aj == 0 && a.sD != a.sA && a.sA->use_dA >= m);
		for (int j = threadIdx.x; j < m; j += NT3)
			inv_s[j] = cached ? a.sA->dA[j] : T(1) / gA.at(ai + j, aj + j);
	}

	for (int j0 = blockIdx.y * LNC; j0 < n; j0 += gridDim.y * LNC)
	{
		const int nc = min(LNC, n - j0);
		__syncthreads();
		g2s<T, PS, NT3, false>(gB, a.bi, a.bj + j0, mb * 16, LNC, m, nc, Xs, LDM);
		{
			const int rb0 = TRANS ? mb - 1 : 0;
			trsml_stage_diag<T, PS, MODE>(gA, ai, aj, rb0 * 16, min(16, m - rb0 * 16), Ad[0]);
		}
		__syncthreads();
		for (int e = threadIdx.x; e < LNC * LDM; e += NT3)
			Xs[e] *= alpha;
		__syncthreads();
		for (int s = 0; s < mb; s++)
		{
			const int rb = TRANS ? mb - 1 - s : s;
			const int r0 = rb * 16;
			const int rv = min(16, m - r0);
			const T *A0 = Ad[s & 1];
			// diagonal triangle: lane c owns column c of the block
			if (w == 0 && lane < nc)
			{
				T x[16];
#pragma unroll
				for (int l = 0; l < 16; l++)
					x[l] = Xs[lane * LDM + r0 + l];
				if (!TRANS)
				{
#pragma unroll
					for (int i = 0; i < 16; i++)
					{
						if (i < rv)
						{
							T c = x[i];
#pragma unroll
							for (int l = 0; l < i; l++)
								c -= A0[i * 17 + l] * x[l];
							x[i] = UNIT ? c : c * inv_s[r0 + i];
						}
					}
				}
				else
				{
#pragma unroll
					for (int i = 15; i >= 0; i--)
					{
						if (i < rv)
						{
							T c = x[i];
#pragma unroll
							for (int l = i + 1; l < 16; l++)
								c -= A0[l * 17 + i] * x[l];
							x[i] = c * inv_s[r0 + i];
						}
					}
				}
#pragma unroll
				for (int l = 0; l < 16; l++)
					if (l < rv)
						Xs[lane * LDM + r0 + l] = x[l];
			}
			__syncthreads();
			// stage the next diagonal block, then update the remaining rows
			if (s + 1 < mb)
			{
				const int rbn = TRANS ? rb - 1 : rb + 1;
				trsml_stage_diag<T, PS, MODE>(gA, ai, aj, rbn * 16, min(16, m - rbn * 16), Ad[(s + 1) & 1]);
			}
			const int ilo = TRANS ? 0 : r0 + 16;
			const int ihi = TRANS ? r0 : m;
			for (int i = ilo + threadIdx.x; i < ihi; i += NT3)
			{
				T av[16];
#pragma unroll
				for (int l = 0; l < 16; l++)
				{
					T v = T(0);
					if (l < rv)
						v = TRANS ? tl_at<T, PS, MODE>(gA, ai, aj, r0 + l, i) : tl_at<T, PS, MODE>(gA, ai, aj, i, r0 + l);
					av[l] = v;
				}
				for (int c = 0; c < nc; c++)
				{
					T acc = Xs[c * LDM + i];
#pragma unroll
					for (int l = 0; l < 16; l++)
						acc -= av[l] * Xs[c * LDM + r0 + l];
					Xs[c * LDM + i] = acc;
				}
			}
			__syncthreads();
		}
		s2g<T, PS, NT3, false>(gD, a.di, a.dj + j0, m, nc, Xs, LDM, 0);
	}
}

// ---------------------------------------------------------------------------
// The same four left-side solves for few right-hand sides (chosen when the hint
// bounds n by 4) and m <= 256: one wavefront per instance, row i of X in lane
// i % 64 (slot i / 64), four columns of X at a time in registers.  Column j of the solve is taken in order (top-down for llnn
// and llnu, bottom-up for lltn and lunn): row j's values are scaled by the
// reciprocal diagonal and broadcast with v_readlane, and every row still to
// be solved drops its multiple of them.  No barriers; the elements of A for
// the next column are loaded while the current one is applied.
// ---------------------------------------------------------------------------
constexpr int THIN_RHS = 4;
constexpr int THIN_Q = 4;

template <typename T, int PS, int MODE>
__device__ __forceinline__ void thin_coef(const PM<T, PS> gA, int ai, int aj, int m, int j, T (&e)[THIN_Q])
{
	constexpr bool BACK = tl_back(MODE);
	const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
	for (int q = 0; q < THIN_Q; q++)
	{
		const int i = lane + q * WAVE;
		T v = T(0);
		if (BACK ? (i < j) : (i > j && i < m))
			v = MODE == TL_LLTN ? gA.at(ai + j, aj + i) : gA.at(ai + i, aj + j);
		e[q] = v;
	}
}

template <typename T, int PS, int MODE>
__global__ __launch_bounds__(WAVE) void k_trsm_left_thin(const typename TrsmArg<T>::type *__restrict__ args)
{
	constexpr bool BACK = tl_back(MODE);
	constexpr bool UNIT = MODE == TL_LLNU;
	const typename TrsmArg<T>::type &a = args[blockIdx.x];
	const int m = a.m;
	if (m <= 0 || a.n <= 0 || m > THIN_Q * WAVE)
		return;
	const PM<T, PS> gA = view_flat<T, PS>(a.sA);
	const PM<T, PS> gB = view_flat<T, PS>(a.sB);
	PM<T, PS> gD = view_flat<T, PS>(a.sD);
	const int ai = a.ai, aj = a.aj;
	const int lane = threadIdx.x;
	const bool cached = !UNIT && ai == 0 && aj == 0 && a.sD != a.sA && a.sA->use_dA >= m;
	const T *dA = a.sA->dA;
	// right-hand sides in groups of THIN_RHS (one group when n <= 4)
	for (int c0 = 0; c0 < a.n; c0 += THIN_RHS)
	{
	const int n = min(THIN_RHS, a.n - c0);
	T x[THIN_Q][THIN_RHS];
#pragma unroll
	for (int q = 0; q < THIN_Q; q++)
#pragma unroll
		for (int c = 0; c < THIN_RHS; c++)
		{
			const int i = lane + q * WAVE;
			x[q][c] = (i < m && c < n) ? a.alpha * gB.at(a.bi + i, a.bj + c0 + c) : T(0);
		}
	T e[THIN_Q];
	thin_coef<T, PS, MODE>(gA, ai, aj, m, BACK ? m - 1 : 0, e);
	// the reciprocal diagonal of column j, also loaded one column ahead
	auto recip = [&](int jj) -> T {
		return UNIT ? T(1) : (cached ? dA[jj] : T(1) / gA.at(ai + jj, aj + jj));
	};
	T d = recip(BACK ? m - 1 : 0);
	// the solve's columns j in order, grouped by the register slot qb = j / 64
	// that holds row j (unrolled, so every register index is a constant)
#pragma unroll
	for (int qq = 0; qq < THIN_Q; qq++)
	{
		const int qb = BACK ? THIN_Q - 1 - qq : qq;
		for (int ll = 0; ll < WAVE; ll++)
		{
			const int lj = BACK ? WAVE - 1 - ll : ll;
			const int j = qb * WAVE + lj;
			if (j >= m)
			{
				if (BACK)
					continue;
				break;
			}
			const int jn = BACK ? j - 1 : j + 1;          // the next column
			const bool more = BACK ? jn >= 0 : jn < m;
			T en[THIN_Q];
			T dn = T(1);
			if (more)
			{
				thin_coef<T, PS, MODE>(gA, ai, aj, m, jn, en);
				dn = recip(jn);
			}
			T xj[THIN_RHS];
#pragma unroll
			for (int c = 0; c < THIN_RHS; c++)
				xj[c] = readlane(x[qb][c], lj) * d;
#pragma unroll
			for (int q = 0; q < THIN_Q; q++)
			{
				// slots whose rows are all solved already (before qb on the way
				// down, after it on the way up) or all past m are left alone
				if (BACK ? q > qb : q < qb)
					continue;
				if (q * WAVE >= m)
					continue;
				const int i = lane + q * WAVE;
				// only the rows still to be solved take the update (a solved
				// row is never touched again, so an Inf or NaN in xj cannot
				// reach it through a zero coefficient); row j itself keeps its
				// unscaled value, scaled by its reciprocal after the sweep
				const bool open = BACK ? i < j : i > j;
#pragma unroll
				for (int c = 0; c < THIN_RHS; c++)
					x[q][c] = open ? x[q][c] - e[q] * xj[c] : x[q][c];
			}
			if (more)
			{
#pragma unroll
				for (int q = 0; q < THIN_Q; q++)
					e[q] = en[q];
				d = dn;
			}
		}
	}
	// row i's solved value is its unscaled value times the reciprocal the
	// sweep used for column i (the same product the broadcast formed)
#pragma unroll
	for (int q = 0; q < THIN_Q; q++)
	{
		const int i = lane + q * WAVE;
		if (i < m)
		{
			const T di = recip(i);
#pragma unroll
			for (int c = 0; c < THIN_RHS; c++)
				if (c < n)
					gD.at(a.di + i, a.dj + c0 + c) = UNIT ? x[q][c] : x[q][c] * di;
		}
	}
	}
}

// Cache update after trsm_rltn / trsm_llnn / trsm_lltn, in the reference's
// order (blasfeo_hp_pm/d_blas3_lib4.c:4157-4178; blasfeo_ref/x_blas3_ref.c:
// 466-480): invalidate D's cache, then refresh A's reciprocal diagonal unless
// it was valid.  Run as its own launch so that no workgroup of the solve
// reads a dA another one is writing.  LEFT: the triangle is m x m, else n x n.
template <typename T, int PS, bool LEFT, bool UNIT = false>
__global__ void k_trsm_dA(const typename TrsmArg<T>::type *__restrict__ args)
{
	const typename TrsmArg<T>::type &a = args[blockIdx.x];
	const int n = LEFT ? a.m : a.n;
	if (a.m <= 0 || a.n <= 0)
		return;
	if (UNIT)
	{
		// trsm_llnu reads no reciprocal diagonal: only D's cache is invalidated
		if (threadIdx.x == 0)
			a.sD->use_dA = 0;
		return;
	}
	const bool zero_off = (a.ai == 0 && a.aj == 0);
	const int use0 = (a.sD == a.sA) ? 0 : a.sA->use_dA;
	__syncthreads();
	if (threadIdx.x == 0)
		a.sD->use_dA = 0;
	if (zero_off && use0 >= n)
		return;
	const PM<T, PS> gA = view<T, PS>(a.sA);
	for (int j = threadIdx.x; j < n; j += blockDim.x)
		a.sA->dA[j] = T(1) / gA.at(a.ai + j, a.aj + j);
	__syncthreads();
	if (threadIdx.x == 0)
		a.sA->use_dA = zero_off ? n : 0;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static int tiles_y(const blasfeo_amd_hint *hint, int tm, int tn)
{
	if (!hint || hint->max_m <= 0)
		return 1;
	int mt = (hint->max_m + tm - 1) / tm;
	int nt = hint->max_n > 0 ? (hint->max_n + tn - 1) / tn : 1;
	int y = mt * nt;
	int cap = tune_value(TUNE_GEMM_GRID_Y, 0);
	if (cap <= 0)
		cap = 64;
	return y < 1 ? 1 : (y > cap ? cap : y);
}

static int check_common(int ps, int batch, const void *args, bool is_float, const char *name)
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

// 1-D grid of batch * slots workgroups rounded up to a multiple of XCDS, for
// kernels that map their workgroup id through xcd_local_index
static int xcd_grid(int batch, int slots, int threads, dim3 *grid, const char *name)
{
	const long long wgs = ((long long)batch * slots + XCDS - 1) / XCDS * XCDS;
	if (wgs * threads > 0x7fffffffLL)
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: batch %d too large for one launch", name, batch);
	*grid = dim3((unsigned)wgs);
	return BLASFEO_AMD_OK;
}

template <typename T, int OP>
static int launch_gemm(int ps, int batch, const typename GemmArg<T>::type *args, const blasfeo_amd_hint *hint, hipStream_t st, const char *name)
{
	int rc = check_common(ps, batch, args, sizeof(T) == 4, name);
	if (rc || batch == 0)
		return rc;
	// sizes bounded by 256 (at most 64 tiles of 32 x 32 per instance): one
	// wavefront per output tile, operands straight from global memory; the
	// LDS-staged 64 x 64 engine for larger or unbounded sizes.  The
	// "gemm_wave" knob forces the wave form off (2).
	const int gwk = tune_value(TUNE_GEMM_WAVE, 0);
	if (gwk != 2 && hint && hint->max_m > 0 && hint->max_m <= 256 &&
	    (OP == OP_SYRK || (hint->max_n > 0 && hint->max_n <= 256)))
	{
		const int mt = (hint->max_m + GW_T - 1) / GW_T;
		const int nt = OP == OP_SYRK ? mt : (hint->max_n + GW_T - 1) / GW_T;
		int gy = (OP == OP_SYRK || OP == OP_SYRK_MN) ? mt * (mt + 1) / 2 : mt * nt;
		gy = gy < 1 ? 1 : gy;
		// one wave per (instance, tile slot), XCD-aware order (k_gemm_wave)
		dim3 gridw;
		if ((rc = xcd_grid(batch, gy, WAVE, &gridw, name)))
			return rc;
		if (ps == 4)
			hipLaunchKernelGGL((k_gemm_wave<T, 4, OP>), gridw, dim3(WAVE), 0, st, args, batch, gy);
		else if (ps == 8)
			hipLaunchKernelGGL((k_gemm_wave<T, 8, OP>), gridw, dim3(WAVE), 0, st, args, batch, gy);
		else
			hipLaunchKernelGGL((k_gemm_wave<T, 16, OP>), gridw, dim3(WAVE), 0, st, args, batch, gy);
		return check_launch(name);
	}
	const int slots = tiles_y(hint, TM, TN);
	dim3 grid;
	if ((rc = xcd_grid(batch, slots, NT3, &grid, name)))
		return rc;
	if (ps == 4)
		hipLaunchKernelGGL((k_gemm<T, 4, OP>), grid, dim3(NT3), 0, st, args, batch, slots);
	else if (ps == 8)
		hipLaunchKernelGGL((k_gemm<T, 8, OP>), grid, dim3(NT3), 0, st, args, batch, slots);
	else
		hipLaunchKernelGGL((k_gemm<T, 16, OP>), grid, dim3(NT3), 0, st, args, batch, slots);
	return check_launch(name);
}

template <typename T>
static int launch_trmm(int ps, int batch, const typename TrsmArg<T>::type *args, const blasfeo_amd_hint *hint, hipStream_t st, const char *name)
{
	int rc = check_common(ps, batch, args, sizeof(T) == 4, name);
	if (rc || batch == 0)
		return rc;
	const int slots = tiles_y(hint, TM, TN);
	dim3 grid;
	if ((rc = xcd_grid(batch, slots, NT3, &grid, name)))
		return rc;
	if (ps == 4)
		hipLaunchKernelGGL((k_trmm_rlnn<T, 4>), grid, dim3(NT3), 0, st, args, batch, slots);
	else if (ps == 8)
		hipLaunchKernelGGL((k_trmm_rlnn<T, 8>), grid, dim3(NT3), 0, st, args, batch, slots);
	else
		hipLaunchKernelGGL((k_trmm_rlnn<T, 16>), grid, dim3(NT3), 0, st, args, batch, slots);
	return check_launch(name);
}

// trsm_rltn's instances sorted by n into buckets of 64 columns, so that each
// bucket's launch sizes its LDS row block for that bucket: 64 columns take
// 37 KB (four workgroups per CU), 128 take 71 KB (two), 256 take 137 KB
// (one).  counts[0..TRSM_NB-1] and the queue heads counts[16..] start at 0.
constexpr int TRSM_NB = TRSM_NMAX / 64;
__global__ void k_trsm_zero(int *ws)
{
	if (threadIdx.x < 64)
		ws[threadIdx.x] = 0;
}
template <typename T>
__global__ void k_trsm_plan(const typename TrsmArg<T>::type *__restrict__ args, int batch, int *counts, int *lists)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	const int n = i < batch ? args[i].n : 0;
	plan_bucket<TRSM_NB>(i, i < batch ? (n <= 64 ? 0 : min(TRSM_NB - 1, (n - 1) / 64)) : -1, batch, counts, lists);
}

// the batched entry: plan, then one persistent launch per n bucket, largest
// first (measured on 16384 calls of m, n ~ U{16..256} against one launch
// sized for the hint's n)
template <typename T, int PS>
static int launch_trsm_ps(int batch, const typename TrsmArg<T>::type *args, int npad, hipStream_t st, const char *name)
{
	const size_t wsbytes = sizeof(int) * (64 + (size_t)TRSM_NB * batch);
	int *ws = (int *)workspace(wsbytes, st);
	if (!ws)
		return set_error(BLASFEO_AMD_ENOMEM, "%s: cannot allocate the %zu-byte plan workspace", name, wsbytes);
	int *counts = ws, *heads = ws + 16, *lists = ws + 64;
	hipLaunchKernelGGL(k_trsm_zero, dim3(1), dim3(64), 0, st, ws);
	hipLaunchKernelGGL((k_trsm_plan<T>), dim3((batch + 255) / 256), dim3(256), 0, st, args, batch, counts, lists);
	int rc = check_launch(name);
	if (rc)
		return rc;
	const int cus = num_cus() > 0 ? num_cus() : 256;
	constexpr size_t static_lds = sizeof(T) * (TRSM_NMAX + 16 * 17) + sizeof(int);
	for (int b = TRSM_NB - 1; b >= 0; b--)
	{
		if (64 * b >= npad)
			continue;   // n beyond the hint: no instance runs (as in the unbucketed launch)
		const int np = min(64 * (b + 1), npad);
		const size_t shm = (size_t)np * LDX * sizeof(T);
		// at most four workgroups per CU (registers: four waves per SIMD)
		const int per_cu = max(1, min(4, (int)(163840 / (shm + static_lds))));
		const int grid = min(batch, per_cu * cus);
		if ((rc = lds_check(reinterpret_cast<const void *>(&k_trsm_rltn<T, PS>), shm, name)))
			return rc;
		hipLaunchKernelGGL((k_trsm_rltn<T, PS>), dim3(grid), dim3(NT3), shm, st, args, np, counts + b,
			lists + (size_t)b * batch, heads + b);
		if ((rc = check_launch(name)))
			return rc;
	}
	hipLaunchKernelGGL((k_trsm_dA<T, PS, false>), dim3(batch), dim3(64), 0, st, args);
	return check_launch(name);
}

// the solve of the rows below the diagonal block in the trapezoid Cholesky:
// reciprocals from sA->dA, no cache update (the caller sets use_dA)
template <int PS>
static int launch_trsm_given(int batch, const blasfeo_amd_dtrsm_arg *args, dim3 grid, size_t shm, int nmax, hipStream_t st, const char *name)
{
	hipLaunchKernelGGL((k_trsm_rltn<double, PS, true>), grid, dim3(NT3), shm, st, args, nmax, (const int *)nullptr,
		(const int *)nullptr, (int *)nullptr);
	return check_launch(name);
}

template <typename T>
static int launch_trsm(int ps, int batch, const typename TrsmArg<T>::type *args, const blasfeo_amd_hint *hint, hipStream_t st, const char *name, bool given = false)
{
	int rc = check_common(ps, batch, args, sizeof(T) == 4, name);
	if (rc || batch == 0)
		return rc;
	if (!hint || hint->max_n <= 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: hint->max_n (upper bound of n) is required", name);
	if (hint->max_n > TRSM_NMAX)
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: n up to %d supported, got %d", name, TRSM_NMAX, hint->max_n);
	const int npad = (hint->max_n + 15) / 16 * 16;
	const size_t shm = (size_t)npad * LDX * sizeof(T);
	const int gy = hint->max_m > 0 ? min(64, (hint->max_m + TM - 1) / TM) : 1;
	dim3 grid(batch, gy);
	if (given)
	{
		if (sizeof(T) != 8 || (ps != 4 && ps != 8))
			return set_error(BLASFEO_AMD_EINVAL, "%s: unsupported panel size %d", name, ps);
		const blasfeo_amd_dtrsm_arg *d = reinterpret_cast<const blasfeo_amd_dtrsm_arg *>(args);
		return ps == 4 ? launch_trsm_given<4>(batch, d, grid, shm, npad, st, name)
		               : launch_trsm_given<8>(batch, d, grid, shm, npad, st, name);
	}
	if (ps == 4)
		return launch_trsm_ps<T, 4>(batch, args, npad, st, name);
	if (ps == 8)
		return launch_trsm_ps<T, 8>(batch, args, npad, st, name);
	return launch_trsm_ps<T, 16>(batch, args, npad, st, name);
}

template <typename T, int PS, int MODE>
static int launch_trsml_ps(int batch, const typename TrsmArg<T>::type *args, dim3 grid, size_t shm, int mmax, bool thin, hipStream_t st, const char *name)
{
	if (thin)
		hipLaunchKernelGGL((k_trsm_left_thin<T, PS, MODE>), dim3(batch), dim3(WAVE), 0, st, args);
	else
		hipLaunchKernelGGL((k_trsm_left<T, PS, MODE>), grid, dim3(NT3), shm, st, args, mmax);
	int rc = check_launch(name);
	if (rc)
		return rc;
	hipLaunchKernelGGL((k_trsm_dA<T, PS, true, MODE == TL_LLNU>), dim3(batch), dim3(64), 0, st, args);
	return check_launch(name);
}

template <typename T, int MODE>
static int launch_trsml(int ps, int batch, const typename TrsmArg<T>::type *args, const blasfeo_amd_hint *hint, hipStream_t st, const char *name)
{
	int rc = check_common(ps, batch, args, sizeof(T) == 4, name);
	if (rc || batch == 0)
		return rc;
	if (!hint || hint->max_m <= 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: hint->max_m (upper bound of m) is required", name);
	if (hint->max_m > trsml_mmax<T>())
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: m up to %d supported, got %d", name, trsml_mmax<T>(), hint->max_m);
	const int mmax = hint->max_m;
	const int mpad = (mmax + 15) / 16 * 16;
	const size_t shm = sizeof(T) * ((size_t)LNC * (mpad + 1) + mpad);
	const int gy = hint->max_n > 0 ? min(64, (hint->max_n + LNC - 1) / LNC) : 1;
	dim3 grid(batch, gy);
	// few right-hand sides (hint->max_n <= 4), m <= 256 and a batch that fills
	// the GPU by itself (16 instances per CU): one wavefront per instance with X
	// in registers.  Smaller batches keep the four-wave block sweep, whose
	// per-instance chain is shorter (2048 solves of n = 256 with 4 right-hand
	// sides: 1.34 ms against 1.48 ms one wavefront each).
	// The "trsm_thin" knob forces it on (1) or off (2) where it applies.
	const int cus = num_cus() > 0 ? num_cus() : 256;
	const int knob = tune_value(TUNE_TRSM_THIN, 0);
	const bool thin = sizeof(T) == 8 && hint->max_n > 0 && hint->max_n <= THIN_RHS && mmax <= THIN_Q * WAVE
	                  && (knob == 1 || (knob == 0 && batch >= 16 * cus));
	if (ps == 4)
		return launch_trsml_ps<T, 4, MODE>(batch, args, grid, shm, mmax, thin, st, name);
	if (ps == 8)
		return launch_trsml_ps<T, 8, MODE>(batch, args, grid, shm, mmax, thin, st, name);
	return launch_trsml_ps<T, 16, MODE>(batch, args, grid, shm, mmax, thin, st, name);
}

int trsm_rltn_given_dA(int ps, int batch, const blasfeo_amd_dtrsm_arg *args, const blasfeo_amd_hint *hint, hipStream_t st, const char *name)
{
	return launch_trsm<double>(ps, batch, args, hint, st, name, true);
}

} // namespace bfa

using namespace bfa;

extern "C" {

int blasfeo_amd_dgemm_nt_batched(int ps, int batch, const blasfeo_amd_dgemm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_gemm<double, OP_NT>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dgemm_nt_batched"); }
int blasfeo_amd_dgemm_nn_batched(int ps, int batch, const blasfeo_amd_dgemm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_gemm<double, OP_NN>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dgemm_nn_batched"); }
int blasfeo_amd_dsyrk_ln_batched(int ps, int batch, const blasfeo_amd_dgemm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_gemm<double, OP_SYRK>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dsyrk_ln_batched"); }
int blasfeo_amd_dsyrk_ln_mn_batched(int ps, int batch, const blasfeo_amd_dgemm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_gemm<double, OP_SYRK_MN>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dsyrk_ln_mn_batched"); }
int blasfeo_amd_dtrmm_rlnn_batched(int ps, int batch, const blasfeo_amd_dtrsm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_trmm<double>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dtrmm_rlnn_batched"); }
int blasfeo_amd_dtrsm_rltn_batched(int ps, int batch, const blasfeo_amd_dtrsm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_trsm<double>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dtrsm_rltn_batched"); }
int blasfeo_amd_dtrsm_llnn_batched(int ps, int batch, const blasfeo_amd_dtrsm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_trsml<double, TL_LLNN>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dtrsm_llnn_batched"); }
int blasfeo_amd_dtrsm_lltn_batched(int ps, int batch, const blasfeo_amd_dtrsm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_trsml<double, TL_LLTN>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dtrsm_lltn_batched"); }
int blasfeo_amd_dtrsm_llnu_batched(int ps, int batch, const blasfeo_amd_dtrsm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_trsml<double, TL_LLNU>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dtrsm_llnu_batched"); }
int blasfeo_amd_dtrsm_lunn_batched(int ps, int batch, const blasfeo_amd_dtrsm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_trsml<double, TL_LUNN>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_dtrsm_lunn_batched"); }

int blasfeo_amd_sgemm_nt_batched(int ps, int batch, const blasfeo_amd_sgemm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_gemm<float, OP_NT>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_sgemm_nt_batched"); }
int blasfeo_amd_sgemm_nn_batched(int ps, int batch, const blasfeo_amd_sgemm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_gemm<float, OP_NN>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_sgemm_nn_batched"); }
int blasfeo_amd_ssyrk_ln_batched(int ps, int batch, const blasfeo_amd_sgemm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_gemm<float, OP_SYRK>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_ssyrk_ln_batched"); }
int blasfeo_amd_strsm_rltn_batched(int ps, int batch, const blasfeo_amd_strsm_arg *args, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{ return launch_trsm<float>(ps, batch, args, hint, (hipStream_t)s, "blasfeo_amd_strsm_rltn_batched"); }

} // extern "C"
