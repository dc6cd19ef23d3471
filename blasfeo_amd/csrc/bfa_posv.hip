// bfa_posv.hip -- batched SPD solve A X = B for gfx950: FP32 Cholesky factor
// resident in LDS, FP64 iterative refinement, FP64 fallback on the device.
//
// One workgroup of 8 waves per instance, n <= 256.  The lower triangle of A
// is converted to FP32 into LDS as packed 16x16 tiles (tile (I,J), I >= J, at
// index I(I+1)/2 + J, column-major inside the tile): 136 tiles = 139,264 B
// for n = 256, where the FP64 triangle (272 KB) would not fit.  The factor
// never leaves LDS:
//   per tile column K: wave 0 factors the diagonal tile in FP64 registers
//   (v_readlane broadcasts, reciprocal square roots) and inverts it in FP64;
//   the FP32 inverse Z_K replaces the tile; the panel L(I,K) = A(I,K) Z_K^T
//   and the trailing update A(I,J) -= L(I,K) L(J,K)^T run on the FP32 matrix
//   cores (v_mfma_f32_16x16x4f32); the next diagonal tile is factored while
//   the other waves apply the trailing update (look-ahead of one column).
// Solves with the factor apply Z_K to the 16-row blocks, so each block step
// is a small matrix-vector product instead of a 16-step substitution.
// Refinement: conjugate gradients in FP64 preconditioned by the FP32 factor
// (CG-based iterative refinement): the products A p are formed in FP64 from
// the FP64 A in global memory (the symmetric product from the stored lower
// triangle), the preconditioner solves run in FP32 from LDS, until the
// normwise backward error max_c |b_c - A x_c|/(|A| |x_c| + |b_c|) (infinity
// norms) of the true residual is <= tol.  Against stationary refinement this
// keeps converging when cond(A) * eps32 approaches 1 (cond ~ 1e8).
// Instances whose FP32 factor meets a non-positive pivot, or whose iteration
// breaks down or does not converge, are solved again by the FP64 kernels
// (potrf_l + trsm_llnn + trsm_lltn) in the same stream.
#include "bfa_chol_core.hpp"
#include "bfa_internal.hpp"

namespace bfa {

constexpr int PV_NT = 512;
constexpr int PV_NW = PV_NT / WAVE;
constexpr int PV_NMAX = 256;
constexpr int PV_R = 4;  // right-hand sides per refinement pass

__device__ __forceinline__ int tix(int I, int J) { return I * (I + 1) / 2 + J; }

struct PosvLds
{
	float *F;      // packed tiles
	double *X;     // [n16][4] current solution (FP64)
	float *R;      // [n16][4] right-hand side / residual of the FP32 solves
	double *S;     // [n16][5] scratch: residual partial sums; factor: L_KK, 1/l_jj
	double *red;   // [PV_NW][12] reductions
	int *flag;     // [4]
};

__host__ __device__ inline size_t posv_lds_bytes(int nb)
{
	const size_t ntile = (size_t)nb * (nb + 1) / 2;
	return ntile * 256 * sizeof(float) + (size_t)nb * 16 * PV_R * (sizeof(double) + sizeof(float))
		+ (size_t)nb * 16 * 5 * sizeof(double) + PV_NW * 12 * sizeof(double) + 4 * sizeof(int);
}

__device__ inline PosvLds posv_carve(unsigned char *base, int nb)
{
	PosvLds L;
	const size_t ntile = (size_t)nb * (nb + 1) / 2;
	L.F = reinterpret_cast<float *>(base);
	unsigned char *p = base + ntile * 256 * sizeof(float);
	L.X = reinterpret_cast<double *>(p);
	p += (size_t)nb * 16 * PV_R * sizeof(double);
	L.S = reinterpret_cast<double *>(p);
	p += (size_t)nb * 16 * 5 * sizeof(double);
	L.red = reinterpret_cast<double *>(p);
	p += PV_NW * 12 * sizeof(double);
	L.R = reinterpret_cast<float *>(p);
	p += (size_t)nb * 16 * PV_R * sizeof(float);
	L.flag = reinterpret_cast<int *>(p);
	return L;
}

// Block-wide maximum of NV values per thread (every thread gets the result).
template <int NV>
__device__ __forceinline__ void block_max(double *v, double *red)
{
	const int lane = lane_id(), w = wave_id();
#pragma unroll
	for (int q = 0; q < NV; q++)
		v[q] = wave_max_d(v[q]);
	__syncthreads();
	if (lane == WAVE - 1)
#pragma unroll
		for (int q = 0; q < NV; q++)
			red[w * 12 + q] = v[q];
	__syncthreads();
#pragma unroll
	for (int q = 0; q < NV; q++)
	{
		double r = red[q];
		for (int ww = 1; ww < PV_NW; ww++)
			r = fmax(r, red[ww * 12 + q]);
		v[q] = r;
	}
}

// Block-wide sum of NV values per thread (every thread gets the same result:
// one lane per wave contributes its wave's total, summed in wave order).
template <int NV>
__device__ __forceinline__ void block_sum(double *v, double *red)
{
	const int lane = lane_id(), w = wave_id();
#pragma unroll
	for (int q = 0; q < NV; q++)
		v[q] = wave_sum_d(v[q]);
	__syncthreads();
	if (lane == WAVE - 1)
#pragma unroll
		for (int q = 0; q < NV; q++)
			red[w * 12 + q] = v[q];
	__syncthreads();
#pragma unroll
	for (int q = 0; q < NV; q++)
	{
		double r = red[q];
		for (int ww = 1; ww < PV_NW; ww++)
			r += red[ww * 12 + q];
		v[q] = r;
	}
}

// Y = A V in FP64 on the vector units, A symmetric from its stored lower
// triangle in global memory (16x16 tiles), V = V[j*4 + c] in LDS, Y[i*4 + c]
// and (ABS) the row sums of |A| in rowabs[i], both in LDS.  Wave w owns tile
// rows and columns {w, 15 - w} (17 tiles of the direct products A_IJ V_J, 15
// of the transposed products V_K^T A_KI for nb = 16), so the two partial sums
// of a row meet in one wave without atomics and the result does not depend on
// timing.  Lane (i, q) = (l % 16, l / 16) takes row i and columns 4u + q of
// every tile: the direct products read a tile as 16-row x 4-column slabs and
// the transposed ones as 4-row x 16-column slabs (a contiguous 512-byte panel
// run at ps = 4); partial sums stay in registers across the whole tile row
// and are reduced over q once per row block.  With four right-hand sides the
// vector units do no padded work, where 16x16x4 matrix-core products would
// waste three quarters.  Ends with a barrier.
template <int PS, bool ABS>
__device__ void posv_matvec(const PM<double, PS> gA, int ai, int aj, int n, int nb,
	const double *V, double *Y, double *rowabs)
{
	const int lane = lane_id(), w = wave_id();
	const int i = lane & 15, q = lane >> 4;
	for (int sel = 0; sel < 2; sel++)
	{
		const int I = sel == 0 ? w : 2 * PV_NW - 1 - w;
		if (I >= nb)
			continue;
		double p[PV_R] = {0, 0, 0, 0};
		double sa = 0.0;
		const int r = 16 * I + i;
		// direct: row r of A_IJ times V_J, J <= I (the diagonal tile completed symmetrically)
		for (int J = 0; J <= I; J++)
		{
			double a[4];
#pragma unroll
			for (int u = 0; u < 4; u++)
			{
				const int cidx = 16 * J + 4 * u + q;
				const bool up = cidx > r;
				const int rr = min(up ? cidx : r, n - 1), cc = min(up ? r : cidx, n - 1);
				const double e = gA.at(ai + rr, aj + cc);
				a[u] = (r < n && cidx < n) ? e : 0.0;
			}
#pragma unroll
			for (int u = 0; u < 4; u++)
			{
				const double *v = V + (16 * J + 4 * u + q) * PV_R;
#pragma unroll
				for (int c = 0; c < PV_R; c++)
					p[c] += a[u] * v[c];
				if (ABS)
					sa += fabs(a[u]);
			}
		}
		// transposed: column r of A_KI (= row r of A_IK) times V_K, K > I
		for (int K = I + 1; K < nb; K++)
		{
			double bt[4];
#pragma unroll
			for (int u = 0; u < 4; u++)
			{
				const int rk = 16 * K + 4 * u + q;
				const double e = gA.at(ai + min(rk, n - 1), aj + min(r, n - 1));
				bt[u] = (rk < n && r < n) ? e : 0.0;
			}
#pragma unroll
			for (int u = 0; u < 4; u++)
			{
				const double *v = V + (16 * K + 4 * u + q) * PV_R;
#pragma unroll
				for (int c = 0; c < PV_R; c++)
					p[c] += bt[u] * v[c];
				if (ABS)
					sa += fabs(bt[u]);
			}
		}
		// reduce over q (lanes i, i + 16, i + 32, i + 48), fixed order
#pragma unroll
		for (int c = 0; c < PV_R; c++)
		{
			p[c] += __shfl_xor(p[c], 16, WAVE);
			p[c] += __shfl_xor(p[c], 32, WAVE);
		}
		if (ABS)
		{
			sa += __shfl_xor(sa, 16, WAVE);
			sa += __shfl_xor(sa, 32, WAVE);
		}
		const double mine = q == 0 ? p[0] : (q == 1 ? p[1] : (q == 2 ? p[2] : p[3]));
		Y[r * PV_R + q] = mine;
		if (ABS && q == 0)
			rowabs[r] = sa;
	}
	__syncthreads();
}

// Wave 0: FP64 Cholesky of the diagonal tile FKK (FP32 in LDS), replaced by
// Z = L_KK^{-1} in FP32.  Lanes 0..15 hold row `lane`; L(jj,j) is broadcast
// with v_readlane.  Sets *flag on a non-positive (or NaN) pivot.
__device__ __forceinline__ void posv_diag(float *FKK, int *flag)
{
	const int lane = lane_id();
	double x[16];
#pragma unroll
	for (int j = 0; j < 16; j++)
		x[j] = lane < 16 ? (double)FKK[j * 16 + lane] : 0.0;
	double iv[16];
	bool bad = false;
	// no per-lane guards: the upper triangle (lanes < j) is updated too, but
	// only the lower triangle is read below
#pragma unroll
	for (int j = 0; j < 16; j++)
	{
		const double c = readlane(x[j], j);
		const bool ok = c > 0.0;
		bad = bad || !ok;
		const double inv = ok ? rsqrt_fast(c) : 0.0;
		iv[j] = inv;
		x[j] *= inv;   // lane j holds c: c * inv, without a per-column lane mask
#pragma unroll
		for (int jj = j + 1; jj < 16; jj++)
		{
			const double l = readlane(x[j], jj);
			x[jj] -= x[j] * l;
		}
	}
	// Z = L_KK^{-1}: lane c computes column c by forward substitution
	double z[16];
#pragma unroll
	for (int i = 0; i < 16; i++)
	{
		double s = (i == lane) ? 1.0 : 0.0;
#pragma unroll
		for (int k = 0; k < i; k++)
			s -= readlane(x[k], i) * z[k];
		z[i] = s * iv[i];
	}
	wave_sync();
	if (lane < 16)
#pragma unroll
		for (int i = 0; i < 16; i++)
			FKK[lane * 16 + i] = (float)z[i];
	if (bad && lane == 0)
		*flag = 1;
}

// In-LDS Cholesky of the nb x nb tile matrix, right-looking with a look-ahead
// of one tile column: in step K all waves form the panel L(I,K) = A(I,K) Z_K^T;
// then wave 0 updates the next diagonal tile and factors it (the FP64 chain)
// while waves 1.. apply the rank-16 update to the rest of the trailing
// matrix.  Two workgroup barriers per tile column.  Returns (uniformly) false
// if a pivot was not positive.
__device__ bool posv_factor(const PosvLds &L, int nb)
{
	typedef f4_t V4;
	const int lane = lane_id();
	const int w = wave_id();
	if (threadIdx.x == 0)
		L.flag[0] = 0;
	__syncthreads();
	if (w == 0)
		posv_diag(L.F, L.flag);
	__syncthreads();
	for (int K = 0; K < nb; K++)
	{
		const float *FKK = L.F + tix(K, K) * 256;
		// panel: L(I,K) = A(I,K) Z_K^T
		for (int I = K + 1 + w; I < nb; I += PV_NW)
		{
			float *FIK = L.F + tix(I, K) * 256;
			V4 acc = V4{0, 0, 0, 0};
#pragma unroll
			for (int kk = 0; kk < 16; kk += 4)
			{
				const int o = (kk + (lane >> 4)) * 16 + (lane & 15);
				acc = mfma16(FIK[o], FKK[o], acc);
			}
			*reinterpret_cast<V4 *>(FIK + (lane & 15) * 16 + 4 * (lane >> 4)) = acc;
		}
		__syncthreads();
		if (K + 1 >= nb)
			break;
		if (w == 0)
		{
			// next diagonal tile: update, then factor
			float *FN = L.F + tix(K + 1, K + 1) * 256;
			const float *FNK = L.F + tix(K + 1, K) * 256;
			V4 acc = *reinterpret_cast<const V4 *>(FN + (lane & 15) * 16 + 4 * (lane >> 4));
#pragma unroll
			for (int kk = 0; kk < 16; kk += 4)
			{
				const int o = (kk + (lane >> 4)) * 16 + (lane & 15);
				acc = mfma16(-FNK[o], FNK[o], acc);
			}
			*reinterpret_cast<V4 *>(FN + (lane & 15) * 16 + 4 * (lane >> 4)) = acc;
			wave_sync();
			posv_diag(FN, L.flag);
		}
		else
		{
			// trailing update without tile (K+1, K+1): tasks (J, group of up to 4 tiles I >= J)
			int ntask = 0;
			for (int J = K + 1; J < nb; J++)
				ntask += (nb - J + 3) / 4;
			for (int task = w - 1; task < ntask; task += PV_NW - 1)
			{
				int J = K + 1, rem = task;
				while (rem >= (nb - J + 3) / 4)
				{
					rem -= (nb - J + 3) / 4;
					J++;
				}
				const int g = J + 4 * rem;
				const int t0 = (J == K + 1 && g == K + 1) ? 1 : 0;   // skip the diagonal tile wave 0 owns
				const float *FJK = L.F + tix(J, K) * 256;
				V4 acc[4];
#pragma unroll
				for (int t = 0; t < 4; t++)
					if (t >= t0 && g + t < nb)
						acc[t] = *reinterpret_cast<const V4 *>(L.F + tix(g + t, J) * 256 + (lane & 15) * 16 + 4 * (lane >> 4));
#pragma unroll
				for (int kk = 0; kk < 16; kk += 4)
				{
					const int o = (kk + (lane >> 4)) * 16 + (lane & 15);
					const float b = FJK[o];
#pragma unroll
					for (int t = 0; t < 4; t++)
						if (t >= t0 && g + t < nb)
							acc[t] = mfma16(-L.F[tix(g + t, K) * 256 + o], b, acc[t]);
				}
#pragma unroll
				for (int t = 0; t < 4; t++)
					if (t >= t0 && g + t < nb)
						*reinterpret_cast<V4 *>(L.F + tix(g + t, J) * 256 + (lane & 15) * 16 + 4 * (lane >> 4)) = acc[t];
			}
		}
		__syncthreads();
	}
	return L.flag[0] == 0;
}

// R <- (L L^T)^{-1} R in FP32 with the factor in LDS (diagonal tiles hold
// Z_K = L_KK^{-1}).  One workgroup barrier per 16-row block: in step K wave 0
// applies the previous block's result to block K and multiplies by Z_K (the
// chain), while waves 1.. apply the previous block's result to the blocks
// beyond K.  Starts and ends with a barrier.
__device__ void posv_solve(const PosvLds &L, int nb)
{
	const int t = threadIdx.x;
	float *R = L.R;
	__syncthreads();
	// forward: y_K = Z_K (r_K - L(K,K-1) y_{K-1}); blocks I > K take -= L(I,K-1) y_{K-1}
	for (int K = 0; K < nb; K++)
	{
		if (t < 64)
		{
			const int i = t & 15, c = t >> 4;
			float v = R[(16 * K + i) * PV_R + c];
			if (K > 0)
			{
				const float *LK = L.F + tix(K, K - 1) * 256 + i;
#pragma unroll
				for (int k = 0; k < 16; k++)
					v -= LK[k * 16] * R[(16 * (K - 1) + k) * PV_R + c];
			}
			R[(16 * K + i) * PV_R + c] = v;
			wave_sync();
			const float *Z = L.F + tix(K, K) * 256;
			float s = 0.0f;
#pragma unroll
			for (int k = 0; k < 16; k++)
				s += Z[k * 16 + i] * R[(16 * K + k) * PV_R + c];
			wave_sync();
			R[(16 * K + i) * PV_R + c] = s;
		}
		else if (K > 0)
		{
			const int np = (nb - K - 1) * 16 * PV_R;
			for (int p = t - 64; p < np; p += PV_NT - 64)
			{
				const int c = p & 3, ir = 16 * (K + 1) + (p >> 2);
				const float *LI = L.F + tix(ir >> 4, K - 1) * 256 + (ir & 15);
				float s = 0.0f;
#pragma unroll
				for (int k = 0; k < 16; k++)
					s += LI[k * 16] * R[(16 * (K - 1) + k) * PV_R + c];
				R[ir * PV_R + c] -= s;
			}
		}
		__syncthreads();
	}
	// backward: x_K = Z_K^T (y_K - L(K+1,K)^T x_{K+1}); blocks J < K take -= L(K+1,J)^T x_{K+1}
	for (int K = nb - 1; K >= 0; K--)
	{
		if (t < 64)
		{
			const int i = t & 15, c = t >> 4;
			float v = R[(16 * K + i) * PV_R + c];
			if (K < nb - 1)
			{
				const float *LK = L.F + tix(K + 1, K) * 256 + i * 16;
#pragma unroll
				for (int k = 0; k < 16; k++)
					v -= LK[k] * R[(16 * (K + 1) + k) * PV_R + c];
			}
			R[(16 * K + i) * PV_R + c] = v;
			wave_sync();
			const float *Z = L.F + tix(K, K) * 256;
			float s = 0.0f;
#pragma unroll
			for (int k = 0; k < 16; k++)
				s += Z[i * 16 + k] * R[(16 * K + k) * PV_R + c];
			wave_sync();
			R[(16 * K + i) * PV_R + c] = s;
		}
		else if (K < nb - 1)
		{
			const int np = K * 16 * PV_R;
			for (int p = t - 64; p < np; p += PV_NT - 64)
			{
				const int c = p & 3, ir = p >> 2;
				const float *LK = L.F + tix(K + 1, ir >> 4) * 256 + (ir & 15) * 16;
				float s = 0.0f;
#pragma unroll
				for (int k = 0; k < 16; k++)
					s += LK[k] * R[(16 * (K + 1) + k) * PV_R + c];
				R[ir * PV_R + c] -= s;
			}
		}
		__syncthreads();
	}
}

template <int PS>
__global__ __launch_bounds__(PV_NT) void k_posv_mixed(const blasfeo_amd_dposv_arg *__restrict__ args, int nmax,
	double tol, int max_iter, int *__restrict__ path_out)
{
	extern __shared__ unsigned char smem_raw[];
	const blasfeo_amd_dposv_arg &a = args[blockIdx.x];
	const int n = a.n, nrhs = a.nrhs;
	const int t = threadIdx.x;
	const int lane = lane_id(), w = wave_id();
	blasfeo_amd_solve_info *info = a.info;
	if (n <= 0 || nrhs <= 0)
	{
		if (t == 0)
		{
			path_out[blockIdx.x] = 0;
			if (info)
				*info = blasfeo_amd_solve_info{0, 0, 0.0};
		}
		return;
	}
	if (n > nmax)
	{
		if (t == 0)
		{
			path_out[blockIdx.x] = 1;
			if (info)
				*info = blasfeo_amd_solve_info{0, 1, -1.0};
		}
		return;
	}
	const int nb = (n + 15) / 16;
	const int n16 = nb * 16;
	const PosvLds L = posv_carve(smem_raw, nb);
	const PM<double, PS> gA = view<double, PS>(a.sA);
	const PM<double, PS> gB = view<double, PS>(a.sB);
	PM<double, PS> gX = view<double, PS>(a.sX);
	const int ai = a.ai, aj = a.aj;

	// A (lower triangle, FP64) -> FP32 tiles; identity beyond n.  Wave w takes
	// tiles w, w + 8, ..., four per step with all sixteen loads issued first.
	{
		const int ntile = nb * (nb + 1) / 2;
		const int i = lane & 15;
		for (int t0 = w; t0 < ntile; t0 += 4 * PV_NW)
		{
			double v[4][4];
			int tI[4], tJ[4];
#pragma unroll
			for (int u = 0; u < 4; u++)
			{
				const int t = t0 + u * PV_NW;
				int I = 0;
				while ((I + 1) * (I + 2) / 2 <= t)
					I++;
				tI[u] = I;
				tJ[u] = t - I * (I + 1) / 2;
				const int r = I * 16 + i;
#pragma unroll
				for (int k = 0; k < 4; k++)
				{
					const int c = tJ[u] * 16 + 4 * k + (lane >> 4);
					const bool in = t < ntile && r < n && c < n && r >= c;
					double e = (r == c) ? 1.0 : 0.0;
					if (in)
						e = gA.at(ai + r, aj + c);
					v[u][k] = e;
				}
			}
#pragma unroll
			for (int u = 0; u < 4; u++)
			{
				const int t = t0 + u * PV_NW;
				if (t < ntile)
				{
					float *FT = L.F + t * 256;
#pragma unroll
					for (int k = 0; k < 4; k++)
						FT[(4 * k + (lane >> 4)) * 16 + i] = (float)v[u][k];
				}
			}
		}
	}
	const bool ok = posv_factor(L, nb);

	// preconditioned conjugate gradients in FP64 per group of 4 right-hand
	// sides, preconditioner M = L L^T with the FP32 factor in LDS; the row
	// owner (t < 256, row t) keeps x, r, p of its row in registers, p is also
	// published in LDS (L.X) for the products A p.  A converged recursive
	// residual is confirmed with the true residual b - A x before stopping;
	// otherwise CG restarts from the true residual.
	const int i = t & 255;
	const bool owner = t < 256 && i < n;
	int iters = 0;
	double berr = 0.0;
	bool conv = ok;
	double anorm = -1.0;
	for (int c0 = 0; conv && c0 < nrhs; c0 += PV_R)
	{
		const int nr = min(PV_R, nrhs - c0);
		double xr[PV_R] = {0, 0, 0, 0}, rr[PV_R] = {0, 0, 0, 0}, pr[PV_R] = {0, 0, 0, 0};
		double bq[PV_R] = {0, 0, 0, 0};
		double v[PV_R];
		if (owner)
#pragma unroll
			for (int c = 0; c < PV_R; c++)
				bq[c] = rr[c] = c < nr ? gB.at(a.bi + i, a.bj + c0 + c) : 0.0;
#pragma unroll
		for (int c = 0; c < PV_R; c++)
			v[c] = fabs(bq[c]);
		block_max<PV_R>(v, L.red);
		double bmax[PV_R];
#pragma unroll
		for (int c = 0; c < PV_R; c++)
			bmax[c] = v[c];
		double rho[PV_R] = {0, 0, 0, 0};
		bool restart = true;
		bool fail = false;
		int it = 0;
		for (;;)
		{
			// z = M^{-1} r (FP32); rho' = r.z; p = z + beta p
			if (t < 256 && i < n16)
#pragma unroll
				for (int c = 0; c < PV_R; c++)
					L.R[i * PV_R + c] = owner ? (float)rr[c] : 0.0f;
			posv_solve(L, nb);
			double z[PV_R] = {0, 0, 0, 0};
			if (owner)
#pragma unroll
				for (int c = 0; c < PV_R; c++)
					z[c] = (double)L.R[i * PV_R + c];
#pragma unroll
			for (int c = 0; c < PV_R; c++)
				v[c] = rr[c] * z[c];
			block_sum<PV_R>(v, L.red);
#pragma unroll
			for (int c = 0; c < PV_R; c++)
			{
				const double beta = (restart || rho[c] == 0.0) ? 0.0 : v[c] / rho[c];
				rho[c] = v[c];
				pr[c] = z[c] + beta * pr[c];
			}
			restart = false;
			if (t < 256 && i < n16)
#pragma unroll
				for (int c = 0; c < PV_R; c++)
					L.X[i * PV_R + c] = owner ? pr[c] : 0.0;
			__syncthreads();
			// q = A p ; alpha = rho / p.q ; x += alpha p ; r -= alpha q
			double qv[PV_R] = {0, 0, 0, 0};
			if (anorm < 0.0)
			{
				posv_matvec<PS, true>(gA, ai, aj, n, nb, L.X, L.S, L.S + n16 * PV_R);
				double an = owner ? L.S[n16 * PV_R + i] : 0.0;
				block_max<1>(&an, L.red);
				anorm = an;
			}
			else
				posv_matvec<PS, false>(gA, ai, aj, n, nb, L.X, L.S, nullptr);
			if (owner)
#pragma unroll
				for (int c = 0; c < PV_R; c++)
					qv[c] = L.S[i * PV_R + c];
#pragma unroll
			for (int c = 0; c < PV_R; c++)
				v[c] = pr[c] * qv[c];
			block_sum<PV_R>(v, L.red);
#pragma unroll
			for (int c = 0; c < PV_R; c++)
			{
				if (c < nr && rho[c] != 0.0 && !(v[c] > 0.0))
					fail = true;  // p^T A p <= 0 (or NaN): not SPD at this precision
				const double alpha = (c < nr && rho[c] != 0.0 && v[c] > 0.0) ? rho[c] / v[c] : 0.0;
				xr[c] += alpha * pr[c];
				rr[c] -= alpha * qv[c];
			}
			it++;
			if (fail)
				break;
			// normwise backward error of the recursive residual
			double m8[2 * PV_R];
#pragma unroll
			for (int c = 0; c < PV_R; c++)
			{
				m8[c] = fabs(rr[c]);
				m8[PV_R + c] = fabs(xr[c]);
			}
			block_max<2 * PV_R>(m8, L.red);
			double eta = 0.0;
#pragma unroll
			for (int c = 0; c < PV_R; c++)
				if (c < nr)
				{
					const double den = anorm * m8[PV_R + c] + bmax[c];
					eta = fmax(eta, den > 0.0 ? m8[c] / den : (m8[c] > 0.0 ? 1e300 : 0.0));
				}
			if (!(eta <= tol) && eta == eta && it < max_iter)
				continue;
			// confirm with the true residual b - A x
			if (t < 256 && i < n16)
#pragma unroll
				for (int c = 0; c < PV_R; c++)
					L.X[i * PV_R + c] = owner ? xr[c] : 0.0;
			__syncthreads();
			double ax[PV_R] = {0, 0, 0, 0};
			posv_matvec<PS, false>(gA, ai, aj, n, nb, L.X, L.S, nullptr);
			if (owner)
#pragma unroll
				for (int c = 0; c < PV_R; c++)
					ax[c] = L.S[i * PV_R + c];
#pragma unroll
			for (int c = 0; c < PV_R; c++)
			{
				rr[c] = owner ? bq[c] - ax[c] : 0.0;
				m8[c] = fabs(rr[c]);
				m8[PV_R + c] = fabs(xr[c]);
			}
			block_max<2 * PV_R>(m8, L.red);
			double eta_t = 0.0;
#pragma unroll
			for (int c = 0; c < PV_R; c++)
				if (c < nr)
				{
					const double den = anorm * m8[PV_R + c] + bmax[c];
					eta_t = fmax(eta_t, den > 0.0 ? m8[c] / den : (m8[c] > 0.0 ? 1e300 : 0.0));
				}
			if (eta_t <= tol)
			{
				berr = fmax(berr, eta_t);
				break;
			}
			if (it >= max_iter || !(eta_t == eta_t))
			{
				fail = true;
				break;
			}
			restart = true;
		}
		iters = max(iters, it);
		if (fail)
		{
			conv = false;
			break;
		}
		if (owner)
#pragma unroll
			for (int c = 0; c < PV_R; c++)
				if (c < nr)
					gX.at(a.xi + i, a.xj + c0 + c) = xr[c];
		__syncthreads();
	}
	if (t == 0)
	{
		path_out[blockIdx.x] = conv ? 0 : 1;
		if (info)
			*info = conv ? blasfeo_amd_solve_info{iters, 0, berr} : blasfeo_amd_solve_info{iters, 1, -1.0};
	}
	if (t == 0)
		a.sX->use_dA = 0;
}

// Records for the FP64 re-solve of the instances that did not converge:
// potrf_l(A -> L), trsm_llnn(L, B -> X), trsm_lltn(L, X -> X); size 0 (a
// no-op) for the others.
__global__ void k_posv_fallback_args(const blasfeo_amd_dposv_arg *__restrict__ args, const int *__restrict__ path,
	int batch, blasfeo_amd_dfact_arg *fa, blasfeo_amd_dtrsm_arg *t1, blasfeo_amd_dtrsm_arg *t2)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= batch)
		return;
	const blasfeo_amd_dposv_arg &a = args[b];
	const int n = (path[b] == 1 && a.nrhs > 0) ? a.n : 0;
	fa[b] = blasfeo_amd_dfact_arg{n, n, a.ai, a.aj, a.li, a.lj, a.sA, a.sL, nullptr};
	t1[b] = blasfeo_amd_dtrsm_arg{n, a.nrhs, a.li, a.lj, a.bi, a.bj, a.xi, a.xj, 1.0, a.sL, a.sB, a.sX};
	t2[b] = blasfeo_amd_dtrsm_arg{n, a.nrhs, a.li, a.lj, a.xi, a.xj, a.xi, a.xj, 1.0, a.sL, a.sX, a.sX};
}

} // namespace bfa

using namespace bfa;

extern "C" int blasfeo_amd_dposv_mixed_batched(int ps, int batch, const blasfeo_amd_dposv_arg *args, double tol,
	int max_iter, const blasfeo_amd_hint *hint, blasfeo_amd_stream_t s)
{
	const char *name = "blasfeo_amd_dposv_mixed_batched";
	hipStream_t st = (hipStream_t)s;
	if (batch < 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: batch=%d < 0", name, batch);
	if (batch > 0 && !args)
		return set_error(BLASFEO_AMD_EINVAL, "%s: args is NULL", name);
	if (ps != 4 && ps != 8)
		return set_error(BLASFEO_AMD_EINVAL, "%s: unsupported panel size %d", name, ps);
	if (!(tol > 0.0) || max_iter < 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: tol must be > 0 and max_iter >= 0", name);
	if (!hint || hint->max_m <= 0)
		return set_error(BLASFEO_AMD_EINVAL, "%s: hint->max_m (upper bound of n) is required", name);
	if (hint->max_m > fact_mmax<double>())
		return set_error(BLASFEO_AMD_ETOOBIG, "%s: n up to %d supported, got %d", name, fact_mmax<double>(), hint->max_m);
	if (batch == 0)
		return BLASFEO_AMD_OK;
	const int nmax = min(hint->max_m, PV_NMAX);
	const int nb = (nmax + 15) / 16;
	const size_t shm = posv_lds_bytes(nb);
	// workspace: per-instance path, then the fallback records
	const size_t off_fa = ((size_t)batch * sizeof(int) + 255) / 256 * 256;
	const size_t off_t1 = off_fa + ((size_t)batch * sizeof(blasfeo_amd_dfact_arg) + 255) / 256 * 256;
	const size_t off_t2 = off_t1 + ((size_t)batch * sizeof(blasfeo_amd_dtrsm_arg) + 255) / 256 * 256;
	const size_t bytes = off_t2 + (size_t)batch * sizeof(blasfeo_amd_dtrsm_arg);
	unsigned char *ws = static_cast<unsigned char *>(workspace(bytes, st));
	if (!ws)
		return set_error(BLASFEO_AMD_ENOMEM, "%s: workspace of %zu bytes not available", name, bytes);
	int *path = reinterpret_cast<int *>(ws);
	auto *fa = reinterpret_cast<blasfeo_amd_dfact_arg *>(ws + off_fa);
	auto *t1 = reinterpret_cast<blasfeo_amd_dtrsm_arg *>(ws + off_t1);
	auto *t2 = reinterpret_cast<blasfeo_amd_dtrsm_arg *>(ws + off_t2);
	if (ps == 4)
		hipLaunchKernelGGL((k_posv_mixed<4>), dim3(batch), dim3(PV_NT), shm, st, args, nmax, tol, max_iter, path);
	else
		hipLaunchKernelGGL((k_posv_mixed<8>), dim3(batch), dim3(PV_NT), shm, st, args, nmax, tol, max_iter, path);
	int rc = check_launch(name);
	if (rc)
		return rc;
	hipLaunchKernelGGL(k_posv_fallback_args, dim3((batch + 255) / 256), dim3(256), 0, st, args, path, batch, fa, t1, t2);
	if ((rc = check_launch(name)))
		return rc;
	blasfeo_amd_hint h2 = {hint->max_m, hint->max_m, 0};
	if ((rc = blasfeo_amd_dpotrf_l_batched(ps, batch, fa, &h2, s)))
		return rc;
	blasfeo_amd_hint h3 = {hint->max_m, hint->max_n > 0 ? hint->max_n : 64, 0};
	if ((rc = blasfeo_amd_dtrsm_llnn_batched(ps, batch, t1, &h3, s)))
		return rc;
	return blasfeo_amd_dtrsm_lltn_batched(ps, batch, t2, &h3, s);
}
