// bfa_common.hpp -- device building blocks shared by the gfx950 kernels.
//
// Panel-major addressing follows include/blasfeo_common.h:112: element (i,j)
// of a matrix with panel size PS and panel stride cn is at
// pA[(i - i%PS)*cn + j*PS + i%PS].  All indices below are absolute (the
// caller's row/column offsets already added).
//
// Matrix-core (MFMA) operand layout for the 16x16x4 shapes used here
// (v_mfma_f64_16x16x4f64, v_mfma_f32_16x16x4f32), wave of 64 lanes:
//   A (16x4):  lane l holds A[l%16][l/16]
//   B (4x16):  lane l holds B[l/16][l%16]
//   D (16x16): lane l, register r holds D[acc_row(l, r)][l%16] with
//              acc_row = l/16 + 4*r for f64 and 4*(l/16) + r for f32
//              (measured on gfx950 with a known-answer product).
#pragma once

// Address space of global memory for the panel-major views (the CPU emulator of
// the kernels defines it empty).
#ifndef BFA_GLOBAL
#define BFA_GLOBAL __attribute__((address_space(1)))
#endif
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/blasfeo_amd.h"

namespace bfa {

constexpr int WAVE = 64;

// MI355X has 8 XCDs (each with its own L2); the dispatcher deals the
// workgroups of a launch to them round-robin in workgroup-id order.
constexpr int XCDS = 8;

// Logical index of this workgroup in a 1-D launch of `total` workgroups
// (`total` a multiple of XCDS) such that each XCD gets one contiguous range
// of logical indices, taken in order: workgroups that share data (the tiles
// of one instance) then run on one XCD, close in time, and hit its L2.
__device__ __forceinline__ unsigned xcd_local_index(unsigned g, unsigned total)
{
	return (g % XCDS) * (total / XCDS) + g / XCDS;
}

// (instance, slot) of this workgroup in a 1-D launch of batch * slots
// workgroups (rounded up to a multiple of XCDS): XCD-aware (the slots of an
// instance consecutive on one XCD) or instance-fastest (slot s of every
// instance, then slot s + 1); inst >= batch for the rounding workgroups.
__device__ __forceinline__ void batch_slot(bool xcd, int batch, int slots, int &inst, int &slot)
{
	if (xcd)
	{
		const unsigned li = xcd_local_index(blockIdx.x, gridDim.x);
		inst = (int)(li / (unsigned)slots);
		slot = (int)(li % (unsigned)slots);
	}
	else
	{
		inst = (int)(blockIdx.x % (unsigned)batch);
		slot = (int)(blockIdx.x / (unsigned)batch);
		if (slot >= slots)
			inst = batch;
	}
}

// Size plans: instance i joins bucket b (b < 0: none), lists[b * batch + k]
// = i for k < counts[b].  The workgroup first counts its own instances per
// bucket with LDS atomics, reserves one range per bucket with a single global
// atomic, then writes its entries, so the global counters see one atomic per
// workgroup and bucket rather than one per instance (per-instance atomics on a
// few counters serialise: tens of microseconds for some ten thousand
// instances).  Every thread of the workgroup calls it.
template <int NB>
__device__ __forceinline__ void plan_bucket(int i, int b, int batch, int *counts, int *lists)
{
	__shared__ int local[NB], base[NB];
	if (threadIdx.x < NB)
		local[threadIdx.x] = 0;
	__syncthreads();
	const int k = b >= 0 ? atomicAdd(&local[b], 1) : 0;
	__syncthreads();
	if (threadIdx.x < NB)
		base[threadIdx.x] = local[threadIdx.x] > 0 ? atomicAdd(&counts[threadIdx.x], local[threadIdx.x]) : 0;
	__syncthreads();
	if (b >= 0)
		lists[(size_t)b * batch + base[b] + k] = i;
}

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T> struct Vec4;
template <> struct Vec4<double> { typedef d4_t type; };
template <> struct Vec4<float> { typedef f4_t type; };

template <typename T> struct MatOf;
template <> struct MatOf<double> { typedef blasfeo_dmat type; };
template <> struct MatOf<float> { typedef blasfeo_smat type; };

// one 16x16x4 matrix-core step: acc += a * b (layout above)
__device__ __forceinline__ d4_t mfma16(double a, double b, d4_t acc)
{
	return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma16(float a, float b, f4_t acc)
{
	return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
}

// Row of the 16x16 accumulator tile that register r of lane l holds.
template <typename T> __device__ __forceinline__ int acc_row(int l, int r);
template <> __device__ __forceinline__ int acc_row<double>(int l, int r) { return (l >> 4) + 4 * r; }
template <> __device__ __forceinline__ int acc_row<float>(int l, int r) { return 4 * (l >> 4) + r; }

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / WAVE; }

// Keeps the compiler from moving LDS/global accesses across this point; waves
// execute their own memory operations in order, so this is the only ordering
// needed when a single wave hands data between its lanes through LDS.
__device__ __forceinline__ void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Panel-major matrix view.
template <typename T, int PS>
struct PM
{
	T *p;
	int cn;
	__device__ __forceinline__ size_t off(int i, int j) const
	{
		return (size_t)(i & ~(PS - 1)) * (size_t)cn + (size_t)j * PS + (i & (PS - 1));
	}
	// through a global-memory pointer: global_load/store instead of flat ones,
	// which also count against the LDS counter (an LDS wait then waits for the
	// wave's global loads too)
	__device__ __forceinline__ BFA_GLOBAL T &at(int i, int j) const { return ((BFA_GLOBAL T *)p)[off(i, j)]; }
};

template <typename T, int PS>
__device__ __forceinline__ PM<T, PS> view(const typename MatOf<T>::type *s)
{
	// the descriptor read as global memory as well (no flat loads): 100000
	// dgemm_nt + dsyrk_ln 6.39 -> 5.81 ms, getrf 7.98 -> 7.90
	const BFA_GLOBAL typename MatOf<T>::type *gs = (const BFA_GLOBAL typename MatOf<T>::type *)s;
	PM<T, PS> v;
	v.p = gs->pA;
	v.cn = gs->cn;
	return v;
}

// The same with the descriptor read through the generic pointer, for the
// one-wavefront left solve: there the global form was slower (16384 LU
// solves, two such launches: 1.31 -> 1.40 ms).
template <typename T, int PS>
__device__ __forceinline__ PM<T, PS> view_flat(const typename MatOf<T>::type *s)
{
	PM<T, PS> v;
	v.p = s->pA;
	v.cn = s->cn;
	return v;
}

// Walks the elements of the R x C block at absolute (r0, c0) of a panel-major
// matrix that this thread copies: the waves take the covering panels in turn
// and the lanes of a wave walk one panel's contiguous run of C*PS elements, so
// a wave reads contiguous memory and no index needs a runtime division.  The
// elements go B at a time: a group's global loads are all issued before any
// of its stores, so a thread keeps up to B loads in flight instead of waiting
// for each (the second pass recomputes the indices, so a group costs B values'
// registers).  Elements with i >= Rv or c >= Cv (and, with LOWER, those above
// the diagonal, i < c) are passed on as zero and never read; with SKIPTILES,
// elements whose 16x16 tile lies above the diagonal are skipped.
// put(i, c, v) stores element (i, c) of the block.
template <typename T, int PS, int NT, bool LOWER, bool SKIPTILES, typename PUT, int W0 = 0, int B = 8>
__device__ __forceinline__ void g2s_walk(const PM<T, PS> g, int r0, int c0, int R, int C, int Rv, int Cv, PUT put)
{
	// waves W0 .. NT/WAVE - 1 take part (all of them by default)
	constexpr int NW = NT / WAVE - W0;
	const int rr = r0 & (PS - 1);
	const int np = (rr + R + PS - 1) / PS;
	const int run = PS * C;
	const BFA_GLOBAL T *base = (const BFA_GLOBAL T *)g.p + (size_t)(r0 - rr) * g.cn + (size_t)c0 * PS;
	const int lane = threadIdx.x & (WAVE - 1);
	int p = (int)(threadIdx.x / WAVE) - W0, e = lane;
	if (p < 0)
		return;
	while (p < np)
	{
		const int p0 = p, e0 = e;
		T v[B];
#pragma unroll
		for (int b = 0; b < B; b++)
		{
			const int c = (unsigned)e / PS;
			const int i = p * PS + (e & (PS - 1)) - rr;
			const bool ok = p < np && e < run && i >= 0 && i < R && (!SKIPTILES || (i >> 4) >= (c >> 4));
			const bool ld = ok && i < Rv && c < Cv && (!LOWER || i >= c);
			v[b] = ld ? base[(size_t)p * PS * g.cn + e] : T(0);
			e += WAVE;
			if (e >= run)
			{
				e = lane;
				p += NW;
			}
		}
		p = p0;
		e = e0;
#pragma unroll
		for (int b = 0; b < B; b++)
		{
			const int c = (unsigned)e / PS;
			const int i = p * PS + (e & (PS - 1)) - rr;
			if (p < np && e < run && i >= 0 && i < R && (!SKIPTILES || (i >> 4) >= (c >> 4)))
				put(i, c, v[b]);
			e += WAVE;
			if (e >= run)
			{
				e = lane;
				p += NW;
			}
		}
	}
}

// Cooperative copy of the R x C block at absolute (r0, c0) of a panel-major
// matrix into LDS, element (i,c) stored at lds[c*ld + i] (TRANS=false) or
// lds[i*ld + c] (TRANS=true); elements outside Rv x Cv (and, with LOWER, above
// the diagonal) are stored as zero.
template <typename T, int PS, int NT, bool TRANS, bool LOWER = false>
__device__ __forceinline__ void g2s(const PM<T, PS> g, int r0, int c0, int R, int C, int Rv, int Cv, T *lds, int ld)
{
	g2s_walk<T, PS, NT, LOWER, false>(g, r0, c0, R, C, Rv, Cv, [=](int i, int c, T v) {
		if (TRANS)
			lds[i * ld + c] = v;
		else
			lds[c * ld + i] = v;
	});
}

// Cooperative copy of LDS (lds[c*ld + i]) back to the Rv x Cv block at
// absolute (r0, c0); LOWER keeps only elements with i - c >= diag.
template <typename T, int PS, int NT, bool LOWER>
__device__ __forceinline__ void s2g(PM<T, PS> g, int r0, int c0, int Rv, int Cv, const T *lds, int ld, int diag)
{
	constexpr int NW = NT / WAVE;
	const int rr = r0 & (PS - 1);
	const int np = (rr + Rv + PS - 1) / PS;
	const int run = PS * Cv;
	BFA_GLOBAL T *base = (BFA_GLOBAL T *)g.p + (size_t)(r0 - rr) * g.cn + (size_t)c0 * PS;
	const int lane = threadIdx.x & (WAVE - 1);
	for (int p = threadIdx.x / WAVE; p < np; p += NW)
	{
		BFA_GLOBAL T *dst = base + (size_t)p * PS * g.cn;
		// with LOWER, columns beyond the panel's last row are never written
		int cend = run;
		if (LOWER)
		{
			const int last = min(p * PS + PS - 1 - rr, Rv - 1);
			cend = min(run, (last - diag + 1) * PS);
		}
		for (int e = lane; e < cend; e += WAVE)
		{
			const int c = (unsigned)e / PS;
			const int q = e & (PS - 1);
			const int i = p * PS + q - rr;
			if (i < 0 || i >= Rv)
				continue;
			if (LOWER && i - c < diag)
				continue;
			dst[e] = lds[c * ld + i];
		}
	}
}

// The same two copies with the LDS side addressed through an accessor
// A(i, c) (element (i, c) of the block).  With LOWER, g2s_acc skips elements
// whose 16x16 tile lies above the diagonal (packed lower-tile storage has no
// such tiles) and zero-fills the upper part of diagonal tiles.
template <typename T, int PS, int NT, bool LOWER, typename ACC, int W0 = 0, int B = 8>
__device__ __forceinline__ void g2s_acc(const PM<T, PS> g, int r0, int c0, int R, int C, int Rv, int Cv, const ACC A)
{
	auto put = [=](int i, int c, T v) { A(i, c) = v; };
	g2s_walk<T, PS, NT, LOWER, LOWER, decltype(put), W0, B>(g, r0, c0, R, C, Rv, Cv, put);
}

template <typename T, int PS, int NT, bool LOWER, typename ACC>
__device__ __forceinline__ void s2g_acc(PM<T, PS> g, int r0, int c0, int Rv, int Cv, const ACC A, int diag)
{
	constexpr int NW = NT / WAVE;
	const int rr = r0 & (PS - 1);
	const int np = (rr + Rv + PS - 1) / PS;
	const int run = PS * Cv;
	BFA_GLOBAL T *base = (BFA_GLOBAL T *)g.p + (size_t)(r0 - rr) * g.cn + (size_t)c0 * PS;
	const int lane = threadIdx.x & (WAVE - 1);
	for (int p = threadIdx.x / WAVE; p < np; p += NW)
	{
		BFA_GLOBAL T *dst = base + (size_t)p * PS * g.cn;
		int cend = run;
		if (LOWER)
		{
			const int last = min(p * PS + PS - 1 - rr, Rv - 1);
			cend = min(run, (last - diag + 1) * PS);
		}
		for (int e = lane; e < cend; e += WAVE)
		{
			const int c = (unsigned)e / PS;
			const int q = e & (PS - 1);
			const int i = p * PS + q - rr;
			if (i < 0 || i >= Rv)
				continue;
			if (LOWER && i - c < diag)
				continue;
			dst[e] = A(i, c);
		}
	}
}

// A DPP move of a double (both 32-bit halves with the same control): lanes
// whose source lane is outside the pattern, or whose row is not in ROWMASK,
// get `old`.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d(double old, double v)
{
	const long long o = __builtin_bit_cast(long long, old), x = __builtin_bit_cast(long long, v);
	const int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, CTRL, ROWMASK, 0xf, false);
	const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), CTRL, ROWMASK, 0xf, false);
	return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// Wave-wide sum and max of a double with DPP moves (quad permutes, row
// shifts by 4 and 8, row broadcasts of lanes 15 and 31): the result is in
// lane 63, in a fixed order.  Register to register, where a shuffle goes
// through the LDS crossbar.  wave_max_d: for values >= 0 or NaN-free inputs
// (the identity used is the lane's own value).
__device__ __forceinline__ double wave_sum_d(double v)
{
	v += dpp_d<0xb1, 0xf>(0.0, v);   // quad_perm 1,0,3,2
	v += dpp_d<0x4e, 0xf>(0.0, v);   // quad_perm 2,3,0,1
	v += dpp_d<0x114, 0xf>(0.0, v);  // row_shr:4
	v += dpp_d<0x118, 0xf>(0.0, v);  // row_shr:8
	v += dpp_d<0x142, 0xa>(0.0, v);  // row_bcast:15 into rows 1, 3
	v += dpp_d<0x143, 0xc>(0.0, v);  // row_bcast:31 into rows 2, 3
	return v;
}
__device__ __forceinline__ double wave_max_d(double v)
{
	v = fmax(v, dpp_d<0xb1, 0xf>(v, v));
	v = fmax(v, dpp_d<0x4e, 0xf>(v, v));
	v = fmax(v, dpp_d<0x114, 0xf>(v, v));
	v = fmax(v, dpp_d<0x118, 0xf>(v, v));
	v = fmax(v, dpp_d<0x142, 0xa>(v, v));
	v = fmax(v, dpp_d<0x143, 0xc>(v, v));
	return v;
}


// DPP move of a 32-bit value (see dpp_d).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i(int old, int v)
{
	return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWMASK, 0xf, false);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_t(float old, float v)
{
	return __builtin_bit_cast(float, dpp_i<CTRL, ROWMASK>(__builtin_bit_cast(int, old), __builtin_bit_cast(int, v)));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_t(double old, double v)
{
	return dpp_d<CTRL, ROWMASK>(old, v);
}

template <typename T, int CTRL, int ROWMASK>
__device__ __forceinline__ void argmax_step(T &best, int &idx)
{
	// a lane without a source in the pattern compares with itself
	const T ob = dpp_t<CTRL, ROWMASK>(best, best);
	const int oi = dpp_i<CTRL, ROWMASK>(idx, idx);
	if (ob > best || (ob == best && oi < idx))
	{
		best = ob;
		idx = oi;
	}
}

// Wave-wide argmax of |v| with the reference's pivot rule: a candidate wins
// only if strictly larger (so NaN never wins), ties go to the smaller row
// index.  Each lane passes its own candidate, which is never NaN (a lane's
// scan only takes strictly larger values), so the rule is a total order and
// the winner does not depend on the reduction tree: DPP moves (quad
// permutes, row shifts, row broadcasts) bring it to lane 63, register to
// register.  Returns the winning row index (uniform across the wave);
// wave_argmax_abs_to_last leaves the winner (value and index) in lane 63.
template <typename T>
__device__ __forceinline__ void wave_argmax_abs_to_last(T &best, int &idx)
{
	argmax_step<T, 0xb1, 0xf>(best, idx);
	argmax_step<T, 0x4e, 0xf>(best, idx);
	argmax_step<T, 0x114, 0xf>(best, idx);
	argmax_step<T, 0x118, 0xf>(best, idx);
	argmax_step<T, 0x142, 0xa>(best, idx);
	argmax_step<T, 0x143, 0xc>(best, idx);
}
template <typename T>
__device__ __forceinline__ int wave_argmax_abs(T best, int idx)
{
	wave_argmax_abs_to_last(best, idx);
	return __builtin_amdgcn_readlane(idx, WAVE - 1);
}

// 1/sqrt(c) for c > 0 from the hardware estimate refined by one Newton step
// y <- y + y*(1/2 - c*y*y/2): within about one ulp of the correctly rounded
// value, with a dependency chain of one transcendental and three FMAs (the
// Cholesky panel's critical path).
__device__ __forceinline__ double rsqrt_fast(double c)
{
	const double y = __builtin_amdgcn_rsq(c);
	const double h = 0.5 * c * y;
	const double r = __builtin_fma(-h, y, 0.5);
	return __builtin_fma(y, r, y);
}
__device__ __forceinline__ float rsqrt_fast(float c)
{
	const float y = __builtin_amdgcn_rsqf(c);
	const float h = 0.5f * c * y;
	const float r = __builtin_fmaf(-h, y, 0.5f);
	return __builtin_fmaf(y, r, y);
}

template <typename T> __device__ __forceinline__ T rsqrt_ref(T c);
// 1/sqrt(c) as the reference forms it (1.0/sqrt(c), correctly rounded sqrt
// followed by a correctly rounded division).
template <> __device__ __forceinline__ double rsqrt_ref<double>(double c) { return 1.0 / __builtin_sqrt(c); }
template <> __device__ __forceinline__ float rsqrt_ref<float>(float c) { return 1.0f / __builtin_sqrtf(c); }

} // namespace bfa
