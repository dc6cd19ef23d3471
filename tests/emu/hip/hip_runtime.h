// CPU emulation of the parts of the HIP device model that the library's
// kernels use, so that a kernel source file can be compiled for the host and
// run under AddressSanitizer.  TEST INFRASTRUCTURE ONLY: nothing in the
// library includes this header; tests/emu/Makefile puts this directory first
// on the include path of the emulator build, where it stands in for the HIP
// runtime header.
//
// Model: a launch runs its workgroups one after another; a workgroup's
// threads are host threads.  __syncthreads() is a barrier over the
// workgroup; the cross-lane builtins (readlane, DPP moves, shuffles, the
// 16x16x4 matrix-core products) exchange values through a per-wave slot
// array between two barriers over the wave's 64 threads, so every lane of a
// wave must reach them, as on the hardware (a barrier that is not reached by
// all its threads within the time limit aborts the run with a message).
// LDS is one host array (dynamic part, ASan-poisoned beyond the launch's
// size) plus function-local statics for the static part.
#pragma once
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <stdio.h>
#include <stdlib.h>
#include <algorithm>
#include <cmath>
#include <tuple>

#define __global__
#define __device__
#define __constant__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __noinline__ __attribute__((noinline))

struct dim3
{
	unsigned x, y, z;
	dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_idx { unsigned x, y, z; };
extern thread_local emu_idx threadIdx, blockIdx;
extern thread_local emu_idx blockDim, gridDim;

struct int2 { int x, y; };

typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
#define hipSuccess 0
#define hipErrorInvalidValue 1
#define hipEventDisableTiming 2
#define hipStreamNonBlocking 1

inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline long min(long a, long b) { return a < b ? a : b; }
inline long max(long a, long b) { return a > b ? a : b; }
inline unsigned long min(unsigned long a, unsigned long b) { return a < b ? a : b; }
inline unsigned long max(unsigned long a, unsigned long b) { return a > b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }
inline float min(float a, float b) { return fminf(a, b); }
inline float max(float a, float b) { return fmaxf(a, b); }
inline double min(double a, double b) { return fmin(a, b); }
inline double max(double a, double b) { return fmax(a, b); }

namespace emu {
void sync_block();
void wave_bar();
uint64_t exch(uint64_t v, int src_lane);           // value of src_lane (after all lanes published)
uint64_t exch_dpp(uint64_t old, uint64_t v, int ctrl, int row_mask, int bank_mask, bool bound_ctrl);
void mfma_f64(double a, double b, double acc[4]);
void mfma_f32(float a, float b, float acc[4]);
int lane();
typedef void (*block_fn)(void *);
void run_grid(dim3 grid, dim3 block, size_t shm, block_fn fn, void *ctx);
}

extern "C" unsigned char emu_lds[];

// cross-lane builtins
inline int emu_readlane_i(int v, int l) { return (int)emu::exch((uint64_t)(uint32_t)v, l); }
#define __builtin_amdgcn_readlane(v, l) emu_readlane_i((v), (l))
#define __builtin_amdgcn_readfirstlane(v) emu_readlane_i((v), 0)
inline int emu_dpp_i(int old, int v, int ctrl, int rm, int bm, bool bc)
{
	return (int)emu::exch_dpp((uint64_t)(uint32_t)old, (uint64_t)(uint32_t)v, ctrl, rm, bm, bc);
}
#define __builtin_amdgcn_update_dpp(old, v, ctrl, rm, bm, bc) emu_dpp_i((old), (v), (ctrl), (rm), (bm), (bc))
#define __builtin_amdgcn_wave_barrier() emu::wave_bar()
#define __builtin_amdgcn_fence(...) ((void)0)
#define __builtin_amdgcn_rsq(c) (1.0 / sqrt((double)(c)))
#define __builtin_amdgcn_rsqf(c) (1.0f / sqrtf((float)(c)))
#define __builtin_amdgcn_rcp(c) (1.0 / (double)(c))
#define __builtin_amdgcn_rcpf(c) (1.0f / (float)(c))

typedef double emu_d4 __attribute__((ext_vector_type(4)));
typedef float emu_f4 __attribute__((ext_vector_type(4)));
inline emu_d4 emu_mfma_d(double a, double b, emu_d4 acc)
{
	double t[4] = {acc[0], acc[1], acc[2], acc[3]};
	emu::mfma_f64(a, b, t);
	return emu_d4{t[0], t[1], t[2], t[3]};
}
inline emu_f4 emu_mfma_f(float a, float b, emu_f4 acc)
{
	float t[4] = {acc[0], acc[1], acc[2], acc[3]};
	emu::mfma_f32(a, b, t);
	return emu_f4{t[0], t[1], t[2], t[3]};
}
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, x, y, z) emu_mfma_d((a), (b), (acc))
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, x, y, z) emu_mfma_f((a), (b), (acc))

template <typename T>
inline T __shfl(T v, int src, int width = 64)
{
	uint64_t u = 0;
	memcpy(&u, &v, sizeof(T));
	const int l = emu::lane();
	const int s = (l & ~(width - 1)) + (src & (width - 1));
	u = emu::exch(u, s);
	T r;
	memcpy(&r, &u, sizeof(T));
	return r;
}
template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64)
{
	uint64_t u = 0;
	memcpy(&u, &v, sizeof(T));
	const int l = emu::lane();
	int s = l ^ mask;
	if ((s & ~(width - 1)) != (l & ~(width - 1)))
		s = l;
	u = emu::exch(u, s);
	T r;
	memcpy(&r, &u, sizeof(T));
	return r;
}

template <typename T>
inline T atomicAdd(T *p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }

inline void __syncthreads() { emu::sync_block(); }

// kernel launch: every workgroup runs the kernel body with the launch's arguments
template <typename... KA, typename... A>
inline void emu_launch(void (*k)(KA...), dim3 grid, dim3 block, size_t shm, A... args)
{
	struct Ctx { void (*k)(KA...); std::tuple<KA...> a; };
	Ctx c{k, std::tuple<KA...>(static_cast<KA>(args)...)};
	emu::run_grid(grid, block, shm, [](void *p) {
		Ctx *c = (Ctx *)p;
		std::apply(c->k, c->a);
	}, &c);
}
#define hipLaunchKernelGGL(k, grid, block, shm, stream, ...) emu_launch(k, dim3(grid), dim3(block), (size_t)(shm), __VA_ARGS__)

#define BFA_GLOBAL   // host memory: no address-space qualifier
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = nullptr; return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = nullptr; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipPeekAtLastError() { return hipSuccess; }
inline const char *hipGetErrorString(hipError_t) { return "emulated"; }
