// Engine of the CPU emulation of the HIP device model (see hip/hip_runtime.h
// in this directory).  TEST INFRASTRUCTURE ONLY.
//
// A workgroup's threads are cooperative fibers on the calling OS thread: each
// runs until it reaches a barrier (the workgroup's, or its wave's: every
// cross-lane operation is a publish / wave barrier / read / wave barrier
// sequence) and then yields.  A wave barrier is released when every live
// thread of the wave waits on it, the workgroup barrier when every live
// thread of the workgroup does; a thread that has returned leaves both, as
// an ended wave leaves the hardware barrier.  Any other state (threads of one
// wave waiting on different barriers, or none can proceed) is a barrier in
// divergent code: the run aborts with a message.
#include "hip/hip_runtime.h"
#include <stdarg.h>
#include <sys/mman.h>
#include <vector>
#include <sanitizer/asan_interface.h>
#include <sanitizer/common_interface_defs.h>

thread_local emu_idx threadIdx, blockIdx, blockDim, gridDim;

#define EMU_LDS_MAX (160 * 1024)
extern "C" {
alignas(256) unsigned char emu_lds[EMU_LDS_MAX + 256];
}

namespace emu {

enum State { READY = 0, WAIT_WAVE = 1, WAIT_BLOCK = 2, DONE = 3 };

// Context switch: saves the callee-saved registers, MXCSR and the x87
// control word on the current stack, stores the stack pointer in *save and
// continues on the stack `next` (System V x86-64).
extern "C" void emu_switch(void **save, void *next);
asm(R"(
	.text
	.globl emu_switch
	.type emu_switch, @function
emu_switch:
	pushq %rbp
	pushq %rbx
	pushq %r12
	pushq %r13
	pushq %r14
	pushq %r15
	subq $8, %rsp
	stmxcsr (%rsp)
	fnstcw 4(%rsp)
	movq %rsp, (%rdi)
	movq %rsi, %rsp
	ldmxcsr (%rsp)
	fldcw 4(%rsp)
	addq $8, %rsp
	popq %r15
	popq %r14
	popq %r13
	popq %r12
	popq %rbx
	popq %rbp
	ret
	.size emu_switch, .-emu_switch
)");

struct Fiber
{
	void *sp = nullptr;
	void *stack = nullptr;
	int tid = 0;
	int state = READY;
};

struct Wave
{
	uint64_t slot[64];
	double fa[64], fb[64];
	float sa[64], sb[64];
};

static const size_t STACK = 1 << 20;
static void *g_sched_sp;
static const void *g_sched_bottom;
static size_t g_sched_size;
static Fiber *g_cur;
static std::vector<Wave> g_waves;
static block_fn g_fn;
static void *g_ctx;
static int t_lane, t_wave;

int lane() { return t_lane; }

static void yield(int st)
{
	Fiber *f = g_cur;
	f->state = st;
	void *fake = nullptr;
	__sanitizer_start_switch_fiber(&fake, g_sched_bottom, g_sched_size);
	emu_switch(&f->sp, g_sched_sp);
	__sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
}

void sync_block() { yield(WAIT_BLOCK); }
void wave_bar() { yield(WAIT_WAVE); }

extern "C" void emu_fiber_main()
{
	__sanitizer_finish_switch_fiber(nullptr, &g_sched_bottom, &g_sched_size);
	g_fn(g_ctx);
	g_cur->state = DONE;
	__sanitizer_start_switch_fiber(nullptr, g_sched_bottom, g_sched_size);
	void *dead;
	emu_switch(&dead, g_sched_sp);
	abort();
}

uint64_t exch(uint64_t v, int src)
{
	Wave &w = g_waves[t_wave];
	w.slot[t_lane] = v;
	wave_bar();
	const uint64_t r = w.slot[src & 63];
	wave_bar();
	return r;
}

uint64_t exch_dpp(uint64_t old, uint64_t v, int ctrl, int row_mask, int bank_mask, bool bound_ctrl)
{
	Wave &w = g_waves[t_wave];
	w.slot[t_lane] = v;
	wave_bar();
	const int l = t_lane, row = l >> 4, r16 = l & 15;
	int src = -1;
	if (ctrl >= 0 && ctrl <= 0xff)
		src = (l & ~3) | ((ctrl >> (2 * (l & 3))) & 3);
	else if (ctrl >= 0x101 && ctrl <= 0x10f)
	{
		const int s = ctrl & 15;
		src = r16 + s < 16 ? l + s : -1;
	}
	else if (ctrl >= 0x111 && ctrl <= 0x11f)
	{
		const int s = ctrl & 15;
		src = r16 >= s ? l - s : -1;
	}
	else if (ctrl >= 0x121 && ctrl <= 0x12f)
	{
		const int s = ctrl & 15;
		src = (l & ~15) | ((r16 - s) & 15);
	}
	else if (ctrl == 0x142)
		src = row >= 1 ? 16 * row - 1 : -1;
	else if (ctrl == 0x143)
		src = row >= 2 ? 31 : -1;
	else
	{
		fprintf(stderr, "emu: DPP control 0x%x not emulated\n", ctrl);
		abort();
	}
	uint64_t r;
	if (!((row_mask >> row) & 1) || !((bank_mask >> ((l >> 2) & 3)) & 1))
		r = old;
	else if (src < 0)
		r = bound_ctrl ? 0 : old;
	else
		r = w.slot[src];
	wave_bar();
	return r;
}

// 16x16x4 products with the operand layouts of bfa_common.hpp
void mfma_f64(double a, double b, double acc[4])
{
	Wave &w = g_waves[t_wave];
	w.fa[t_lane] = a;
	w.fb[t_lane] = b;
	wave_bar();
	const int col = t_lane & 15;
	for (int r = 0; r < 4; r++)
	{
		const int row = (t_lane >> 4) + 4 * r;
		double s = acc[r];
		for (int k = 0; k < 4; k++)
			s = fma(w.fa[k * 16 + row], w.fb[k * 16 + col], s);
		acc[r] = s;
	}
	wave_bar();
}

void mfma_f32(float a, float b, float acc[4])
{
	Wave &w = g_waves[t_wave];
	w.sa[t_lane] = a;
	w.sb[t_lane] = b;
	wave_bar();
	const int col = t_lane & 15;
	for (int r = 0; r < 4; r++)
	{
		const int row = 4 * (t_lane >> 4) + r;
		float s = acc[r];
		for (int k = 0; k < 4; k++)
			s = fmaf(w.sa[k * 16 + row], w.sb[k * 16 + col], s);
		acc[r] = s;
	}
	wave_bar();
}

static void run_block(std::vector<Fiber> &fib, dim3 grid, dim3 block, unsigned bx, unsigned by, unsigned bz)
{
	const int nt = (int)fib.size();
	for (int t = 0; t < nt; t++)
	{
		Fiber &f = fib[t];
		f.tid = t;
		f.state = READY;
		// initial frame: MXCSR / x87 control word, six registers, the entry
		// point as return address, and a zero slot so that the entry runs
		// with the ABI's stack alignment
		uint64_t *top = (uint64_t *)((char *)f.stack + STACK);
		top[-1] = 0;
		top[-2] = (uint64_t)(uintptr_t)&emu_fiber_main;
		for (int r = 3; r <= 8; r++)
			top[-r] = 0;
		uint32_t *cw = (uint32_t *)&top[-9];
		cw[0] = 0x1f80;
		cw[1] = 0x037f;
		f.sp = &top[-9];
	}
	blockIdx = emu_idx{bx, by, bz};
	blockDim = emu_idx{block.x, block.y, block.z};
	gridDim = emu_idx{grid.x, grid.y, grid.z};
	const int nw = (nt + 63) / 64;
	for (;;)
	{
		for (int t = 0; t < nt; t++)
		{
			Fiber &f = fib[t];
			if (f.state != READY)
				continue;
			g_cur = &f;
			threadIdx = emu_idx{(unsigned)t, 0, 0};
			t_lane = t & 63;
			t_wave = t >> 6;
			void *fake = nullptr;
			__sanitizer_start_switch_fiber(&fake, f.stack, STACK);
			emu_switch(&g_sched_sp, f.sp);
			__sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
		}
		// release the barriers every live thread has reached
		bool any = false, live = false, all_block = true;
		for (int w = 0; w < nw; w++)
		{
			int nlive = 0, nwave = 0;
			for (int t = 64 * w; t < std::min(nt, 64 * w + 64); t++)
			{
				if (fib[t].state == DONE)
					continue;
				nlive++;
				nwave += fib[t].state == WAIT_WAVE;
				all_block = all_block && fib[t].state == WAIT_BLOCK;
			}
			live = live || nlive > 0;
			if (nlive > 0 && nwave == nlive)
			{
				for (int t = 64 * w; t < std::min(nt, 64 * w + 64); t++)
					fib[t].state = READY;
				any = true;
			}
		}
		if (!live)
			return;
		if (any)
			continue;
		if (all_block)
		{
			for (auto &f : fib)
				if (f.state == WAIT_BLOCK)
					f.state = READY;
			continue;
		}
		fprintf(stderr, "emu: block (%u,%u): threads wait on different barriers (a barrier or "
			"cross-lane operation in divergent code)\n", bx, by);
		abort();
	}
}

void run_grid(dim3 grid, dim3 block, size_t shm, block_fn fn, void *ctx)
{
	const int nt = (int)(block.x * block.y * block.z);
	if (shm > EMU_LDS_MAX || nt <= 0 || nt > 1024 || block.y != 1 || block.z != 1)
	{
		fprintf(stderr, "emu: launch shape not supported (shm %zu, block %u x %u x %u)\n", shm, block.x, block.y, block.z);
		abort();
	}
	g_fn = fn;
	g_ctx = ctx;
	g_waves.assign((nt + 63) / 64, Wave());
	std::vector<Fiber> fib(nt);
	for (auto &f : fib)
		f.stack = mmap(nullptr, STACK, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
	ASAN_UNPOISON_MEMORY_REGION(emu_lds, sizeof(emu_lds));
	ASAN_POISON_MEMORY_REGION(emu_lds + shm, sizeof(emu_lds) - shm);
	for (unsigned bz = 0; bz < grid.z; bz++)
		for (unsigned by = 0; by < grid.y; by++)
			for (unsigned bx = 0; bx < grid.x; bx++)
			{
				memset(emu_lds, 0xff, shm);  // uninitialised LDS reads as NaN
				run_block(fib, grid, block, bx, by, bz);
			}
	ASAN_UNPOISON_MEMORY_REGION(emu_lds, sizeof(emu_lds));
	for (auto &f : fib)
		munmap(f.stack, STACK);
}

} // namespace emu

// host-side helpers of the library (bfa_internal.hpp), emulator versions
#include "../../blasfeo_amd/csrc/bfa_internal.hpp"
namespace bfa {
static thread_local char g_err[512];
int set_error(int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof g_err, fmt, ap);
	va_end(ap);
	return code;
}
int check_launch(const char *) { return BLASFEO_AMD_OK; }
int num_cus() { return 256; }
int lds_check(const void *, size_t, const char *) { return BLASFEO_AMD_OK; }
void *workspace(size_t bytes, hipStream_t)
{
	static std::vector<unsigned char> ws;
	if (ws.size() < bytes)
		ws.resize(bytes);
	return ws.data();
}
void *plan_workspace(size_t bytes, hipStream_t)
{
	static std::vector<unsigned char> ws;
	if (ws.size() < bytes)
		ws.resize(bytes);
	return ws.data();
}
}
extern "C" const char *emu_last_error() { return bfa::g_err; }
