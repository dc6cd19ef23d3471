// bfa_host.cpp -- host side of the library: error reporting, panel-major
// layout helpers (memsize / create / pack / unpack, auxiliary/d_aux_lib4.c,
// d_aux_lib8.c, d_aux_common.c), GPU allocation of descriptors
// (auxiliary/x_aux_ext_dep.c:40-59) and the single-instance struct API, which
// stages the caller's descriptors to the GPU and runs a batch of one.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "bfa_internal.hpp"

namespace bfa {

static thread_local char g_err[512] = "";
static int g_ps = 4;

int set_error(int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof(g_err), fmt, ap);
	va_end(ap);
	return code;
}

int check_launch(const char *what)
{
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		return set_error(BLASFEO_AMD_ELAUNCH, "%s: %s", what, hipGetErrorString(e));
	return BLASFEO_AMD_OK;
}

// Whether a launch of kernel fn with dyn bytes of dynamic LDS fits the
// device: its static LDS plus dyn at most the per-workgroup limit.  A launch
// beyond it aborts the queue at dispatch instead of failing the launch call.
int lds_check(const void *fn, size_t dyn, const char *what)
{
	// static LDS per (device, kernel), in a fixed table (no allocation: the
	// allocation-free entries use this check too)
	struct Entry { int dev; const void *fn; size_t stat; };
	constexpr int NCACHE = 256;
	static std::mutex mu;
	static Entry cache[NCACHE];
	static int ncache = 0;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess)
		return set_error(BLASFEO_AMD_ELAUNCH, "%s: no device", what);
	size_t stat = 0, limit = 0;
	{
		std::lock_guard<std::mutex> lock(mu);
		int i = 0;
		while (i < ncache && !(cache[i].dev == dev && cache[i].fn == fn))
			i++;
		if (i == ncache)
		{
			hipFuncAttributes at;
			if (hipFuncGetAttributes(&at, fn) != hipSuccess)
				return check_launch(what);
			if (ncache < NCACHE)
				cache[ncache++] = Entry{dev, fn, at.sharedSizeBytes};
			stat = at.sharedSizeBytes;
		}
		else
			stat = cache[i].stat;
	}
	// the per-block attribute can report less than a CU's LDS, which one
	// workgroup may use whole on gfx950; take the larger of the two
	int lb = 0, lm = 0;
	if (hipDeviceGetAttribute(&lb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
		hipDeviceGetAttribute(&lm, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess)
		return check_launch(what);
	limit = (size_t)(lb > lm ? lb : lm);
	if (stat + dyn > limit)
		return set_error(BLASFEO_AMD_ELAUNCH, "%s: %zu bytes of LDS (%zu static + %zu dynamic) exceed the %zu-byte limit",
			what, stat + dyn, stat, dyn, limit);
	return BLASFEO_AMD_OK;
}

int num_cus()
{
	int dev = 0, n = 0;
	if (hipGetDevice(&dev) != hipSuccess)
		return 0;
	if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
		return 0;
	return n;
}

struct WsBlock
{
	void *ptr = nullptr;
	size_t bytes = 0;
};
static std::mutex g_ws_mutex;
static std::map<std::pair<int, hipStream_t>, WsBlock> g_ws, g_plan_ws;
// Blocks that a larger request replaced.  They are never freed: a HIP graph
// captured while one was current keeps its address in its kernel arguments,
// and replaying that graph must still find the memory there.
static std::vector<void *> g_ws_retired;

static void *ws_block(std::map<std::pair<int, hipStream_t>, WsBlock> &blocks, size_t bytes, hipStream_t stream)
{
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess)
		return nullptr;
	std::lock_guard<std::mutex> lock(g_ws_mutex);
	WsBlock &b = blocks[std::make_pair(dev, stream)];
	if (b.bytes >= bytes)
		return b.ptr;
	size_t want = bytes < ((size_t)1 << 20) ? ((size_t)1 << 20) : bytes;
	void *p = nullptr;
	if (hipMalloc(&p, want) != hipSuccess)
		return nullptr;
	if (b.ptr)
		g_ws_retired.push_back(b.ptr);
	b.ptr = p;
	b.bytes = want;
	return b.ptr;
}

void *workspace(size_t bytes, hipStream_t stream) { return ws_block(g_ws, bytes, stream); }
void *plan_workspace(size_t bytes, hipStream_t stream) { return ws_block(g_plan_ws, bytes, stream); }

// panel leading dimension (column rounding) of the reference builds:
// double lib4 -> 4, double lib8 -> 8 (include/blasfeo_block_size.h:54,100),
// float -> 4 (S_PLD)
static inline int pld_d(int ps) { return ps == 8 ? 8 : 4; }
static const size_t CACHE_LINE = 64;

static bool have_device()
{
	static int cached = -1;
	if (cached < 0)
	{
		int n = 0;
		cached = (hipGetDeviceCount(&n) == hipSuccess && n > 0) ? 1 : 0;
		(void)hipGetLastError();
	}
	return cached == 1;
}

static bool is_device_ptr(const void *p)
{
	if (!p || !have_device())
		return false;
	hipPointerAttribute_t at;
	if (hipPointerGetAttributes(&at, p) != hipSuccess)
	{
		(void)hipGetLastError();
		return false;
	}
	return at.type == hipMemoryTypeDevice;
}

// -------------------------------------------------------------------------
// host staging for the single-instance API: per thread, one small device
// block for the descriptors and the argument record
// -------------------------------------------------------------------------
struct Staging
{
	int dev = -1;
	unsigned char *d = nullptr;
	~Staging()
	{
		// device memory is released with the context at process exit
	}
};
static thread_local Staging g_stage;
static const size_t STAGE_BYTES = 8192;  // descriptors, argument record, pivots

static unsigned char *stage_buffer()
{
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess)
		return nullptr;
	if (g_stage.d && g_stage.dev == dev)
		return g_stage.d;
	void *p = nullptr;
	if (hipMalloc(&p, STAGE_BYTES) != hipSuccess)
		return nullptr;
	g_stage.d = (unsigned char *)p;
	g_stage.dev = dev;
	return g_stage.d;
}

} // namespace bfa

using namespace bfa;

extern "C" {

const char *blasfeo_amd_version(void) { return "blasfeo_amd 0.1.0 (gfx950)"; }
const char *blasfeo_amd_last_error(void) { return g_err; }

int blasfeo_amd_set_ps(int ps)
{
	if (ps != 4 && ps != 8)
		return set_error(BLASFEO_AMD_EINVAL, "blasfeo_amd_set_ps: panel size %d not in {4, 8}", ps);
	g_ps = ps;
	return BLASFEO_AMD_OK;
}
int blasfeo_amd_get_ps(void) { return g_ps; }

// ---- layout: auxiliary/d_aux_lib4.c:1519-1570, d_aux_lib8.c:64-115 ----
size_t blasfeo_amd_memsize_dmat(int ps, int m, int n)
{
	const int nc = pld_d(ps);
	const int al = ps * nc;
	const int pm = (m + ps - 1) / ps * ps;
	const int cn = (n + nc - 1) / nc * nc;
	const int tmp = m < n ? (m + al - 1) / al * al : (n + al - 1) / al * al;
	size_t memsize = ((size_t)pm * cn + tmp) * sizeof(double);
	return (memsize + CACHE_LINE - 1) / CACHE_LINE * CACHE_LINE;
}

void blasfeo_amd_create_dmat(int ps, int m, int n, struct blasfeo_dmat *sA, void *memory)
{
	const int nc = pld_d(ps);
	const int al = ps * nc;
	sA->mem = (double *)memory;
	sA->m = m;
	sA->n = n;
	const int pm = (m + ps - 1) / ps * ps;
	const int cn = (n + nc - 1) / nc * nc;
	sA->pm = pm;
	sA->cn = cn;
	double *ptr = (double *)memory;
	sA->pA = ptr;
	ptr += (size_t)pm * cn;
	sA->dA = ptr;
	const int tmp = m < n ? (m + al - 1) / al * al : (n + al - 1) / al * al;
	size_t memsize = ((size_t)pm * cn + tmp) * sizeof(double);
	sA->memsize = (int)((memsize + CACHE_LINE - 1) / CACHE_LINE * CACHE_LINE);
	sA->use_dA = 0;
}

// float panels: ps (4, 8, 16) with column rounding 4 (S_PLD)
size_t blasfeo_amd_memsize_smat(int ps, int m, int n)
{
	const int nc = 4;
	const int al = ps * nc;
	const int pm = (m + ps - 1) / ps * ps;
	const int cn = (n + nc - 1) / nc * nc;
	const int tmp = m < n ? (m + al - 1) / al * al : (n + al - 1) / al * al;
	size_t memsize = ((size_t)pm * cn + tmp) * sizeof(float);
	return (memsize + CACHE_LINE - 1) / CACHE_LINE * CACHE_LINE;
}

void blasfeo_amd_create_smat(int ps, int m, int n, struct blasfeo_smat *sA, void *memory)
{
	const int nc = 4;
	const int al = ps * nc;
	sA->mem = (float *)memory;
	sA->m = m;
	sA->n = n;
	const int pm = (m + ps - 1) / ps * ps;
	const int cn = (n + nc - 1) / nc * nc;
	sA->pm = pm;
	sA->cn = cn;
	float *ptr = (float *)memory;
	sA->pA = ptr;
	ptr += (size_t)pm * cn;
	sA->dA = ptr;
	const int tmp = m < n ? (m + al - 1) / al * al : (n + al - 1) / al * al;
	size_t memsize = ((size_t)pm * cn + tmp) * sizeof(float);
	sA->memsize = (int)((memsize + CACHE_LINE - 1) / CACHE_LINE * CACHE_LINE);
	sA->use_dA = 0;
}

size_t blasfeo_memsize_dmat(int m, int n) { return blasfeo_amd_memsize_dmat(g_ps, m, n); }

size_t blasfeo_memsize_diag_dmat(int m, int n)
{
	const int al = g_ps * pld_d(g_ps);
	const int tmp = m < n ? (m + al - 1) / al * al : (n + al - 1) / al * al;
	return (size_t)tmp * sizeof(double);
}

size_t blasfeo_memsize_dvec(int m)
{
	const int pm = (m + g_ps - 1) / g_ps * g_ps;
	return (size_t)pm * sizeof(double);
}

void blasfeo_create_dmat(int m, int n, struct blasfeo_dmat *sA, void *memory) { blasfeo_amd_create_dmat(g_ps, m, n, sA, memory); }

void blasfeo_create_dvec(int m, struct blasfeo_dvec *sa, void *memory)
{
	sa->mem = (double *)memory;
	sa->m = m;
	const int pm = (m + g_ps - 1) / g_ps * g_ps;
	sa->pm = pm;
	sa->pa = (double *)memory;
	sa->memsize = pm * (int)sizeof(double);
}

// auxiliary/d_aux_common.c:49-83 (column rounding hard-coded to 4)
size_t blasfeo_pm_memsize_dmat(int ps, int m, int n)
{
	const int nc = 4;
	const int al = ps * nc;
	const int pm = (m + ps - 1) / ps * ps;
	const int cn = (n + nc - 1) / nc * nc;
	const int tmp = m < n ? (m + al - 1) / al * al : (n + al - 1) / al * al;
	return ((size_t)pm * cn + tmp) * sizeof(double);
}

void blasfeo_pm_create_dmat(int ps, int m, int n, struct blasfeo_pm_dmat *sA, void *memory)
{
	const int nc = 4;
	const int al = ps * nc;
	sA->mem = (double *)memory;
	sA->ps = ps;
	sA->m = m;
	sA->n = n;
	const int pm = (m + ps - 1) / ps * ps;
	const int cn = (n + nc - 1) / nc * nc;
	sA->pm = pm;
	sA->cn = cn;
	double *ptr = (double *)memory;
	sA->pA = ptr;
	ptr += (size_t)pm * cn;
	sA->dA = ptr;
	const int tmp = m < n ? (m + al - 1) / al * al : (n + al - 1) / al * al;
	sA->memsize = (int)(((size_t)pm * cn + tmp) * sizeof(double));
	sA->use_dA = 0;
}

// ---- pack / unpack on host or device memory ----
// The panel rows touched by rows [ai, ai+m) are one contiguous range of pA;
// device matrices are converted through a host copy of that range.
static void pm_transfer(int m, int n, double *A, int lda, struct blasfeo_dmat *sA, int ai, int aj, int dir, bool tran)
{
	// pack_tran clears the cache even for an empty block, pack and unpack
	// return first (auxiliary/d_aux_lib4.c:2090, :1609, :2238)
	if (dir == 0 && tran)
		sA->use_dA = 0;
	if (m <= 0 || n <= 0)
		return;
	const int ps = g_ps;
	const int sda = sA->cn;
	const int p0 = ai / ps;
	const int p1 = (ai + (tran ? n : m) - 1) / ps;
	const size_t off = (size_t)p0 * ps * sda;
	const size_t cnt = (size_t)(p1 - p0 + 1) * ps * sda;
	const bool dev = is_device_ptr(sA->pA);
	std::vector<double> tmp;
	double *base;
	if (dev)
	{
		tmp.resize(cnt);
		if (hipMemcpy(tmp.data(), sA->pA + off, cnt * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
		{
			set_error(BLASFEO_AMD_ELAUNCH, "pack/unpack: device copy failed");
			return;
		}
		base = tmp.data() - off;
	}
	else
		base = sA->pA;
	auto el = [&](int i, int j) -> double & {
		return base[(size_t)(i - i % ps) * sda + (size_t)j * ps + i % ps];
	};
	if (dir == 0)
	{
		// pack: column-major A (m x n) -> sA, transposed when tran
		for (int j = 0; j < n; j++)
			for (int i = 0; i < m; i++)
			{
				if (tran)
					el(ai + j, aj + i) = A[i + (size_t)j * lda];
				else
					el(ai + i, aj + j) = A[i + (size_t)j * lda];
			}
		sA->use_dA = 0;
		if (dev && hipMemcpy(sA->pA + off, tmp.data(), cnt * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
			set_error(BLASFEO_AMD_ELAUNCH, "pack: device copy failed");
	}
	else
	{
		for (int j = 0; j < n; j++)
			for (int i = 0; i < m; i++)
				A[i + (size_t)j * lda] = el(ai + i, aj + j);
	}
}

void blasfeo_pack_dmat(int m, int n, double *A, int lda, struct blasfeo_dmat *sA, int ai, int aj)
{ pm_transfer(m, n, A, lda, sA, ai, aj, 0, false); }
void blasfeo_pack_tran_dmat(int m, int n, double *A, int lda, struct blasfeo_dmat *sA, int ai, int aj)
{ pm_transfer(m, n, A, lda, sA, ai, aj, 0, true); }
void blasfeo_unpack_dmat(int m, int n, struct blasfeo_dmat *sA, int ai, int aj, double *A, int lda)
{ pm_transfer(m, n, A, lda, sA, ai, aj, 1, false); }

void blasfeo_pack_dvec(int m, double *x, int incx, struct blasfeo_dvec *sy, int yi)
{
	if (m <= 0)
		return;
	std::vector<double> tmp(m);
	for (int i = 0; i < m; i++)
		tmp[i] = x[(size_t)i * incx];
	if (is_device_ptr(sy->pa))
	{
		if (hipMemcpy(sy->pa + yi, tmp.data(), m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
			set_error(BLASFEO_AMD_ELAUNCH, "blasfeo_pack_dvec: device copy failed");
	}
	else
		memcpy(sy->pa + yi, tmp.data(), m * sizeof(double));
}

void blasfeo_unpack_dvec(int m, struct blasfeo_dvec *sx, int xi, double *y, int incy)
{
	if (m <= 0)
		return;
	std::vector<double> tmp(m);
	if (is_device_ptr(sx->pa))
	{
		if (hipMemcpy(tmp.data(), sx->pa + xi, m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
		{
			set_error(BLASFEO_AMD_ELAUNCH, "blasfeo_unpack_dvec: device copy failed");
			return;
		}
	}
	else
		memcpy(tmp.data(), sx->pa + xi, m * sizeof(double));
	for (int i = 0; i < m; i++)
		y[(size_t)i * incy] = tmp[i];
}

// ---- GPU allocation (auxiliary/x_aux_ext_dep.c:40-59: zeroed memory) ----
void blasfeo_allocate_dmat(int m, int n, struct blasfeo_dmat *sA)
{
	size_t size = blasfeo_memsize_dmat(m, n);
	void *mem = nullptr;
	if (hipMalloc(&mem, size) != hipSuccess || hipMemset(mem, 0, size) != hipSuccess)
	{
		set_error(BLASFEO_AMD_ENOMEM, "blasfeo_allocate_dmat: %zu bytes", size);
		memset(sA, 0, sizeof(*sA));
		return;
	}
	blasfeo_create_dmat(m, n, sA, mem);
}

void blasfeo_free_dmat(struct blasfeo_dmat *sA)
{
	if (sA && sA->mem)
		(void)hipFree(sA->mem);
}

void blasfeo_allocate_dvec(int m, struct blasfeo_dvec *sa)
{
	size_t size = blasfeo_memsize_dvec(m);
	void *mem = nullptr;
	if (hipMalloc(&mem, size) != hipSuccess || hipMemset(mem, 0, size) != hipSuccess)
	{
		set_error(BLASFEO_AMD_ENOMEM, "blasfeo_allocate_dvec: %zu bytes", size);
		memset(sa, 0, sizeof(*sa));
		return;
	}
	blasfeo_create_dvec(m, sa, mem);
}

void blasfeo_free_dvec(struct blasfeo_dvec *sa)
{
	if (sa && sa->mem)
		(void)hipFree(sa->mem);
}

} // extern "C"

// ---------------------------------------------------------------------------
// single-instance struct API: stage the (distinct) descriptors and the
// argument record, run a batch of one on the null stream, wait, and copy the
// descriptors' use_dA back (the only field the routines change).
// ---------------------------------------------------------------------------
namespace {

struct Single
{
	struct blasfeo_dmat *host[4];
	int slot[4];
	int nd = 0;
	unsigned char *d = nullptr;

	// returns the device address of the staged copy of descriptor s
	struct blasfeo_dmat *add(struct blasfeo_dmat *s)
	{
		for (int i = 0; i < nd; i++)
			if (host[i] == s)
				return (struct blasfeo_dmat *)(d + i * sizeof(struct blasfeo_dmat));
		host[nd] = s;
		nd++;
		return (struct blasfeo_dmat *)(d + (nd - 1) * sizeof(struct blasfeo_dmat));
	}
};

static const size_t ARG_OFF = 512;
static const size_t PIV_OFF = 1024;

template <typename Arg, typename Fn>
static void run_single(Single &S, const Arg &arg, Fn launch, const char *name)
{
	unsigned char buf[bfa::STAGE_BYTES];
	for (int i = 0; i < S.nd; i++)
		memcpy(buf + i * sizeof(struct blasfeo_dmat), S.host[i], sizeof(struct blasfeo_dmat));
	memcpy(buf + ARG_OFF, &arg, sizeof(Arg));
	if (hipMemcpy(S.d, buf, ARG_OFF + sizeof(Arg), hipMemcpyHostToDevice) != hipSuccess)
	{
		bfa::set_error(BLASFEO_AMD_ELAUNCH, "%s: staging copy failed", name);
		return;
	}
	int rc = launch((const Arg *)(S.d + ARG_OFF));
	if (rc != BLASFEO_AMD_OK)
	{
		fprintf(stderr, "%s: %s\n", name, blasfeo_amd_last_error());
		return;
	}
	if (hipStreamSynchronize(0) != hipSuccess)
	{
		bfa::set_error(BLASFEO_AMD_ELAUNCH, "%s: execution failed", name);
		fprintf(stderr, "%s: %s\n", name, blasfeo_amd_last_error());
		return;
	}
	if (hipMemcpy(buf, S.d, S.nd * sizeof(struct blasfeo_dmat), hipMemcpyDeviceToHost) != hipSuccess)
		return;
	for (int i = 0; i < S.nd; i++)
		S.host[i]->use_dA = ((struct blasfeo_dmat *)(buf + i * sizeof(struct blasfeo_dmat)))->use_dA;
}

static bool single_init(Single &S, const char *name)
{
	S.d = bfa::stage_buffer();
	if (!S.d)
	{
		fprintf(stderr, "%s: no GPU staging buffer\n", name);
		return false;
	}
	return true;
}

} // namespace

extern "C" {

static void gemm_single(int op, int m, int n, int k, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, double beta, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj, const char *name)
{
	Single S;
	if (!single_init(S, name))
		return;
	blasfeo_amd_dgemm_arg a;
	a.m = m; a.n = n; a.k = k;
	a.ai = ai; a.aj = aj; a.bi = bi; a.bj = bj; a.ci = ci; a.cj = cj; a.di = di; a.dj = dj;
	a.alpha = alpha; a.beta = beta;
	a.sA = S.add(sA); a.sB = S.add(sB); a.sC = S.add(sC); a.sD = S.add(sD);
	blasfeo_amd_hint h = {m, n, k};
	const int ps = g_ps;
	run_single(S, a, [&](const blasfeo_amd_dgemm_arg *d) {
		switch (op)
		{
		case 0: return blasfeo_amd_dgemm_nt_batched(ps, 1, d, &h, 0);
		case 1: return blasfeo_amd_dgemm_nn_batched(ps, 1, d, &h, 0);
		case 2: return blasfeo_amd_dsyrk_ln_batched(ps, 1, d, &h, 0);
		default: return blasfeo_amd_dsyrk_ln_mn_batched(ps, 1, d, &h, 0);
		}
	}, name);
}

void blasfeo_dgemm_nt(int m, int n, int k, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, double beta, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj)
{ gemm_single(0, m, n, k, alpha, sA, ai, aj, sB, bi, bj, beta, sC, ci, cj, sD, di, dj, "blasfeo_dgemm_nt"); }
void blasfeo_dgemm_nn(int m, int n, int k, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, double beta, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj)
{ gemm_single(1, m, n, k, alpha, sA, ai, aj, sB, bi, bj, beta, sC, ci, cj, sD, di, dj, "blasfeo_dgemm_nn"); }
void blasfeo_dsyrk_ln(int m, int k, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, double beta, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj)
{ gemm_single(2, m, m, k, alpha, sA, ai, aj, sB, bi, bj, beta, sC, ci, cj, sD, di, dj, "blasfeo_dsyrk_ln"); }
void blasfeo_dsyrk_ln_mn(int m, int n, int k, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, double beta, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj)
{ gemm_single(3, m, n, k, alpha, sA, ai, aj, sB, bi, bj, beta, sC, ci, cj, sD, di, dj, "blasfeo_dsyrk_ln_mn"); }

static void trxm_single(int op, int m, int n, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, struct blasfeo_dmat *sD, int di, int dj, const char *name)
{
	Single S;
	if (!single_init(S, name))
		return;
	blasfeo_amd_dtrsm_arg a;
	a.m = m; a.n = n;
	a.ai = ai; a.aj = aj; a.bi = bi; a.bj = bj; a.di = di; a.dj = dj;
	a.alpha = alpha;
	a.sA = S.add(sA); a.sB = S.add(sB); a.sD = S.add(sD);
	blasfeo_amd_hint h = {m, n, n};
	const int ps = g_ps;
	run_single(S, a, [&](const blasfeo_amd_dtrsm_arg *d) {
		return op == 0 ? blasfeo_amd_dtrsm_rltn_batched(ps, 1, d, &h, 0) : blasfeo_amd_dtrmm_rlnn_batched(ps, 1, d, &h, 0);
	}, name);
}

void blasfeo_dtrsm_rltn(int m, int n, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, struct blasfeo_dmat *sD, int di, int dj)
{ trxm_single(0, m, n, alpha, sA, ai, aj, sB, bi, bj, sD, di, dj, "blasfeo_dtrsm_rltn"); }
void blasfeo_dtrmm_rlnn(int m, int n, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, struct blasfeo_dmat *sD, int di, int dj)
{ trxm_single(1, m, n, alpha, sA, ai, aj, sB, bi, bj, sD, di, dj, "blasfeo_dtrmm_rlnn"); }

static void fact_single(int op, int m, int n, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj, int *ipiv, const char *name)
{
	Single S;
	if (!single_init(S, name))
		return;
	blasfeo_amd_dfact_arg a;
	a.m = m; a.n = n; a.ci = ci; a.cj = cj; a.di = di; a.dj = dj;
	a.sC = S.add(sC); a.sD = S.add(sD);
	a.ipiv = nullptr;
	int *dpiv = nullptr;
	const int p = m < n ? m : n;
	if (op == 2 && p > 0)
	{
		if ((size_t)p * sizeof(int) > bfa::STAGE_BYTES - PIV_OFF)
		{
			fprintf(stderr, "%s: min(m,n)=%d too large for the pivot staging buffer\n", name, p);
			return;
		}
		dpiv = (int *)(S.d + PIV_OFF);
		a.ipiv = dpiv;
	}
	blasfeo_amd_hint h = {m, n, 0};
	const int ps = g_ps;
	run_single(S, a, [&](const blasfeo_amd_dfact_arg *d) {
		if (op == 0)
			return blasfeo_amd_dpotrf_l_batched(ps, 1, d, &h, 0);
		if (op == 1)
			return blasfeo_amd_dpotrf_l_mn_batched(ps, 1, d, &h, 0);
		return blasfeo_amd_dgetrf_rp_batched(ps, 1, d, &h, 0);
	}, name);
	if (dpiv)
	{
		if (hipMemcpy(ipiv, dpiv, sizeof(int) * p, hipMemcpyDeviceToHost) != hipSuccess)
			fprintf(stderr, "%s: pivot copy failed\n", name);
	}
}

void blasfeo_dpotrf_l(int m, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj)
{ fact_single(0, m, m, sC, ci, cj, sD, di, dj, nullptr, "blasfeo_dpotrf_l"); }
void blasfeo_dpotrf_l_mn(int m, int n, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj)
{ fact_single(1, m, n, sC, ci, cj, sD, di, dj, nullptr, "blasfeo_dpotrf_l_mn"); }
void blasfeo_dgetrf_rp(int m, int n, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj, int *ipiv)
{ fact_single(2, m, n, sC, ci, cj, sD, di, dj, ipiv, "blasfeo_dgetrf_rp"); }

void blasfeo_dsyrk_dpotrf_ln(int m, int k, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj)
{
	const char *name = "blasfeo_dsyrk_dpotrf_ln";
	Single S;
	if (!single_init(S, name))
		return;
	blasfeo_amd_dsyrk_dpotrf_arg a;
	a.m = m; a.n = m; a.k = k;
	a.ai = ai; a.aj = aj; a.bi = bi; a.bj = bj; a.ci = ci; a.cj = cj; a.di = di; a.dj = dj;
	a.sA = S.add(sA); a.sB = S.add(sB); a.sC = S.add(sC); a.sD = S.add(sD);
	blasfeo_amd_hint h = {m, m, k};
	const int ps = g_ps;
	run_single(S, a, [&](const blasfeo_amd_dsyrk_dpotrf_arg *d) {
		return blasfeo_amd_dsyrk_dpotrf_ln_batched(ps, 1, d, &h, 0);
	}, name);
}

} // extern "C"
