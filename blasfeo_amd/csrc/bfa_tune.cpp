// bfa_tune.cpp -- process-wide launch-shape knobs (blasfeo_amd_tune_set/get).
// The batched entries read them at launch; an autotuner fills them from
// measurements on the caller's workload.
#include <atomic>
#include <cstring>

#include "bfa_internal.hpp"

namespace bfa {

namespace {
constexpr int NCLASS = 8;
std::atomic<int> g_ric_per_cu[NCLASS];
std::atomic<int> g_ric_threads[NCLASS];
std::atomic<int> g_gemm_grid_y{0};
std::atomic<int> g_potrf_bw{0};
std::atomic<int> g_trsm_thin{0};
std::atomic<int> g_gemm_wave{0};
std::atomic<int> g_lds_threads[NCLASS];

std::atomic<int> *slot(const char *name, int index)
{
	if (!name)
		return nullptr;
	if (!std::strcmp(name, "ric_per_cu"))
		return (index >= 0 && index < NCLASS) ? &g_ric_per_cu[index] : nullptr;
	if (!std::strcmp(name, "ric_threads"))
		return (index >= 0 && index < NCLASS) ? &g_ric_threads[index] : nullptr;
	if (!std::strcmp(name, "gemm_grid_y"))
		return index == 0 ? &g_gemm_grid_y : nullptr;
	if (!std::strcmp(name, "potrf_bw"))
		return index == 0 ? &g_potrf_bw : nullptr;
	if (!std::strcmp(name, "trsm_thin"))
		return index == 0 ? &g_trsm_thin : nullptr;
	if (!std::strcmp(name, "gemm_wave"))
		return index == 0 ? &g_gemm_wave : nullptr;
	if (!std::strcmp(name, "lds_threads"))
		return (index >= 0 && index < NCLASS) ? &g_lds_threads[index] : nullptr;
	return nullptr;
}

bool valid(const char *name, int index, int value)
{
	if (value == 0)
		return true;
	if (!std::strcmp(name, "ric_per_cu") || !std::strcmp(name, "gemm_grid_y"))
		return value >= 1 && value <= 64;
	// potrf_bw: tile columns per block panel of potrf_l for 128 < n <= 256
	if (!std::strcmp(name, "potrf_bw"))
		return value == 1 || value == 2 || value == 4;
	// trsm_thin: 1 = always the one-wavefront left solve when it applies, 2 = never
	// gemm_wave: 1 = the one-wavefront-per-tile gemm when the sizes allow it, 2 = never
	if (!std::strcmp(name, "trsm_thin") || !std::strcmp(name, "gemm_wave"))
		return value == 1 || value == 2;
	// lds_threads: threads per workgroup of the LDS-resident potrf_l(_mn) and
	// dsyrk_dpotrf_ln launch of 16-row bucket index (rows 16 index + 1 ..
	// 16 index + 16); the buckets up to 32 rows always run one wave.  0: 256
	if (!std::strcmp(name, "lds_threads"))
		return index >= 2 && (value == 256 || value == 512);
	// ric_threads: one wave only for the classes whose stages have nu+nx <= 64
	return value == 128 || value == 256 || (value == 64 && index <= 4);
}
} // namespace

int tune_value(TuneKnob knob, int index)
{
	switch (knob)
	{
	case TUNE_RIC_PER_CU:
		return (index >= 0 && index < NCLASS) ? g_ric_per_cu[index].load() : 0;
	case TUNE_RIC_THREADS:
		return (index >= 0 && index < NCLASS) ? g_ric_threads[index].load() : 0;
	case TUNE_GEMM_GRID_Y:
		return g_gemm_grid_y.load();
	case TUNE_POTRF_BW:
		return g_potrf_bw.load();
	case TUNE_TRSM_THIN:
		return g_trsm_thin.load();
	case TUNE_GEMM_WAVE:
		return g_gemm_wave.load();
	case TUNE_LDS_THREADS:
		return (index >= 0 && index < NCLASS) ? g_lds_threads[index].load() : 0;
	}
	return 0;
}

} // namespace bfa

using namespace bfa;

extern "C" int blasfeo_amd_tune_set(const char *name, int index, int value)
{
	std::atomic<int> *s = slot(name, index);
	if (!s)
		return set_error(BLASFEO_AMD_EINVAL, "blasfeo_amd_tune_set: unknown knob %s[%d]", name ? name : "(null)", index);
	if (!valid(name, index, value))
		return set_error(BLASFEO_AMD_EINVAL, "blasfeo_amd_tune_set: value %d not allowed for %s[%d]", value, name, index);
	s->store(value);
	return BLASFEO_AMD_OK;
}

extern "C" int blasfeo_amd_tune_get(const char *name, int index)
{
	std::atomic<int> *s = slot(name, index);
	if (!s)
		return set_error(BLASFEO_AMD_EINVAL, "blasfeo_amd_tune_get: unknown knob %s[%d]", name ? name : "(null)", index);
	return s->load();
}
