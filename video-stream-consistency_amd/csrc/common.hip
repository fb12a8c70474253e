// Error state, version and defaults of libvsc.
#include "vsc_internal.h"

#include <cstdio>
#include <vector>

namespace vsc {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return VSC_OK;
    return set_error(e == hipErrorOutOfMemory ? VSC_E_NOMEM : VSC_E_HIP,
                     std::string(what) + ": " + hipGetErrorString(e));
}

int check_launch(const char* what)
{
    return check_hip(hipGetLastError(), what);
}

struct ProfSlot {
    hipEvent_t a = nullptr, b = nullptr;
    int kernel = 0;
    double bytes = 0;
};
static std::vector<ProfSlot> g_prof;
static int g_prof_used = 0;
static bool g_prof_on = false;

int prof_start(ProfKernel k, double algorithmic_bytes, hipStream_t s)
{
    if (!g_prof_on || g_prof_used >= (int)g_prof.size()) return -1;
    const int i = g_prof_used++;
    g_prof[i].kernel = k;
    g_prof[i].bytes = algorithmic_bytes;
    if (hipEventRecord(g_prof[i].a, s) != hipSuccess) return -1;
    return i;
}

void prof_stop(int slot, hipStream_t s)
{
    if (slot >= 0) (void)hipEventRecord(g_prof[slot].b, s);
}

}  // namespace vsc

extern "C" int vsc_profile_begin(int capacity)
{
    using namespace vsc;
    if (capacity < 1) return set_error(VSC_E_INVALID, "vsc_profile_begin: capacity < 1");
    for (auto& e : g_prof) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    g_prof.assign(capacity, ProfSlot{});
    for (auto& e : g_prof) {
        int rc = check_hip(hipEventCreate(&e.a), "vsc_profile_begin");
        if (!rc) rc = check_hip(hipEventCreate(&e.b), "vsc_profile_begin");
        if (rc) return rc;
    }
    g_prof_used = 0;
    g_prof_on = true;
    return VSC_OK;
}

extern "C" int vsc_profile_end(const char* kernel, int* launches, double* total_ms, double* algorithmic_bytes)
{
    using namespace vsc;
    g_prof_on = false;
    int k = -1;
    if (kernel && std::string(kernel) == "consist_blocked") k = PROF_CONSIST_BLOCKED;
    if (kernel && std::string(kernel) == "front_pass") k = PROF_FRONT_PASS;
    if (k < 0) return set_error(VSC_E_INVALID, "vsc_profile_end: unknown kernel name");
    int n = 0;
    double ms = 0, bytes = 0;
    for (int i = 0; i < g_prof_used; ++i) {
        if (g_prof[i].kernel != k) continue;
        int rc = check_hip(hipEventSynchronize(g_prof[i].b), "vsc_profile_end");
        if (rc) return rc;
        float t = 0;
        rc = check_hip(hipEventElapsedTime(&t, g_prof[i].a, g_prof[i].b), "vsc_profile_end");
        if (rc) return rc;
        ++n;
        ms += t;
        bytes += g_prof[i].bytes;
    }
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
    if (algorithmic_bytes) *algorithmic_bytes = bytes;
    return VSC_OK;
}

extern "C" const char* vsc_last_error(void) { return vsc::g_last_error.c_str(); }

extern "C" const char* vsc_version(void) { return "vsc-mi355x 0.1 (gfx950)"; }

extern "C" int vsc_default_params(vsc_params* out)
{
    if (!out) return vsc::set_error(VSC_E_INVALID, "vsc_default_params: null output");
    // VideoStabilizer::initHyperParams, videostabilizer.cpp:104-113
    out->alpha = 6800.0f;
    out->beta = 6800.0f;
    out->gamma = 2.0f;
    out->pyramid_levels = 2;
    out->num_iter = 150;
    out->step_size = 0.15f;
    out->mom_fac = 0.15f;
    return VSC_OK;
}
