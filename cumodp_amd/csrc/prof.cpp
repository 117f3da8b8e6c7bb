// prof.cpp -- optional per-kernel timing with HIP events.
//
// Off by default (no events, no overhead).  Enabled by
// cumodp_profile_enable(1) or CUMODP_PROFILE=1 in the environment, every
// instrumented launch records an event pair on its own stream; reading the
// totals synchronises those events.  This is the library's counterpart of the
// reference's cudaEvent timer macros (include/cudautils.h:49-74) and per-call
// "FFT-based:" print (src/stockham.cu:585-600), silent unless asked for.
#include <cstdlib>
#include <mutex>
#include <vector>

#include "../../include/cumodp_amd.h"
#include "internal.h"

namespace cm {

namespace {

const char *kNames[kProfCount] = {
    "k_col_fwd", "k_row_mul", "k_col_inv", "k_mul_lds", "k_mul_mixed",
    "k_ntt_lds", "k_row_out", "k_gather_scatter", "k_twoconv_split", "k_twoconv_crt",
    "k_col_ntt", "k_elementwise", "k_plain_mul", "k_list_persistent",
};

struct Rec {
    int id;
    hipEvent_t a, b;
};

std::mutex g_m;
std::vector<Rec> g_open, g_done;
std::vector<hipEvent_t> g_pool;   // recorded-and-read events, reused

// Timing events without the system-scope release fence (the bracketed
// kernels' results stay in device memory), taken from a pool so that no
// event is created inside a timed region once the pool is warm.
hipEvent_t take_event_locked() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}
double g_total[kProfCount];
int64_t g_count[kProfCount];
int g_enabled = -1;

bool enabled_locked() {
    if (g_enabled < 0) {
        const char *e = getenv("CUMODP_PROFILE");
        g_enabled = (e && *e && *e != '0') ? 1 : 0;
    }
    return g_enabled == 1;
}

void drain_locked() {
    for (auto &r : g_done) {
        float ms = 0.f;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            g_total[r.id] += ms;
            g_count[r.id] += 1;
        }
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
    }
    g_done.clear();
}

}  // namespace

thread_local bool t_suspended = false;

void prof_suspend(bool on) { t_suspended = on; }

bool prof_on() {
    std::lock_guard<std::mutex> lk(g_m);
    return enabled_locked();
}

ProfScope::ProfScope(int id, hipStream_t s) : id_(-1), s_(s) {
    if (t_suspended) return;
    std::lock_guard<std::mutex> lk(g_m);
    if (!enabled_locked() || id < 0 || id >= kProfCount) return;
    if (stream_capturing(s)) return;   // graphs are timed as a whole by the caller
    hipEvent_t a = take_event_locked();
    if (!a) return;
    if (hipEventRecord(a, s) != hipSuccess) { g_pool.push_back(a); return; }
    id_ = id;
    ev_ = (void *)a;
}

ProfScope::~ProfScope() {
    if (id_ < 0) return;
    std::lock_guard<std::mutex> lk(g_m);
    hipEvent_t b = take_event_locked();
    if (!b) return;
    if (hipEventRecord(b, s_) != hipSuccess) { g_pool.push_back(b); return; }
    g_done.push_back(Rec{id_, (hipEvent_t)ev_, b});
    if (g_done.size() > 4096) drain_locked();
}

}  // namespace cm

extern "C" {

cumodp_err cumodp_profile_enable(int on) {
    std::lock_guard<std::mutex> lk(cm::g_m);
    cm::g_enabled = on ? 1 : 0;
    return CUMODP_SUCCESS;
}

void cumodp_profile_reset(void) {
    std::lock_guard<std::mutex> lk(cm::g_m);
    cm::drain_locked();
    for (int i = 0; i < cm::kProfCount; ++i) {
        cm::g_total[i] = 0;
        cm::g_count[i] = 0;
    }
}

int cumodp_profile_count(void) { return cm::kProfCount; }

const char *cumodp_profile_name(int id) {
    return (id >= 0 && id < cm::kProfCount) ? cm::kNames[id] : "";
}

cumodp_err cumodp_profile_read(int id, double *total_ms, int64_t *count) {
    if (id < 0 || id >= cm::kProfCount) return CUMODP_ASSUMPTION_ERROR;
    std::lock_guard<std::mutex> lk(cm::g_m);
    cm::drain_locked();
    if (total_ms) *total_ms = cm::g_total[id];
    if (count) *count = cm::g_count[id];
    return CUMODP_SUCCESS;
}

}  // extern "C"
