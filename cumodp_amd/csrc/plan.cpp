// plan.cpp -- device-resident twiddle tables and scratch buffers.
//
// The reference rebuilds its twiddle table W[j] = w^j on every call
// (get_powers_binary, src/fft_aux.cu:87-132, one 4 KB host copy plus e-10
// doubling launches) and cudaMallocs every temporary.  Here each root family
// is built once per process, on the host, with Shoup companions, and kept in
// device memory; scratch space grows and is reused.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "internal.h"

namespace cm {

std::mutex &api_mutex() {
    static std::mutex m;
    return m;
}

namespace {

std::mutex g_plan_mutex;
std::map<std::tuple<u32, u32, int>, Family *> g_families;
std::map<std::tuple<const Family *, int, int>, Outer *> g_outer;
int g_split[kMaxLog + 1];   // chosen k2 per k (0: default)
std::map<u32, const Family *> g_default;

u32 *upload(const std::vector<u32> &h) {
    u32 *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(u32) * h.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(u32) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    return d;
}

// powers r^0 .. r^(cnt-1) in Montgomery form
std::vector<u32> power_table(u32 r, size_t cnt, u32 p) {
    std::vector<u32> t(cnt);
    u64 x = cm_host_mont(1, p);
    for (size_t j = 0; j < cnt; ++j) {
        t[j] = (u32)x;
        x = x * r % p;
    }
    return t;
}

Family *make_family(u32 p, u32 wN, int K) {
    Family *f = new Family();
    f->p = p;
    f->pinv = cm_host_pinv(p);
    f->np = cm_host_npinv(p);
    f->K = K;
    f->wN = wN;
    f->LT = K < kMaxLevelLog ? K : kMaxLevelLog;
    // level tables: level M = 2^j (j = 1 .. LT) at offset 2^(j-1) - 1
    std::vector<u32> lv, lvi;
    for (int j = 1; j <= f->LT; ++j) {
        u32 wm = cm_host_pow(wN, (u64)1 << (K - j), p);   // order 2^j
        std::vector<u32> a = power_table(wm, (size_t)1 << (j - 1), p);
        std::vector<u32> b = power_table(cm_host_inv(wm, p), (size_t)1 << (j - 1), p);
        lv.insert(lv.end(), a.begin(), a.end());
        lvi.insert(lvi.end(), b.begin(), b.end());
    }
    if (lv.empty()) { lv.push_back(cm_host_mont(1, p)); lvi.push_back(cm_host_mont(1, p)); }
    f->lv = upload(lv);
    f->lvi = upload(lvi);
    if (!f->lv || !f->lvi) {
        delete f;
        return nullptr;
    }
    return f;
}

}  // namespace

const Family *family_for_root(u32 p, u32 w, int k) {
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    auto key = std::make_tuple(p, w, k);
    auto it = g_families.find(key);
    if (it != g_families.end()) return it->second;
    Family *f = make_family(p, w, k);
    if (f) g_families[key] = f;
    return f;
}

const Family *family_default(u32 p) {
    {
        std::lock_guard<std::mutex> lk(g_plan_mutex);
        auto it = g_default.find(p);
        if (it != g_default.end()) return it->second;
    }
    int K = cm_fourier_degree(p);
    if (K > 30) K = 30;
    u32 g = cm_host_generator(p);
    u32 wN = cm_host_pow(g, (u64)(p - 1) >> K, p);
    const Family *f = family_for_root(p, wN, K);
    if (f) {
        std::lock_guard<std::mutex> lk(g_plan_mutex);
        g_default[p] = f;
    }
    return f;
}

const Family *family_default_cached(u32 p) {
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    auto it = g_default.find(p);
    return it != g_default.end() ? it->second : nullptr;
}

void four_step_split(int k, int &k1, int &k2) {
    int c = (k >= 0 && k <= kMaxLog) ? g_split[k] : 0;
    if (!c) {
        char name[32];
        snprintf(name, sizeof(name), "CUMODP_SPLIT_%d", k);
        const char *e = getenv(name);
        if (e) c = atoi(e);
    }
    // measured defaults on MI355X (cumodp_amd.autotune, back-to-back 2^k
    // multiplies): 2^12-word rows up to 2^20, 2^13 for 2^21..2^24, 2^14 above
    if (c < 12 || c > 14 || k - c < 1 || k - c > kMaxLdsLog)
        c = k >= 25 ? 14 : (k >= 21 ? 13 : 12);
    k2 = c;
    k1 = k - k2;
}

bool set_four_step_split(int k, int k2) {
    if (k <= kMaxLdsLog || k > kMaxLog || k2 < 12 || k2 > 14 || k - k2 < 1 ||
        k - k2 > kMaxLdsLog)
        return false;
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    g_split[k] = k2;
    return true;
}

const Outer *outer_for(const Family *f, int k) {
    int k1, k2;
    four_step_split(k, k1, k2);
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    auto key = std::make_tuple(f, k, k2);
    auto it = g_outer.find(key);
    if (it != g_outer.end()) return it->second;
    if (k > f->K || k <= kOuterLoBits) return nullptr;
    u32 p = f->p;
    u32 wn = cm_host_pow(f->wN, (u64)1 << (f->K - k), p);   // order 2^k
    u32 wni = cm_host_inv(wn, p);
    size_t nlo = (size_t)1 << kOuterLoBits, nhi = (size_t)1 << (k - kOuterLoBits);
    Outer *o = new Outer();
    o->k = k;
    o->k1 = k1;
    o->k2 = k2;
    o->lo_f = upload(power_table(wn, nlo, p));
    o->hi_f = upload(power_table(cm_host_pow(wn, nlo, p), nhi, p));
    o->lo_i = upload(power_table(wni, nlo, p));
    o->hi_i = upload(power_table(cm_host_pow(wni, nlo, p), nhi, p));
    // row-pass tables; the inverse rb table also carries 2^32 / n, the scale
    // the row pass leaves out of its pointwise product (one Montgomery
    // product gives F G / 2^32, the inverse transform a factor n)
    const size_t n1 = (size_t)1 << k1, T = ((size_t)1 << k2) / 16;
    const u32 unscale = cm_host_mul(cm_host_r1(p), cm_host_inv(cm_host_pow(2, k, p), p), p);
    for (int dir = 0; dir < 2; ++dir) {
        const u32 w = dir ? wni : wn;
        std::vector<u32> rb(n1 * T), rc(n1 * 16);
        for (size_t r = 0; r < n1; ++r) {
            const u32 wr = cm_host_pow(w, r, p);                 // w^k1
            u64 x = cm_host_mont(dir ? unscale : 1u, p);
            for (size_t t = 0; t < T; ++t) {
                rb[r * T + t] = (u32)x;
                x = x * wr % p;
            }
            const u32 wrT = cm_host_pow(wr, T, p);              // w^(k1 T)
            x = cm_host_mont(1, p);
            for (size_t m = 0; m < 16; ++m) {
                rc[r * 16 + m] = (u32)x;
                x = x * wrT % p;
            }
        }
        (dir ? o->rb_i : o->rb_f) = upload(rb);
        (dir ? o->rc_i : o->rc_f) = upload(rc);
    }
    if (!o->lo_f || !o->hi_f || !o->lo_i || !o->hi_i || !o->rb_f || !o->rc_f || !o->rb_i ||
        !o->rc_i) {
        delete o;
        return nullptr;
    }
    g_outer[key] = o;
    return o;
}

const Outer *outer_cached(const Family *f, int k) {
    int k1, k2;
    four_step_split(k, k1, k2);
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    auto it = g_outer.find(std::make_tuple(f, k, k2));
    return it != g_outer.end() ? it->second : nullptr;
}

namespace {
std::mutex g_ws_mutex;
std::map<std::pair<int, hipStream_t>, std::pair<void *, size_t>> g_ws;
// Scratch buffers that captured graphs refer to: buffer -> number of live
// graphs.  Such a buffer is never freed while a graph holds it; when its
// stream needs a larger one (or is destroyed) it is only detached
// (g_ws_detached) and freed when the last graph holding it is destroyed.
std::map<void *, int> g_ws_graph_refs;
std::map<void *, bool> g_ws_detached;

// drops buffer b from its (slot, stream) entry: freed at once unless a live
// graph refers to it (caller holds g_ws_mutex; work queued on the stream
// that may still use b has been synchronised)
void retire_buffer_locked(void *b) {
    if (!b) return;
    auto r = g_ws_graph_refs.find(b);
    if (r != g_ws_graph_refs.end() && r->second > 0) {
        g_ws_detached[b] = true;
        return;
    }
    (void)hipFree(b);
}
}  // namespace

bool stream_capturing(hipStream_t s) {
    if (!s) return false;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st == hipStreamCaptureStatusActive;
}

void *workspace(int slot, size_t bytes, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    auto &b = g_ws[std::make_pair(slot, s)];
    if (b.second >= bytes && b.first) return b.first;
    // a graph being captured cannot grow scratch (it would synchronise and
    // free memory the graph refers to): the same calls must run once eagerly
    // on this stream first
    if (stream_capturing(s)) return nullptr;
    if (b.first) {
        // work already queued on s may still use the old buffer
        (void)hipStreamSynchronize(s);
        retire_buffer_locked(b.first);
        b.first = nullptr;
        b.second = 0;
    }
    size_t sz = bytes < 4096 ? 4096 : bytes;
    void *d = nullptr;
    if (hipMalloc(&d, sz) != hipSuccess) return nullptr;
    b.first = d;
    b.second = sz;
    return d;
}

namespace {
// Pinned staging for descriptor uploads, one per stream; reused once the
// previous upload from it has finished.
struct Staging {
    void *host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
};
// A ring of staging buffers per stream: a buffer is rewritten only once the
// copy that last read it has run, so the host can queue several calls ahead
// of the device without waiting.
constexpr int kStagingRing = 8;
struct StagingRing {
    Staging b[kStagingRing];
    int next = 0;
};
std::map<hipStream_t, StagingRing> g_staging;   // guarded by g_ws_mutex
// pinned buffers whose copies were recorded into a graph being captured on a
// stream; they belong to that graph once the capture ends
std::map<hipStream_t, std::vector<void *>> g_captured;
}  // namespace

hipError_t stage_upload(const void *src, size_t bytes, void *dst, hipStream_t s) {
    return stage_upload_fill(bytes, dst, s, [&](char *h) { memcpy(h, src, bytes); });
}

hipError_t stage_upload_fill(size_t bytes, void *dst, hipStream_t s,
                             const std::function<void(char *)> &fill) {
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    hipError_t e;
    if (stream_capturing(s)) {
        // the recorded copy must keep reading these bytes at every replay
        void *h = nullptr;
        e = hipHostMalloc(&h, bytes ? bytes : 1);
        if (e != hipSuccess) return e;
        fill((char *)h);
        g_captured[s].push_back(h);
        return hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, s);
    }
    // prefer a buffer that is already large enough and whose last copy has
    // run; else an unused slot; else wait for the oldest
    StagingRing &ring = g_staging[s];
    int pick = -1;
    for (int i = 0; i < kStagingRing && pick < 0; ++i) {
        const int j = (ring.next + i) % kStagingRing;
        Staging &b = ring.b[j];
        if (b.done && b.cap >= bytes && hipEventQuery(b.done) == hipSuccess) pick = j;
    }
    for (int i = 0; i < kStagingRing && pick < 0; ++i) {
        const int j = (ring.next + i) % kStagingRing;
        if (!ring.b[j].done) pick = j;
    }
    if (pick < 0) pick = ring.next;
    ring.next = (pick + 1) % kStagingRing;
    Staging &st = ring.b[pick];
    if (!st.done) {
        e = hipEventCreateWithFlags(&st.done, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    } else {
        e = hipEventSynchronize(st.done);
        if (e != hipSuccess) return e;
    }
    if (st.cap < bytes) {
        if (st.host) (void)hipHostFree(st.host);
        st.host = nullptr;
        st.cap = 0;
        e = hipHostMalloc(&st.host, bytes);
        if (e != hipSuccess) return e;
        st.cap = bytes;
    }
    fill((char *)st.host);
    e = hipMemcpyAsync(dst, st.host, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipEventRecord(st.done, s);
}

namespace {
std::map<hipStream_t, AuxStreams *> g_aux;   // guarded by g_ws_mutex
}  // namespace

AuxStreams *aux_streams(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    auto it = g_aux.find(s);
    if (it != g_aux.end()) return it->second;
    if (stream_capturing(s)) return nullptr;   // created eagerly, never inside a capture
    AuxStreams *a = new AuxStreams();
    bool ok = hipEventCreateWithFlags(&a->fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < kAuxStreams && ok; ++i) {
        ok = hipStreamCreate(&a->s[i]) == hipSuccess &&
             hipEventCreateWithFlags(&a->join[i], hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) return nullptr;   // leaks the partial set; the caller runs serially
    g_aux[s] = a;
    return a;
}

std::vector<void *> take_captured_buffers(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    std::vector<void *> v;
    auto it = g_captured.find(s);
    if (it != g_captured.end()) {
        v.swap(it->second);
        g_captured.erase(it);
    }
    return v;
}

std::vector<void *> hold_workspace_for_graph(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    std::vector<hipStream_t> streams{s};
    auto a = g_aux.find(s);
    if (a != g_aux.end())
        for (int i = 0; i < kAuxStreams; ++i) streams.push_back(a->second->s[i]);
    std::vector<void *> held;
    for (auto &kv : g_ws) {
        if (!kv.second.first) continue;
        for (hipStream_t t : streams) {
            if (kv.first.second == t) {
                held.push_back(kv.second.first);
                ++g_ws_graph_refs[kv.second.first];
                break;
            }
        }
    }
    return held;
}

void release_graph_workspace(const std::vector<void *> &held) {
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    for (void *b : held) {
        auto r = g_ws_graph_refs.find(b);
        if (r == g_ws_graph_refs.end()) continue;
        if (--r->second > 0) continue;
        g_ws_graph_refs.erase(r);
        auto d = g_ws_detached.find(b);
        if (d != g_ws_detached.end()) {
            g_ws_detached.erase(d);
            (void)hipFree(b);
        }
    }
}

void release_workspace(hipStream_t s) {
    AuxStreams *a = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_ws_mutex);
        auto it = g_aux.find(s);
        if (it != g_aux.end()) {
            a = it->second;
            g_aux.erase(it);
        }
    }
    if (a) {
        for (int i = 0; i < kAuxStreams; ++i) {
            (void)hipStreamSynchronize(a->s[i]);
            release_workspace(a->s[i]);
            (void)hipStreamDestroy(a->s[i]);
            (void)hipEventDestroy(a->join[i]);
        }
        (void)hipEventDestroy(a->fork);
        delete a;
    }
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    for (auto it = g_ws.begin(); it != g_ws.end();) {
        if (it->first.second == s) {
            retire_buffer_locked(it->second.first);
            it = g_ws.erase(it);
        } else {
            ++it;
        }
    }
    auto st = g_staging.find(s);
    if (st != g_staging.end()) {
        for (Staging &b : st->second.b) {
            if (b.done) (void)hipEventDestroy(b.done);
            if (b.host) (void)hipHostFree(b.host);
        }
        g_staging.erase(st);
    }
}

}  // namespace cm
