// plan.cpp -- device-resident twiddle tables and scratch buffers.
//
// The reference rebuilds its twiddle table W[j] = w^j on every call
// (get_powers_binary, src/fft_aux.cu:87-132, one 4 KB host copy plus e-10
// doubling launches) and cudaMallocs every temporary.  Here each root family
// is built once per process, on the host, with Shoup companions, and kept in
// device memory; scratch space grows and is reused.
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
std::map<std::pair<const Family *, int>, Outer *> g_outer;
std::map<u32, const Family *> g_default;

uint2 *upload(const std::vector<uint2> &h) {
    uint2 *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(uint2) * h.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(uint2) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    return d;
}

// powers r^0 .. r^(cnt-1) with Shoup companions
std::vector<uint2> power_table(u32 r, size_t cnt, u32 p) {
    std::vector<uint2> t(cnt);
    u64 x = 1;
    for (size_t j = 0; j < cnt; ++j) {
        t[j] = make_uint2((u32)x, cm_host_shoup_q((u32)x, p));
        x = x * r % p;
    }
    return t;
}

Family *make_family(u32 p, u32 wN, int K) {
    Family *f = new Family();
    f->p = p;
    f->pinv = cm_host_pinv(p);
    f->K = K;
    f->wN = wN;
    f->LT = K < kMaxLdsLog ? K : kMaxLdsLog;
    u32 wt = cm_host_pow(wN, (u64)1 << (K - f->LT), p);   // order 2^LT
    size_t half = f->LT >= 1 ? ((size_t)1 << (f->LT - 1)) : 1;
    f->fwd = upload(power_table(wt, half, p));
    f->inv = upload(power_table(cm_host_inv(wt, p), half, p));
    if (!f->fwd || !f->inv) {
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

const Outer *outer_for(const Family *f, int k) {
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    auto key = std::make_pair(f, k);
    auto it = g_outer.find(key);
    if (it != g_outer.end()) return it->second;
    if (k > f->K || k <= kOuterLoBits) return nullptr;
    u32 p = f->p;
    u32 wn = cm_host_pow(f->wN, (u64)1 << (f->K - k), p);   // order 2^k
    u32 wni = cm_host_inv(wn, p);
    size_t nlo = (size_t)1 << kOuterLoBits, nhi = (size_t)1 << (k - kOuterLoBits);
    Outer *o = new Outer();
    o->k = k;
    o->lo_f = upload(power_table(wn, nlo, p));
    o->hi_f = upload(power_table(cm_host_pow(wn, nlo, p), nhi, p));
    o->lo_i = upload(power_table(wni, nlo, p));
    o->hi_i = upload(power_table(cm_host_pow(wni, nlo, p), nhi, p));
    if (!o->lo_f || !o->hi_f || !o->lo_i || !o->hi_i) {
        delete o;
        return nullptr;
    }
    g_outer[key] = o;
    return o;
}

void *workspace(int slot, size_t bytes) {
    static std::mutex m;
    static std::map<int, std::pair<void *, size_t>> bufs;
    std::lock_guard<std::mutex> lk(m);
    auto &b = bufs[slot];
    if (b.second >= bytes && b.first) return b.first;
    if (b.first) {
        (void)hipDeviceSynchronize();
        (void)hipFree(b.first);
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

}  // namespace cm
