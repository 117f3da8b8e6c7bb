// api.cpp -- the C ABI (include/cumodp.h, include/cumodp_amd.h) and the
// C++-linkage drop-ins (include/stockham.h, list_stockham.h, fft_aux.h,
// list_inv_stockham.h) over the gfx950 kernels.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cumodp.h"
#include "../../include/cumodp_amd.h"
#include "../../include/fastPolyEvaluation.h"
#include "../../include/fft_aux.h"
#include "../../include/list_inv_stockham.h"
#include "../../include/list_stockham.h"
#include "../../include/opt_plain_mul.h"
#include "../../include/stockham.h"
#include "internal.h"

using namespace cm;

namespace {

thread_local std::string g_last_error;

cumodp_err fail(cumodp_err code, const char *what, hipError_t e = hipSuccess) {
    g_last_error = what;
    if (e != hipSuccess) {
        g_last_error += ": ";
        g_last_error += hipGetErrorString(e);
    }
    return code;
}

int ceil_log2(int64_t s) {
    return s <= 1 ? 0 : 64 - __builtin_clzll((unsigned long long)(s - 1));
}

hipStream_t as_stream(void *s) { return (hipStream_t)s; }

// Checks shared by every transform entry point.
cumodp_err check_prime_size(u32 p, int k) {
    if (p < 3 || (p & 1) == 0 || p >= (1u << 31)) return fail(CUMODP_ASSUMPTION_ERROR, "modulus must be an odd prime below 2^31");
    if (k > kMaxLog) return fail(CUMODP_FFT_SIZE_TOO_LARGE, "transform size above 2^26");
    if (k > cm_fourier_degree(p)) return fail(CUMODP_FOURIER_DEGREE_TOO_SMALL, "2^k does not divide p-1");
    return CUMODP_SUCCESS;
}

// ---- univariate multiply --------------------------------------------------
// Hd[0 .. nout) = Fd * Gd mod p; Fd, Gd hold lf, lg words (device).
cumodp_err mul_dev_impl(u32 *Hd, int64_t nout, const u32 *Fd, int64_t lf, const u32 *Gd,
                        int64_t lg, u32 p, hipStream_t s) {
    const int64_t slen = lf + lg - 1;
    int e = ceil_log2(slen);
    if (e < 4) e = 4;   // smallest workgroup transform; zero padding is harmless
    cumodp_err ce = check_prime_size(p, e);
    if (ce != CUMODP_SUCCESS) return ce;
    const Family *fam = family_default(p);
    if (!fam) return fail(CUMODP_OUT_OF_MEMORY, "twiddle tables");
    const int64_t n = (int64_t)1 << e;
    hipError_t he;
    if (e <= kMaxLdsLog) {
        MulDesc d;
        d.f = 0; d.g = 0; d.h = 0;
        d.lf = (int)lf; d.lg = (int)lg; d.lh = (int)(nout < n ? nout : n); d.logn = e;
        he = launch_mul_lds_one(e, Fd, Gd, Hd, nullptr, 1, d, fam, s);
    } else {
        // the column kernels read the operands with zero padding and write
        // the first nout product words straight to Hd
        u32 *wf = (u32 *)workspace(0, sizeof(u32) * (size_t)n, s);
        u32 *wg = (u32 *)workspace(1, sizeof(u32) * (size_t)n, s);
        if (!wf || !wg) return fail(CUMODP_OUT_OF_MEMORY, "multiply workspace");
        const int64_t nw = nout < n ? nout : n;
        FourStepIO io{Fd, Gd, lf, lg, 0, wf, wg, Hd, nw, 0, 1};
        he = launch_mul_four_step(e, io, fam, s);
    }
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "multiply", he);
    return CUMODP_SUCCESS;
}

cumodp_err mul_host_impl(sfixn dh, sfixn *H, sfixn df, const sfixn *F, sfixn dg, const sfixn *G,
                         sfixn p) {
    if (df < 0 || dg < 0 || dh < 0) return fail(CUMODP_ASSUMPTION_ERROR, "negative degree");
    const int64_t s = (int64_t)df + dg + 1;
    const int e = ceil_log2(s);
    const int64_t n = (int64_t)1 << e;
    if (n < 256) return fail(CUMODP_FFT_SIZE_TOO_SMALL, "fft size < 256 is not supported by gpu");
    cumodp_err ce = check_prime_size((u32)p, e);
    if (ce != CUMODP_SUCCESS) return ce;
    std::lock_guard<std::mutex> lk(api_mutex());
    const int64_t nout = (int64_t)dh + 1 < n ? (int64_t)dh + 1 : n;
    u32 *buf = (u32 *)workspace(2, sizeof(u32) * (size_t)(df + 1 + dg + 1 + nout), 0);
    if (!buf) return fail(CUMODP_OUT_OF_MEMORY, "staging buffer");
    u32 *dF = buf, *dG = buf + df + 1, *dH = dG + dg + 1;
    hipStream_t st = 0;
    hipError_t he = hipMemcpyAsync(dF, F, sizeof(u32) * (size_t)(df + 1), hipMemcpyHostToDevice, st);
    if (he == hipSuccess) he = hipMemcpyAsync(dG, G, sizeof(u32) * (size_t)(dg + 1), hipMemcpyHostToDevice, st);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy in", he);
    ce = mul_dev_impl(dH, nout, dF, df + 1, dG, dg + 1, (u32)p, st);
    if (ce != CUMODP_SUCCESS) return ce;
    he = hipMemcpyAsync(H, dH, sizeof(u32) * (size_t)nout, hipMemcpyDeviceToHost, st);
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy out", he);
    return CUMODP_SUCCESS;
}

// ---- schoolbook multiply (any odd modulus) --------------------------------
// Products up to 2^30 output coefficients; operands must be canonical.
constexpr int64_t kPlainMaxOut = (int64_t)1 << 30;

cumodp_err check_plain_modulus(u32 p) {
    if (p < 3 || (p & 1) == 0 || p >= (1u << 31))
        return fail(CUMODP_ASSUMPTION_ERROR, "modulus must be odd, in [3, 2^31)");
    return CUMODP_SUCCESS;
}

// descs[i] for F[f .. f+lf) x G[g .. g+lg) -> H[h .. h+lf+lg-1)
cumodp_err plain_dev_impl(std::vector<MulDesc> &descs, const u32 *F, const u32 *G, u32 *H, u32 p,
                          hipStream_t s) {
    cumodp_err ce = check_plain_modulus(p);
    if (ce != CUMODP_SUCCESS) return ce;
    for (MulDesc &d : descs) {
        if (d.lf < 1 || d.lg < 1) return fail(CUMODP_ASSUMPTION_ERROR, "empty operand");
        if ((int64_t)d.lf + d.lg - 1 > kPlainMaxOut)
            return fail(CUMODP_ASSUMPTION_ERROR, "plain product longer than 2^30");
        d.lh = d.lf + d.lg - 1;
        d.logn = d.lg > d.lf ? 1 : 0;   // sweep the shorter operand
    }
    hipError_t he = launch_plain_list(descs.data(), (int64_t)descs.size(), F, G, H, p, s);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "plain multiply", he);
    return CUMODP_SUCCESS;
}

// Shorter operands than this many coefficients go to the schoolbook kernel
// in mulCUDA; longer ones to the NTT multiply when p has the roots for it.
constexpr int64_t kPlainCrossover = 128;

// mulCUDA semantics (src/opt_plain_mul.cu:258-384): C[0 .. n+m-2] = A * B
// mod p for host arrays of n and m coefficients.
cumodp_err plain_host_impl(const sfixn *A, sfixn n, const sfixn *B, sfixn m, sfixn *C, sfixn p) {
    if (n < 1 || m < 1) return fail(CUMODP_ASSUMPTION_ERROR, "empty operand");
    cumodp_err ce = check_plain_modulus((u32)p);
    if (ce != CUMODP_SUCCESS) return ce;
    const int64_t nc = (int64_t)n + m - 1;
    if (nc > kPlainMaxOut) return fail(CUMODP_ASSUMPTION_ERROR, "product longer than 2^30");
    const int e = ceil_log2(nc);
    const bool ntt = std::min(n, m) >= kPlainCrossover && e <= kMaxLog &&
                     e <= cm_fourier_degree((u32)p);
    std::lock_guard<std::mutex> lk(api_mutex());
    u32 *buf = (u32 *)workspace(2, sizeof(u32) * (size_t)(n + m + nc), 0);
    if (!buf) return fail(CUMODP_OUT_OF_MEMORY, "staging buffer");
    u32 *dA = buf, *dB = buf + n, *dC = dB + m;
    hipStream_t st = 0;
    hipError_t he = hipMemcpyAsync(dA, A, sizeof(u32) * (size_t)n, hipMemcpyHostToDevice, st);
    if (he == hipSuccess) he = hipMemcpyAsync(dB, B, sizeof(u32) * (size_t)m, hipMemcpyHostToDevice, st);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy in", he);
    if (ntt) {
        ce = mul_dev_impl(dC, nc, dA, n, dB, m, (u32)p, st);
    } else {
        std::vector<MulDesc> d(1);
        d[0].f = 0; d[0].g = 0; d[0].h = 0; d[0].lf = n; d[0].lg = m;
        ce = plain_dev_impl(d, dA, dB, dC, (u32)p, st);
    }
    if (ce != CUMODP_SUCCESS) return ce;
    he = hipMemcpyAsync(C, dC, sizeof(u32) * (size_t)nc, hipMemcpyDeviceToHost, st);
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy out", he);
    return CUMODP_SUCCESS;
}

// ---- transforms ------------------------------------------------------------
// m rows of 2^k words in place: natural-order DFT with root w (inverse: root
// w^-1 and scaling by 2^-k).
cumodp_err ntt_dev_impl(u32 *X, int64_t m, int k, u32 w, u32 p, bool inverse, hipStream_t s) {
    if (k < 4) return fail(CUMODP_FFT_SIZE_TOO_SMALL, "transform size below 16");
    cumodp_err ce = check_prime_size(p, k);
    if (ce != CUMODP_SUCCESS) return ce;
    const Family *fam = family_for_root(p, w % p, k);
    if (!fam) return fail(CUMODP_OUT_OF_MEMORY, "twiddle tables");
    hipError_t he;
    if (k <= kMaxLdsLog) {
        const u32 sc = inverse ? cm_host_inv(cm_host_pow(2, k, p), p) : 1u;
        he = launch_ntt_lds(k, X, m, inverse, sc, fam, s);
    } else {
        const int64_t n = (int64_t)1 << k;
        if (m > INT_MAX) return fail(CUMODP_FFT_SIZE_TOO_LARGE, "too many rows");
        u32 *Y = (u32 *)workspace(3, sizeof(u32) * (size_t)(n * m), s);
        if (!Y) return fail(CUMODP_OUT_OF_MEMORY, "transform workspace");
        he = launch_ntt_four_step(k, X, Y, (int)m, inverse, fam, s);
    }
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "transform", he);
    return CUMODP_SUCCESS;
}

cumodp_err ntt_host_impl(sfixn *X, int64_t m, int k, u32 w, u32 p, bool inverse) {
    if (k < 4) return fail(CUMODP_FFT_SIZE_TOO_SMALL, "transform size below 16");
    std::lock_guard<std::mutex> lk(api_mutex());
    const size_t bytes = sizeof(u32) * (size_t)(m << k);
    u32 *d = (u32 *)workspace(4, bytes, 0);
    if (!d) return fail(CUMODP_OUT_OF_MEMORY, "staging buffer");
    hipError_t he = hipMemcpy(d, X, bytes, hipMemcpyHostToDevice);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy in", he);
    cumodp_err ce = ntt_dev_impl(d, m, k, w, p, inverse, 0);
    if (ce != CUMODP_SUCCESS) return ce;
    he = hipMemcpy(X, d, bytes, hipMemcpyDeviceToHost);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy out", he);
    return CUMODP_SUCCESS;
}

// DFT_{m,n} on a row-major 2^em x 2^en device array: rows with root wn,
// then columns with root wm (src/list_stockham.cu:1161-1204); the inverse
// uses the inverse roots and scales by (mn)^-1.
// q such arrays, one after the other (I_q (x) DFT_{m,n},
// src/list_stockham.cu:1803-1846): the row transforms of all q m rows in one
// call, the column transforms of all q arrays in one launch per 65535 arrays.
cumodp_err biv_dev_impl(u32 *X, int em, u32 wm, int en, u32 wn, u32 p, bool inverse,
                        hipStream_t s, int64_t q = 1) {
    if (em < 1 || em > kMaxLdsLog) return fail(CUMODP_FFT_SIZE_TOO_LARGE, "2^em must lie in [2, 2^13]");
    if (q < 1) return fail(CUMODP_ASSUMPTION_ERROR, "need at least one array");
    cumodp_err ce = check_prime_size(p, em);
    if (ce != CUMODP_SUCCESS) return ce;
    if (inverse) { wm = cm_host_inv(wm % p, p); wn = cm_host_inv(wn % p, p); }
    ce = ntt_dev_impl(X, q << em, en, wn, p, false, s);
    if (ce != CUMODP_SUCCESS) return ce;
    const Family *fm = family_for_root(p, wm % p, em);
    if (!fm) return fail(CUMODP_OUT_OF_MEMORY, "twiddle tables");
    const u32 sc = inverse ? cm_host_inv(cm_host_pow(2, em + en, p), p) : 1u;
    const int64_t words = (int64_t)1 << (em + en);
    hipError_t he = hipSuccess;
    for (int64_t b0 = 0; b0 < q && he == hipSuccess; b0 += 65535) {
        const int nb = (int)std::min<int64_t>(65535, q - b0);
        he = launch_col_ntt(em, X + b0 * words, (int64_t)1 << en, sc, fm, s, nb, words);
    }
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "bivariate transform", he);
    return CUMODP_SUCCESS;
}

cumodp_err biv_host_impl(sfixn *X, int em, u32 wm, int en, u32 wn, u32 p, bool inverse) {
    if (em < 1 || en < 4 || em + en > 30) return fail(CUMODP_ASSUMPTION_ERROR, "bivariate sizes");
    std::lock_guard<std::mutex> lk(api_mutex());
    const size_t bytes = sizeof(u32) << (em + en);
    u32 *d = (u32 *)workspace(4, bytes, 0);
    if (!d) return fail(CUMODP_OUT_OF_MEMORY, "staging buffer");
    hipError_t he = hipMemcpy(d, X, bytes, hipMemcpyHostToDevice);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy in", he);
    cumodp_err ce = biv_dev_impl(d, em, wm, en, wn, p, inverse, 0);
    if (ce != CUMODP_SUCCESS) return ce;
    he = hipMemcpy(X, d, bytes, hipMemcpyDeviceToHost);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy out", he);
    return CUMODP_SUCCESS;
}

// root of order 2^k of the library's default family
u32 default_root(u32 p, int k) {
    const Family *f = family_default(p);
    if (!f || k > f->K) return 0;
    return cm_host_pow(f->wN, (u64)1 << (f->K - k), p);
}

u32 read_root(const sfixn *W) {
    sfixn w = 1;
    (void)hipMemcpy(&w, W + 1, sizeof(w), hipMemcpyDeviceToHost);
    return (u32)w;
}

// the root of order 2^k of a power table [1, w, ..., w^(2^(k-1)-1)]: W[1],
// except for k = 1, whose one-word table [1] leaves w = -1 implied
u32 read_root_k(const sfixn *W, int k, u32 p) { return k == 1 ? p - 1 : read_root(W); }

void report(const char *fn, cumodp_err e) {
    if (e != CUMODP_SUCCESS) fprintf(stderr, "%s: %s\n", fn, g_last_error.c_str());
}

// The reference prints timings to stdout on every call
// (src/stockham.cu:600, src/two_convolution_poly_mul.cu:30,50,
// src/opt_plain_mul.cu:367-369); here they appear only with CUMODP_VERBOSE
// set to a non-zero value.
bool verbose() {
    static const bool v = [] {
        const char *e = getenv("CUMODP_VERBOSE");
        return e && *e && *e != '0';
    }();
    return v;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

// ===========================================================================
// C linkage: include/cumodp.h
// ===========================================================================
extern "C" {

int is_cuda_enabled(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c > 0;
}

int num_of_cuda_devices(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int is_double_float_enabled(void) { return is_cuda_enabled(); }

static size_t total_mem() {
    size_t fr = 0, tot = 0;
    if (!is_cuda_enabled() || hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
    return tot;
}

unsigned int global_memory_in_bytes(void) {
    size_t t = total_mem();
    return t > UINT_MAX ? UINT_MAX : (unsigned int)t;
}

float global_memory_in_megabytes(void) { return (float)(total_mem() >> 20); }

int can_call_fftmul_uni(sfixn df, sfixn dg) {
    size_t mem = total_mem();
    if (!mem || df < 0 || dg < 0) return 0;
    int b = ceil_log2((int64_t)df + dg + 1);
    if (b > kMaxLog) return 0;
    int lm = 0;
    while (((size_t)2 << lm) <= mem / 4) ++lm;   // floor(log2(mem/4))
    return lm > b + 2;
}

int can_call_subres2(sfixn, sfixn, sfixn) { return 0; }
int can_call_subres3(sfixn, sfixn, sfixn, sfixn) { return 0; }

cumodp_err cumodp_fft_uni(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p) {
    if (n != ((sfixn)1 << k)) return fail(CUMODP_ASSUMPTION_ERROR, "n != 2^k");
    return ntt_host_impl(X, 1, k, (u32)w, (u32)p, false);
}

cumodp_err cumodp_invfft_uni(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p) {
    if (n != ((sfixn)1 << k)) return fail(CUMODP_ASSUMPTION_ERROR, "n != 2^k");
    return ntt_host_impl(X, 1, k, (u32)w, (u32)p, true);
}

cumodp_err cumodp_fftmul_uni(sfixn dh, sfixn *H, sfixn df, const sfixn *F, sfixn dg,
                             const sfixn *G, sfixn p) {
    return mul_host_impl(dh, H, df, F, dg, G, p);
}

cumodp_err cumodp_fft_bivariate(sfixn *X, sfixn em, sfixn wm, sfixn en, sfixn wn, sfixn p) {
    return biv_host_impl(X, em, (u32)wm, en, (u32)wn, (u32)p, false);
}

cumodp_err cumodp_invfft_bivariate(sfixn *X, sfixn em, sfixn wm, sfixn en, sfixn wn, sfixn p) {
    return biv_host_impl(X, em, (u32)wm, en, (u32)wn, (u32)p, true);
}

sfixn cumodp_subres_chain2_fine(sfixn *, sfixn, sfixn, sfixn, sfixn, const sfixn *, sfixn, sfixn,
                                const sfixn *, sfixn) { return -1; }
sfixn cumodp_subres_chain2_coarse(sfixn *, sfixn, sfixn, sfixn, sfixn, const sfixn *, sfixn,
                                  sfixn, const sfixn *, sfixn) { return -1; }
sfixn cumodp_subres_chain3_fine(sfixn *, sfixn, sfixn, sfixn, sfixn, sfixn, sfixn, sfixn,
                                const sfixn *, sfixn, sfixn, sfixn, const sfixn *, sfixn) { return -1; }
sfixn cumodp_subres_chain3_coarse(sfixn *, sfixn, sfixn, sfixn, sfixn, sfixn, sfixn, sfixn,
                                  const sfixn *, sfixn, sfixn, sfixn, const sfixn *, sfixn) { return -1; }
void *init_cuda_scube(sfixn, const sfixn *, const sfixn *, sfixn) { return nullptr; }
void free_cuda_scube(void *) {}
void print_cuda_scube(const void *) {}
cumodp_err build_cuda_scube(void *, const sfixn *, const sfixn *, const sfixn *, const sfixn *) {
    return CUMODP_FAILURE;
}
const sfixn *interp_subres_coeff2(sfixn *, void *, sfixn, sfixn) { return nullptr; }
const sfixn *interp_subres_coeff3(sfixn *, sfixn *, void *, sfixn, sfixn) { return nullptr; }
float gcdPrimeField(sfixn *, sfixn *, sfixn *, sfixn *, sfixn, sfixn, sfixn, sfixn) {
    fprintf(stderr, "gcdPrimeField: not provided by this library (outputs untouched)\n");
    return -1.0f;
}
float divPrimeField(sfixn *, sfixn *, sfixn *, sfixn *, sfixn, sfixn, sfixn) {
    fprintf(stderr, "divPrimeField: not provided by this library (outputs untouched)\n");
    return -1.0f;
}

// ===========================================================================
// C linkage: include/cumodp_amd.h
// ===========================================================================
void *cumodp_dev_alloc(size_t bytes) {
    void *d = nullptr;
    hipError_t e = hipMalloc(&d, bytes ? bytes : 4);
    if (e != hipSuccess) { fail(CUMODP_OUT_OF_MEMORY, "hipMalloc", e); return nullptr; }
    return d;
}

cumodp_err cumodp_dev_free(void *ptr) {
    hipError_t e = hipFree(ptr);
    return e == hipSuccess ? CUMODP_SUCCESS : fail(CUMODP_FAILURE, "hipFree", e);
}

cumodp_err cumodp_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    return e == hipSuccess ? CUMODP_SUCCESS : fail(CUMODP_FAILURE, "memcpy h2d", e);
}

cumodp_err cumodp_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? CUMODP_SUCCESS : fail(CUMODP_FAILURE, "memcpy d2h", e);
}

cumodp_err cumodp_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice);
    return e == hipSuccess ? CUMODP_SUCCESS : fail(CUMODP_FAILURE, "memcpy d2d", e);
}

cumodp_err cumodp_synchronize(void) {
    hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? CUMODP_SUCCESS : fail(CUMODP_FAILURE, "synchronize", e);
}

void *cumodp_stream_create(void) {
    hipStream_t s = nullptr;
    if (hipStreamCreate(&s) != hipSuccess) return nullptr;
    return (void *)s;
}

cumodp_err cumodp_stream_destroy(void *stream) {
    hipError_t e = hipStreamSynchronize(as_stream(stream));
    if (e != hipSuccess) return fail(CUMODP_FAILURE, "stream destroy", e);
    release_workspace(as_stream(stream));
    e = hipStreamDestroy(as_stream(stream));
    return e == hipSuccess ? CUMODP_SUCCESS : fail(CUMODP_FAILURE, "stream destroy", e);
}

int cumodp_stream_query(void *stream) {
    hipError_t e = hipStreamQuery(as_stream(stream));
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    fail(CUMODP_FAILURE, "stream query", e);
    return -1;
}

cumodp_err cumodp_stream_synchronize(void *stream) {
    hipError_t e = hipStreamSynchronize(as_stream(stream));
    return e == hipSuccess ? CUMODP_SUCCESS : fail(CUMODP_FAILURE, "stream synchronize", e);
}

const char *cumodp_last_error(void) { return g_last_error.c_str(); }

cumodp_err cumodp_fftmul_uni_dev(sfixn dh, sfixn *Hd, sfixn df, const sfixn *Fd, sfixn dg,
                                 const sfixn *Gd, sfixn p, void *stream) {
    if (df < 0 || dg < 0 || dh < 0) return fail(CUMODP_ASSUMPTION_ERROR, "negative degree");
    std::lock_guard<std::mutex> lk(api_mutex());
    const int64_t s = (int64_t)df + dg + 1;
    int e = ceil_log2(s);
    if (e < 4) e = 4;
    const int64_t n = (int64_t)1 << e;
    const int64_t nout = (int64_t)dh + 1 < n ? (int64_t)dh + 1 : n;
    return mul_dev_impl((u32 *)Hd, nout, (const u32 *)Fd, df + 1, (const u32 *)Gd, dg + 1,
                        (u32)p, as_stream(stream));
}

// The same product with no allocation: twiddle tables are looked up, never
// built, and the four-step work arrays are the file-scope device scratch of
// noalloc.hip.  Per-kernel timing is suspended (it would create events).
cumodp_err cumodp_fftmul_uni_dev_noalloc(sfixn dh, sfixn *Hd, sfixn df, const sfixn *Fd,
                                         sfixn dg, const sfixn *Gd, sfixn p, void *stream) {
    if (df < 0 || dg < 0 || dh < 0) return fail(CUMODP_ASSUMPTION_ERROR, "negative degree");
    const int64_t slen = (int64_t)df + dg + 1;
    int e = ceil_log2(slen);
    if (e < 4) e = 4;
    if (e > kNoAllocMaxLog)
        return fail(CUMODP_FFT_SIZE_TOO_LARGE,
                    "non-allocating multiply: df + dg + 1 above 2^22 (its scratch holds two "
                    "2^22-word arrays)");
    cumodp_err ce = check_prime_size((u32)p, e);
    if (ce != CUMODP_SUCCESS) return ce;
    std::lock_guard<std::mutex> lk(api_mutex());
    const Family *fam = family_default_cached((u32)p);
    const Outer *o = (fam && e > kMaxLdsLog) ? outer_cached(fam, e) : nullptr;
    if (!fam || (e > kMaxLdsLog && !o))
        return fail(CUMODP_ASSUMPTION_ERROR,
                    "non-allocating multiply: the twiddle tables of this prime and size are not "
                    "built yet (one cumodp_fftmul_uni_dev call of the same size builds them)");
    const int64_t n = (int64_t)1 << e;
    const int64_t nout = (int64_t)dh + 1 < n ? (int64_t)dh + 1 : n;
    hipStream_t s = as_stream(stream);
    prof_suspend(true);
    hipError_t he;
    if (e <= kMaxLdsLog) {
        MulDesc d;
        d.f = 0; d.g = 0; d.h = 0;
        d.lf = df + 1; d.lg = dg + 1; d.lh = (int)nout; d.logn = e;
        he = launch_mul_lds_one(e, (const u32 *)Fd, (const u32 *)Gd, (u32 *)Hd, nullptr, 1, d,
                                fam, s);
    } else {
        u32 *ws = nullptr;
        he = noalloc_scratch(&ws);
        if (he == hipSuccess) {
            FourStepIO io{(const u32 *)Fd, (const u32 *)Gd, df + 1, dg + 1, 0, ws, ws + n,
                          (u32 *)Hd, nout, 0, 1};
            he = launch_mul_four_step_with(io, fam, o, s);
        }
    }
    prof_suspend(false);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "non-allocating multiply", he);
    return CUMODP_SUCCESS;
}

// ---- two-convolution integer multiply -------------------------------------
static int tc_out_bits(int bits_a, int bits_b, int64_t na, int64_t nb) {
    int lg = 0;
    while (((int64_t)1 << lg) < std::min(na, nb)) ++lg;
    return bits_a + bits_b + lg + 1;
}

// caller holds api_mutex
static cumodp_err tc_dev_impl(const uint64_t *a, int64_t na, int wa, const uint64_t *b,
                              int64_t nb, int wb, uint64_t *c, int wc, int max_bits,
                              hipStream_t s) {
    int bits = max_bits;
    if (bits <= 0) {
        std::vector<u64> ha((size_t)(na * wa)), hb((size_t)(nb * wb));
        hipError_t he = hipMemcpy(ha.data(), a, ha.size() * 8, hipMemcpyDeviceToHost);
        if (he == hipSuccess) he = hipMemcpy(hb.data(), b, hb.size() * 8, hipMemcpyDeviceToHost);
        if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "reading inputs", he);
        bits = std::max(tc_bit_width(ha.data(), na, wa), tc_bit_width(hb.data(), nb, wb));
    }
    if (64 * wc < tc_out_bits(bits, bits, na, nb))
        return fail(CUMODP_ASSUMPTION_ERROR, "output words too narrow for the product");
    int err = 0;
    hipError_t he = launch_two_convolution((const u64 *)a, na, wa, (const u64 *)b, nb, wb,
                                           (u64 *)c, wc, bits, s, &err);
    if (err) return fail(CUMODP_FFT_SIZE_TOO_LARGE, "no two-convolution base fits these sizes");
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "two-convolution multiply", he);
    return CUMODP_SUCCESS;
}

cumodp_err cumodp_two_convolution_poly_mul_dev(const uint64_t *a, int64_t na, int wa,
                                               const uint64_t *b, int64_t nb, int wb, uint64_t *c,
                                               int wc, int max_bits, void *stream) {
    if (na < 1 || nb < 1 || wa < 1 || wb < 1 || wc < 1)
        return fail(CUMODP_ASSUMPTION_ERROR, "empty polynomial or word count");
    std::lock_guard<std::mutex> lk(api_mutex());
    return tc_dev_impl(a, na, wa, b, nb, wb, c, wc, max_bits, as_stream(stream));
}

cumodp_err cumodp_two_convolution_poly_mul(const uint64_t *a, int64_t na, int wa,
                                           const uint64_t *b, int64_t nb, int wb, uint64_t *c,
                                           int wc) {
    if (na < 1 || nb < 1 || wa < 1 || wb < 1 || wc < 1)
        return fail(CUMODP_ASSUMPTION_ERROR, "empty polynomial or word count");
    const int ba = tc_bit_width((const u64 *)a, na, wa), bb = tc_bit_width((const u64 *)b, nb, wb);
    if (64 * wc < tc_out_bits(ba, bb, na, nb))
        return fail(CUMODP_ASSUMPTION_ERROR, "output words too narrow for the product");
    const int64_t nc = na + nb - 1;
    std::lock_guard<std::mutex> lk(api_mutex());
    u64 *d = (u64 *)workspace(12, sizeof(u64) * (size_t)(na * wa + nb * wb + nc * wc), 0);
    if (!d) return fail(CUMODP_OUT_OF_MEMORY, "staging buffer");
    u64 *da = d, *db = da + na * wa, *dc = db + nb * wb;
    hipError_t he = hipMemcpy(da, a, sizeof(u64) * (size_t)(na * wa), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(db, b, sizeof(u64) * (size_t)(nb * wb), hipMemcpyHostToDevice);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy in", he);
    const double t0 = verbose() ? now_s() : 0.0;
    cumodp_err ce = tc_dev_impl(da, na, wa, db, nb, wb, dc, wc, std::max(ba, bb), nullptr);
    if (ce != CUMODP_SUCCESS) return ce;
    he = hipMemcpy(c, dc, sizeof(u64) * (size_t)(nc * wc), hipMemcpyDeviceToHost);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy out", he);
    if (verbose()) printf("\ttwo_convolution_poly_mul took %.3f ms\n", (now_s() - t0) * 1e3);
    return CUMODP_SUCCESS;
}

cumodp_err cumodp_two_convolution_list_dev(const cumodp_zmul_item *items, int64_t count,
                                           const uint64_t *A, const uint64_t *B, uint64_t *C,
                                           void *stream) {
    static_assert(sizeof(cumodp_zmul_item) == sizeof(TcItem), "item layout");
    if (count < 0) return fail(CUMODP_ASSUMPTION_ERROR, "negative count");
    if (count == 0) return CUMODP_SUCCESS;
    for (int64_t i = 0; i < count; ++i) {
        const cumodp_zmul_item &it = items[i];
        if (it.na < 1 || it.nb < 1 || it.wa < 1 || it.wb < 1 || it.wc < 1 || it.bits < 1 ||
            it.a_off < 0 || it.b_off < 0 || it.c_off < 0)
            return fail(CUMODP_ASSUMPTION_ERROR, "bad integer product item");
        if (it.bits > 64 * std::max(it.wa, it.wb))
            return fail(CUMODP_ASSUMPTION_ERROR, "bits exceed the operand words");
        if (64 * it.wc < tc_out_bits(it.bits, it.bits, it.na, it.nb))
            return fail(CUMODP_ASSUMPTION_ERROR, "output words too narrow for the product");
    }
    std::lock_guard<std::mutex> lk(api_mutex());
    int err = 0;
    hipError_t he = launch_two_convolution_list((const TcItem *)items, count, (const u64 *)A,
                                                (const u64 *)B, (u64 *)C, as_stream(stream), &err);
    if (err) return fail(CUMODP_FFT_SIZE_TOO_LARGE, "no two-convolution base fits an item");
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "two-convolution list", he);
    return CUMODP_SUCCESS;
}

void *cumodp_host_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
    return p;
}

cumodp_err cumodp_host_free(void *ptr) {
    hipError_t e = hipHostFree(ptr);
    return e == hipSuccess ? CUMODP_SUCCESS : fail(CUMODP_FAILURE, "host free", e);
}

cumodp_err cumodp_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *stream) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream));
    return e == hipSuccess ? CUMODP_SUCCESS : fail(CUMODP_FAILURE, "copy to device", e);
}

cumodp_err cumodp_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *stream) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream));
    return e == hipSuccess ? CUMODP_SUCCESS : fail(CUMODP_FAILURE, "copy to host", e);
}

int cumodp_find_largest_bit_width(const uint64_t *a, int64_t na, int wa, const uint64_t *b,
                                  int64_t nb, int wb) {
    return std::max(tc_bit_width((const u64 *)a, na, wa), tc_bit_width((const u64 *)b, nb, wb));
}

cumodp_err cumodp_convert_to_modular_bivariate(const uint64_t *a, int64_t na, int wa, int K,
                                               int M, sfixn prime, sfixn *out) {
    if (K < 1 || M < 1 || M > 62 || prime < 2 || wa < 1)
        return fail(CUMODP_ASSUMPTION_ERROR, "bivariate base");
    const u32 P = (u32)prime;
    std::vector<u64> mag((size_t)wa + 1);
    for (int64_t i = 0; i < na; ++i) {
        const u64 *x = a + i * wa;
        const bool neg = (int64_t)x[wa - 1] < 0;
        u64 carry = 1;
        for (int q = 0; q < wa; ++q) {
            u64 v = x[q];
            if (neg) { v = ~v + carry; carry = (carry && v == 0); }
            mag[(size_t)q] = v;
        }
        mag[(size_t)wa] = 0;
        for (int j = 0; j < K; ++j) {
            const int64_t pos = (int64_t)j * M;
            const int64_t wd = pos >> 6;
            const int off = (int)(pos & 63);
            u64 v = 0;
            if (wd < wa) {
                v = mag[(size_t)wd] >> off;
                if (off && wd + 1 <= wa) v |= mag[(size_t)wd + 1] << (64 - off);
            }
            v &= (M >= 64) ? ~0ull : ((1ull << M) - 1);
            const u32 r = (u32)(v % P);
            out[i * K + j] = (sfixn)(neg ? (r ? P - r : 0u) : r);
        }
    }
    return CUMODP_SUCCESS;
}

cumodp_err cumodp_scale_x_argument_dev(sfixn *A, int64_t len, int K, sfixn theta, sfixn prime,
                                       void *stream) {
    if (K < 1 || prime < 2) return fail(CUMODP_ASSUMPTION_ERROR, "scale_x arguments");
    std::lock_guard<std::mutex> lk(api_mutex());
    hipError_t he = launch_scale_x((u32 *)A, len, K, (u32)theta, (u32)prime, as_stream(stream));
    if (he == hipSuccess) he = hipStreamSynchronize(as_stream(stream));
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "scale_x", he);
    return CUMODP_SUCCESS;
}

int cumodp_two_convolution_plan(int bits, int64_t na, int64_t nb, int *nprimes, int *K, int *M) {
    return tc_plan_info(bits, na, nb, nprimes, K, M);
}

// ---- subproduct tree -------------------------------------------------------
int64_t cumodp_subproduct_tree_words(int k) {
    return (k < 0 || k > 30) ? -1 : tree_level_offset(k, k + 1);
}

int64_t cumodp_subproduct_tree_offset(int k, int i) {
    return (k < 0 || k > 30 || i < 0 || i > k) ? -1 : tree_level_offset(k, i);
}

static cumodp_err tree_check(int k, sfixn p) {
    if (k < 1) return fail(CUMODP_ASSUMPTION_ERROR, "tree needs 2^k points, k >= 1");
    return check_prime_size((u32)p, k < 4 ? 4 : k);
}

cumodp_err cumodp_subproduct_tree_dev(const sfixn *points, int k, sfixn p, sfixn *tree,
                                      void *stream) {
    cumodp_err ce = tree_check(k, p);
    if (ce != CUMODP_SUCCESS) return ce;
    std::lock_guard<std::mutex> lk(api_mutex());
    const Family *fam = family_default((u32)p);
    if (!fam) return fail(CUMODP_OUT_OF_MEMORY, "twiddle tables");
    hipError_t he = launch_subproduct_tree((const u32 *)points, k, (u32)p, (u32 *)tree, fam,
                                           as_stream(stream));
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "subproduct tree", he);
    return CUMODP_SUCCESS;
}

static cumodp_err eval_impl(const u32 *F, const u32 *pts, int k, u32 p, u32 *values,
                            hipStream_t s) {
    const Family *fam = family_default(p);
    if (!fam) return fail(CUMODP_OUT_OF_MEMORY, "twiddle tables");
    u32 *scratch = (u32 *)workspace(15, sizeof(u32) * (size_t)eval_scratch_words(k), s);
    if (!scratch) return fail(CUMODP_OUT_OF_MEMORY, "evaluation scratch");
    hipError_t he = launch_fast_evaluation(F, pts, k, p, values, scratch, fam, s);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "fast evaluation", he);
    return CUMODP_SUCCESS;
}

cumodp_err cumodp_fast_evaluation_dev(const sfixn *F, const sfixn *points, int k, sfixn p,
                                      sfixn *values, void *stream) {
    cumodp_err ce = tree_check(k, p);
    if (ce != CUMODP_SUCCESS) return ce;
    std::lock_guard<std::mutex> lk(api_mutex());
    return eval_impl((const u32 *)F, (const u32 *)points, k, (u32)p, (u32 *)values,
                     as_stream(stream));
}

cumodp_err cumodp_fast_evaluation(const sfixn *F, const sfixn *points, int k, sfixn p,
                                  sfixn *values) {
    cumodp_err ce = tree_check(k, p);
    if (ce != CUMODP_SUCCESS) return ce;
    const int64_t n = (int64_t)1 << k;
    std::lock_guard<std::mutex> lk(api_mutex());
    u32 *d = (u32 *)workspace(16, sizeof(u32) * (size_t)(3 * n), 0);
    if (!d) return fail(CUMODP_OUT_OF_MEMORY, "staging buffer");
    hipError_t he = hipMemcpy(d, F, sizeof(u32) * (size_t)n, hipMemcpyHostToDevice);
    if (he == hipSuccess)
        he = hipMemcpy(d + n, points, sizeof(u32) * (size_t)n, hipMemcpyHostToDevice);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy in", he);
    ce = eval_impl(d, d + n, k, (u32)p, d + 2 * n, 0);
    if (ce != CUMODP_SUCCESS) return ce;
    he = hipMemcpy(values, d + 2 * n, sizeof(u32) * (size_t)n, hipMemcpyDeviceToHost);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy out", he);
    return CUMODP_SUCCESS;
}

cumodp_err cumodp_subproduct_tree(const sfixn *points, int k, sfixn p, sfixn *tree) {
    cumodp_err ce = tree_check(k, p);
    if (ce != CUMODP_SUCCESS) return ce;
    const int64_t n = (int64_t)1 << k, words = tree_level_offset(k, k + 1);
    std::lock_guard<std::mutex> lk(api_mutex());
    u32 *d = (u32 *)workspace(13, sizeof(u32) * (size_t)(n + words), 0);
    if (!d) return fail(CUMODP_OUT_OF_MEMORY, "staging buffer");
    const Family *fam = family_default((u32)p);
    if (!fam) return fail(CUMODP_OUT_OF_MEMORY, "twiddle tables");
    hipError_t he = hipMemcpy(d, points, sizeof(u32) * (size_t)n, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = launch_subproduct_tree(d, k, (u32)p, d + n, fam, 0);
    if (he == hipSuccess)
        he = hipMemcpy(tree, d + n, sizeof(u32) * (size_t)words, hipMemcpyDeviceToHost);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "subproduct tree", he);
    return CUMODP_SUCCESS;
}

namespace {
struct CapturedGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void *> host_bufs;   // pinned descriptor copies the replays read
    std::vector<void *> scratch;     // device scratch the replays use (held, see plan.cpp)
};
}  // namespace

cumodp_err cumodp_graph_begin(void *stream) {
    if (!stream) return fail(CUMODP_ASSUMPTION_ERROR, "graphs need a created stream");
    hipError_t e = hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeRelaxed);
    return e == hipSuccess ? CUMODP_SUCCESS : fail(CUMODP_FAILURE, "begin capture", e);
}

void *cumodp_graph_end(void *stream) {
    if (!stream) {
        fail(CUMODP_ASSUMPTION_ERROR, "graphs need a created stream");
        return nullptr;
    }
    hipStream_t s = as_stream(stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    std::vector<void *> bufs = take_captured_buffers(s);
    hipGraphExec_t x = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        fail(CUMODP_FAILURE, "end capture", e);
        if (g) (void)hipGraphDestroy(g);
        for (void *h : bufs) (void)hipHostFree(h);
        return nullptr;
    }
    CapturedGraph *cg = new CapturedGraph();
    cg->graph = g;
    cg->exec = x;
    cg->host_bufs.swap(bufs);
    cg->scratch = hold_workspace_for_graph(s);
    return cg;
}

cumodp_err cumodp_graph_launch(void *graph, void *stream) {
    if (!graph) return fail(CUMODP_ASSUMPTION_ERROR, "null graph");
    hipError_t e = hipGraphLaunch(((CapturedGraph *)graph)->exec, as_stream(stream));
    return e == hipSuccess ? CUMODP_SUCCESS : fail(CUMODP_FAILURE, "graph launch", e);
}

cumodp_err cumodp_graph_destroy(void *graph) {
    if (!graph) return CUMODP_SUCCESS;
    CapturedGraph *cg = (CapturedGraph *)graph;
    // replays still queued may read the graph's descriptor copies
    (void)hipDeviceSynchronize();
    if (cg->exec) (void)hipGraphExecDestroy(cg->exec);
    if (cg->graph) (void)hipGraphDestroy(cg->graph);
    for (void *h : cg->host_bufs) (void)hipHostFree(h);
    release_graph_workspace(cg->scratch);
    delete cg;
    return CUMODP_SUCCESS;
}

int cumodp_tune_get_split(int k) {
    if (k <= kMaxLdsLog || k > kMaxLog) return 0;
    int k1, k2;
    four_step_split(k, k1, k2);
    return k2;
}

cumodp_err cumodp_tune_set_list_tile(int words) {
    std::lock_guard<std::mutex> lk(api_mutex());
    return set_list_tile_words(words) ? CUMODP_SUCCESS
                                      : fail(CUMODP_ASSUMPTION_ERROR, "tile words: 2048, 4096 or 8192");
}

int cumodp_tune_get_list_tile(void) { return list_tile_words(); }

cumodp_err cumodp_tune_set_zmul_merge(int merge) {
    std::lock_guard<std::mutex> lk(api_mutex());
    return set_zmul_merge(merge) ? CUMODP_SUCCESS
                                 : fail(CUMODP_ASSUMPTION_ERROR, "zmul merge: 0 or 1");
}

int cumodp_tune_get_zmul_merge(void) { return zmul_merge(); }

cumodp_err cumodp_tune_set_list_plain(int cutoff) {
    std::lock_guard<std::mutex> lk(api_mutex());
    return set_list_plain_cutoff(cutoff) ? CUMODP_SUCCESS
                                         : fail(CUMODP_ASSUMPTION_ERROR, "plain cutoff: 0 .. 1024");
}

int cumodp_tune_get_list_plain(void) { return list_plain_cutoff(); }

cumodp_err cumodp_tune_set_list_persistent(int on) {
    std::lock_guard<std::mutex> lk(api_mutex());
    return set_list_persistent(on) ? CUMODP_SUCCESS
                                   : fail(CUMODP_ASSUMPTION_ERROR, "list persistent: 0 or 1");
}

int cumodp_tune_get_list_persistent(void) { return list_persistent(); }

cumodp_err cumodp_tune_set_col_tile(int words) {
    std::lock_guard<std::mutex> lk(api_mutex());
    return set_col_tile_words(words) ? CUMODP_SUCCESS
                                     : fail(CUMODP_ASSUMPTION_ERROR, "column tile words: 8192 or 16384");
}

int cumodp_tune_get_col_tile(void) { return col_tile_words(); }

cumodp_err cumodp_tune_set_split(int k, int k2) {
    return set_four_step_split(k, k2) ? CUMODP_SUCCESS
                                      : fail(CUMODP_ASSUMPTION_ERROR, "unsupported four-step split");
}

cumodp_err cumodp_list_fftmul_dev(const cumodp_mul_item *items, int64_t count, const sfixn *F,
                                  const sfixn *G, sfixn *H, sfixn p, void *stream) {
    if (count < 0) return fail(CUMODP_ASSUMPTION_ERROR, "negative count");
    if (count == 0) return CUMODP_SUCCESS;
    // the descriptor array is reused from call to call (this thread's)
    thread_local std::vector<MulDesc> d;
    d.resize((size_t)count);
    int emax = 0;
    for (int64_t i = 0; i < count; ++i) {
        const cumodp_mul_item &it = items[i];
        if (it.df < 0 || it.dg < 0) return fail(CUMODP_ASSUMPTION_ERROR, "negative degree");
        int e = ceil_log2((int64_t)it.df + it.dg + 1);
        if (e < 4) e = 4;
        emax = e > emax ? e : emax;
        d[(size_t)i].f = it.f_off; d[(size_t)i].g = it.g_off; d[(size_t)i].h = it.h_off;
        d[(size_t)i].lf = it.df + 1; d[(size_t)i].lg = it.dg + 1;
        d[(size_t)i].lh = it.df + it.dg + 1; d[(size_t)i].logn = e;
    }
    cumodp_err ce = check_prime_size((u32)p, emax);
    if (ce != CUMODP_SUCCESS) return ce;
    std::lock_guard<std::mutex> lk(api_mutex());
    const Family *fam = family_default((u32)p);
    if (!fam) return fail(CUMODP_OUT_OF_MEMORY, "twiddle tables");
    hipError_t he = launch_list_mul(d.data(), count, (const u32 *)F, (const u32 *)G, (u32 *)H, fam,
                                    as_stream(stream));
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "list multiply", he);
    return CUMODP_SUCCESS;
}

cumodp_err cumodp_list_plainmul_dev(const cumodp_mul_item *items, int64_t count, const sfixn *F,
                                    const sfixn *G, sfixn *H, sfixn p, void *stream) {
    if (count < 0) return fail(CUMODP_ASSUMPTION_ERROR, "negative count");
    if (count == 0) return CUMODP_SUCCESS;
    if (count > INT32_MAX) return fail(CUMODP_ASSUMPTION_ERROR, "too many items");
    std::vector<MulDesc> d((size_t)count);
    for (int64_t i = 0; i < count; ++i) {
        const cumodp_mul_item &it = items[i];
        if (it.df < 0 || it.dg < 0) return fail(CUMODP_ASSUMPTION_ERROR, "negative degree");
        d[(size_t)i].f = it.f_off; d[(size_t)i].g = it.g_off; d[(size_t)i].h = it.h_off;
        d[(size_t)i].lf = it.df + 1; d[(size_t)i].lg = it.dg + 1;
    }
    std::lock_guard<std::mutex> lk(api_mutex());
    return plain_dev_impl(d, (const u32 *)F, (const u32 *)G, (u32 *)H, (u32)p, as_stream(stream));
}

cumodp_err cumodp_plain_mul_dev(sfixn *C, const sfixn *A, sfixn n, const sfixn *B, sfixn m,
                                sfixn p, void *stream) {
    if (n < 1 || m < 1) return fail(CUMODP_ASSUMPTION_ERROR, "empty operand");
    std::vector<MulDesc> d(1);
    d[0].f = 0; d[0].g = 0; d[0].h = 0; d[0].lf = n; d[0].lg = m;
    std::lock_guard<std::mutex> lk(api_mutex());
    return plain_dev_impl(d, (const u32 *)A, (const u32 *)B, (u32 *)C, (u32)p, as_stream(stream));
}

// One cumodp_fftmul_uni_dev-equivalent call per item (the plain per-call form
// a caller would write as a loop), for comparison with the batched engine.
cumodp_err cumodp_list_fftmul_dev_percall(const cumodp_mul_item *items, int64_t count,
                                          const sfixn *F, const sfixn *G, sfixn *H, sfixn p,
                                          void *stream) {
    for (int64_t i = 0; i < count; ++i) {
        const cumodp_mul_item &it = items[i];
        cumodp_err ce = cumodp_fftmul_uni_dev(it.df + it.dg, H + it.h_off, it.df, F + it.f_off,
                                              it.dg, G + it.g_off, p, stream);
        if (ce != CUMODP_SUCCESS) return ce;
    }
    return CUMODP_SUCCESS;
}

}  // extern "C"

// ===========================================================================
// C++ linkage drop-ins
// ===========================================================================

// ---- the reference's evaluation drivers (include/fastPolyEvaluation.h) ----
// One fast-evaluation run (tree, subinverse tree, remainder tree) in the
// library's layout, then device copies into separately allocated arrays in
// the reference's half-tree layout (src/fastPolyEvaluation.cu:512-656,
// :1371-1470).
namespace {

enum class StepsWanted { kTree, kTrees, kEval };

// a new allocation holding `rows` rows of `width` words read `pitch` words
// apart from src (the caller of the drop-in frees it)
hipError_t steps_copy(sfixn **dst, const u32 *src, int64_t rows, int64_t width, int64_t pitch) {
    void *d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(u32) * (size_t)(rows * width));
    if (e != hipSuccess) return e;
    *dst = (sfixn *)d;
    if (pitch == width || rows == 1)
        return hipMemcpyAsync(d, src, sizeof(u32) * (size_t)(rows * width),
                              hipMemcpyDeviceToDevice, 0);
    return hipMemcpy2DAsync(d, sizeof(u32) * (size_t)width, src, sizeof(u32) * (size_t)pitch,
                            sizeof(u32) * (size_t)width, (size_t)rows, hipMemcpyDeviceToDevice,
                            0);
}

void steps_free(PolyEvalSteps &r) {
    for (int i = 0; i < MAX_LEVEL; ++i) {
        for (sfixn *q : {r.Ml[i], r.Mr[i], r.InvMl[i], r.InvMr[i], r.fL[i], r.fR[i]})
            if (q) (void)hipFree(q);
    }
    memset(&r, 0, sizeof(r));
}

cumodp_err steps_run(PolyEvalSteps &r, const sfixn *F, const sfixn *points, int k, sfixn p,
                     StepsWanted what, bool keep_all) {
    if (k < 1 || k >= MAX_LEVEL) return fail(CUMODP_ASSUMPTION_ERROR, "k must be in 1 .. 26");
    cumodp_err ce = check_prime_size((u32)p, k < 4 ? 4 : k);
    if (ce != CUMODP_SUCCESS) return ce;
    std::lock_guard<std::mutex> lk(api_mutex());
    const Family *fam = family_default((u32)p);
    if (!fam) return fail(CUMODP_OUT_OF_MEMORY, "twiddle tables");
    const int64_t n = (int64_t)1 << k, nh = n / 2;
    const bool eval = what == StepsWanted::kEval;
    u32 *io = (u32 *)workspace(19, sizeof(u32) * (size_t)(3 * n), 0);
    u32 *scr = (u32 *)workspace(15, sizeof(u32) * (size_t)eval_scratch_words(k), 0);
    u32 *kept = keep_all ? (u32 *)workspace(18, sizeof(u32) * (size_t)(k * n), 0) : nullptr;
    if (!io || !scr || (keep_all && !kept)) return fail(CUMODP_OUT_OF_MEMORY, "scratch");
    u32 *dpts = io, *dF = io + n, *dval = io + 2 * n;
    hipError_t he = hipMemcpy(dpts, points, sizeof(u32) * (size_t)n, hipMemcpyHostToDevice);
    if (he == hipSuccess && eval)
        he = hipMemcpy(dF, F, sizeof(u32) * (size_t)n, hipMemcpyHostToDevice);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "copy in", he);
    EvalKeep keep;
    keep.trees_only = !eval;
    keep.rem = kept;
    he = launch_fast_evaluation(dF, dpts, k, (u32)p, dval, scr, fam, 0, &keep);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "fast evaluation", he);
    const bool trees = !eval || keep_all;
    if (trees) {
        he = hipMalloc((void **)&r.Ml[0], sizeof(u32) * (size_t)(2 * nh));
        if (he == hipSuccess) he = hipMalloc((void **)&r.Mr[0], sizeof(u32) * (size_t)(2 * nh));
        if (he == hipSuccess)
            he = launch_ref_leaves(dpts, nh, (u32 *)r.Ml[0], (u32 *)r.Mr[0], 0);
        for (int i = 1; i < k && he == hipSuccess; ++i) {
            const int64_t nodes = nh >> i, len = ((int64_t)1 << i) + 1;
            const u32 *lv = scr + tree_level_offset(k, i);
            if (i < plainMulLimit) {
                he = steps_copy(&r.Ml[i], lv, 1, nodes * len, nodes * len);
                if (he == hipSuccess)
                    he = steps_copy(&r.Mr[i], lv + nodes * len, 1, nodes * len, nodes * len);
            } else {
                he = steps_copy(&r.Ml[i], lv, nodes, len - 1, len);
                if (he == hipSuccess) he = steps_copy(&r.Mr[i], lv + nodes * len, nodes, len - 1, len);
            }
        }
        // subinverse levels: plainMulLimit only for the tree alone, else
        // plainMulLimit .. k-1 (the library keeps levels min(k,4) .. k-1)
        const int top = what == StepsWanted::kTree ? plainMulLimit : k - 1;
        const int L0 = k < 4 ? k : 4;
        for (int i = plainMulLimit; i <= top && i < k && he == hipSuccess; ++i) {
            const u32 *iv = scr + eval_inverse_offset(k) + (int64_t)(i - L0) * n;
            he = steps_copy(&r.InvMl[i], iv, 1, nh, nh);
            if (he == hipSuccess) he = steps_copy(&r.InvMr[i], iv + nh, 1, nh, nh);
        }
    }
    if (eval && he == hipSuccess) {
        he = steps_copy(&r.fL[1], dval, 1, nh, nh);
        if (he == hipSuccess) he = steps_copy(&r.fR[1], dval + nh, 1, nh, nh);
        for (int j = 2; keep_all && j <= k && he == hipSuccess; ++j) {
            const u32 *rv = kept + (int64_t)(j - 1) * n;
            he = steps_copy(&r.fL[j], rv, 1, nh, nh);
            if (he == hipSuccess) he = steps_copy(&r.fR[j], rv + nh, 1, nh, nh);
        }
    }
    if (he == hipSuccess) he = hipStreamSynchronize(0);
    if (he != hipSuccess) return fail(CUMODP_FFT_ERROR, "evaluation steps", he);
    return CUMODP_SUCCESS;
}

PolyEvalSteps steps_entry(const char *fn, const sfixn *F, const sfixn *points, int k, sfixn p,
                          StepsWanted what, bool keep_all) {
    PolyEvalSteps r;
    memset(&r, 0, sizeof(r));
    cumodp_err e = steps_run(r, F, points, k, p, what, keep_all);
    if (e != CUMODP_SUCCESS) {
        steps_free(r);
        report(fn, e);
    }
    return r;
}

}  // namespace

struct PolyEvalSteps fastEvaluation(sfixn *F, sfixn *points, sfixn k, sfixn p, sfixn flag) {
    return steps_entry("fastEvaluation", F, points, k, p, StepsWanted::kEval, flag == 1);
}

struct PolyEvalSteps onlySubproductTree(sfixn *, sfixn *points, sfixn k, sfixn p, sfixn) {
    return steps_entry("onlySubproductTree", nullptr, points, k, p, StepsWanted::kTree, false);
}

struct PolyEvalSteps onlyTrees(sfixn *points, sfixn k, sfixn p, sfixn) {
    return steps_entry("onlyTrees", nullptr, points, k, p, StepsWanted::kTrees, false);
}

void stockham_dev(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p) {
    (void)n;
    report("stockham_dev", ntt_dev_impl((u32 *)X, 1, k, (u32)w, (u32)p, false, 0));
}

void stockham_dev(sfixn *X, sfixn n, sfixn k, const sfixn *W, sfixn p) {
    stockham_dev(X, n, k, (sfixn)read_root(W), p);
}

void inverse_stockham_dev(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p) {
    (void)n;
    report("inverse_stockham_dev", ntt_dev_impl((u32 *)X, 1, k, (u32)w, (u32)p, true, 0));
}

void inverse_stockham_dev(sfixn *X, sfixn n, sfixn k, const sfixn *W, sfixn p) {
    // W holds the powers of the inverse root: DFT with W[1], then scaling
    u32 wi = read_root(W);
    inverse_stockham_dev(X, n, k, (sfixn)cm_host_inv(wi % (u32)p, (u32)p), p);
}

void stockham_host(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p) {
    (void)n;
    report("stockham_host", ntt_host_impl(X, 1, k, (u32)w, (u32)p, false));
}

void inverse_stockham_host(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p) {
    (void)n;
    report("inverse_stockham_host", ntt_host_impl(X, 1, k, (u32)w, (u32)p, true));
}

void stockham_poly_mul_dev(sfixn n, sfixn e, sfixn *F, sfixn *G, sfixn p) {
    std::lock_guard<std::mutex> lk(api_mutex());
    cumodp_err ce = check_prime_size((u32)p, e);
    if (ce == CUMODP_SUCCESS && e < 4) ce = fail(CUMODP_FFT_SIZE_TOO_SMALL, "transform size below 16");
    if (ce == CUMODP_SUCCESS) {
        // cyclic product of length n = 2^e, as the reference computes it
        const Family *fam = family_default((u32)p);
        hipError_t he = hipErrorOutOfMemory;
        if (fam && e <= kMaxLdsLog) {
            MulDesc d;
            d.f = 0; d.g = 0; d.h = 0;
            d.lf = d.lg = d.lh = (int)n; d.logn = e;
            he = launch_mul_lds_one(e, (const u32 *)F, (const u32 *)G, (u32 *)F, nullptr, 1, d,
                                    fam, 0);
        } else if (fam) {
            FourStepIO io{(const u32 *)F, (const u32 *)G, n, n, 0, (u32 *)F, (u32 *)G, (u32 *)F,
                          n, 0, 1};
            he = launch_mul_four_step(e, io, fam, 0);
        }
        if (he == hipSuccess) he = hipStreamSynchronize(0);
        if (he != hipSuccess) ce = fail(CUMODP_FFT_ERROR, "multiply", he);
    }
    report("stockham_poly_mul_dev", ce);
}

void stockham_poly_mul_host(sfixn dh, sfixn *H, sfixn df, const sfixn *F, sfixn dg,
                            const sfixn *G, sfixn p) {
    const double t0 = verbose() ? now_s() : 0.0;
    const cumodp_err e = mul_host_impl(dh, H, df, F, dg, G, p);
    report("stockham_poly_mul_host", e);
    if (verbose() && e == CUMODP_SUCCESS) printf("FFT-based: %g\n", now_s() - t0);
}

void mulCUDA(sfixn *A, sfixn *B, sfixn *C, sfixn n, sfixn m, sfixn Prime, sfixn verify) {
    const double t0 = verbose() ? now_s() : 0.0;
    const cumodp_err e = plain_host_impl(A, n, B, m, C, Prime);
    report("mulCUDA", e);
    // the reference prints "n m seconds" unless verify == 1
    if (verbose() && verify != 1 && e == CUMODP_SUCCESS)
        printf("%d %d %g\n", (int)n, (int)m, now_s() - t0);
}

void list_stockham_host(sfixn *X, sfixn m, sfixn k, sfixn w, sfixn p) {
    report("list_stockham_host", ntt_host_impl(X, m, k, (u32)w, (u32)p, false));
}

void list_stockham_dev(sfixn *X, sfixn m, sfixn k, sfixn w, sfixn p) {
    report("list_stockham_dev", ntt_dev_impl((u32 *)X, m, k, (u32)w, (u32)p, false, 0));
}

void list_stockham_dev(sfixn *X, sfixn m, sfixn k, const sfixn *W, sfixn p) {
    list_stockham_dev(X, m, k, (sfixn)read_root(W), p);
}

void list_inv_stockham_host(sfixn *X, sfixn m, sfixn k, sfixn w, sfixn p) {
    // src/list_inv_stockham.cu:71-88: forward with w, forward with w^-1,
    // scaling by n^-1 (net: identity for a primitive root w)
    cumodp_err ce = ntt_host_impl(X, m, k, (u32)w, (u32)p, false);
    if (ce == CUMODP_SUCCESS) ce = ntt_host_impl(X, m, k, cm_host_inv((u32)w % (u32)p, (u32)p), (u32)p, false);
    if (ce == CUMODP_SUCCESS) {
        u32 ni = cm_host_inv(cm_host_pow(2, k, (u32)p), (u32)p);
        for (int64_t i = 0; i < ((int64_t)m << k); ++i) X[i] = (sfixn)cm_host_mul((u32)X[i], ni, (u32)p);
    }
    report("list_inv_stockham_host", ce);
}

void get_powers_binary(sfixn e, sfixn *W, sfixn w, sfixn p) {
    hipError_t he = launch_powers((u32 *)W, (int64_t)1 << e, (u32)w, (u32)p, 0);
    if (he == hipSuccess) he = hipStreamSynchronize(0);
    if (he != hipSuccess) report("get_powers_binary", fail(CUMODP_FFT_ERROR, "powers", he));
}

void pointwise_mul_dev(sfixn n, sfixn k, sfixn *Y, const sfixn *X, sfixn p) {
    (void)k;
    hipError_t he = launch_pointwise((u32 *)Y, (const u32 *)X, n, (u32)p, 0);
    if (he == hipSuccess) he = hipStreamSynchronize(0);
    if (he != hipSuccess) report("pointwise_mul_dev", fail(CUMODP_FFT_ERROR, "pointwise", he));
}

void scale_vector_dev(sfixn a, sfixn n, sfixn *X, sfixn p) {
    hipError_t he = launch_scale((u32 *)X, n, (u32)a, (u32)p, 0);
    if (he == hipSuccess) he = hipStreamSynchronize(0);
    if (he != hipSuccess) report("scale_vector_dev", fail(CUMODP_FFT_ERROR, "scale", he));
}

void expand_to_fft_dev(sfixn m, sfixn n, sfixn *X) {
    hipError_t he = launch_fill((u32 *)X + n, (int64_t)m - n, 0u, 0);
    if (he == hipSuccess) he = hipStreamSynchronize(0);
    if (he != hipSuccess) report("expand_to_fft_dev", fail(CUMODP_FFT_ERROR, "fill", he));
}

void reset_vector_dev(sfixn n, sfixn *X) {
    hipError_t he = launch_fill((u32 *)X, n, 0u, 0);
    if (he == hipSuccess) he = hipStreamSynchronize(0);
    if (he != hipSuccess) report("reset_vector_dev", fail(CUMODP_FFT_ERROR, "fill", he));
}

void bivariate_stockham_dev(sfixn *X, sfixn em, const sfixn *Wm, sfixn en, const sfixn *Wn,
                            sfixn p) {
    report("bivariate_stockham_dev",
           biv_dev_impl((u32 *)X, em, read_root_k(Wm, em, (u32)p), en,
                        read_root_k(Wn, en, (u32)p), (u32)p, false, 0));
}

void inverse_bivariate_stockham_dev(sfixn *X, sfixn em, const sfixn *Wm, sfixn en,
                                    const sfixn *Wn, sfixn p) {
    // src/list_stockham.cu:1251-1263: the tables already hold inverse roots
    const u32 P = (u32)p;
    report("inverse_bivariate_stockham_dev",
           biv_dev_impl((u32 *)X, em, cm_host_inv(read_root_k(Wm, em, P) % P, P), en,
                        cm_host_inv(read_root_k(Wn, en, P) % P, P), P, true, 0));
}

void bivariate_stockham_host(sfixn *X, sfixn em, sfixn wm, sfixn en, sfixn wn, sfixn p) {
    report("bivariate_stockham_host", biv_host_impl(X, em, (u32)wm, en, (u32)wn, (u32)p, false));
}

void inverse_bivariate_stockham_host(sfixn *X, sfixn em, sfixn wm, sfixn en, sfixn wn, sfixn p) {
    report("inverse_bivariate_stockham_host",
           biv_host_impl(X, em, (u32)wm, en, (u32)wn, (u32)p, true));
}

void list_bivariate_stockham_dev(sfixn *X, sfixn q, sfixn em, sfixn wm, sfixn en, sfixn wn,
                                 sfixn p) {
    std::lock_guard<std::mutex> lk(api_mutex());
    report("list_bivariate_stockham_dev",
           biv_dev_impl((u32 *)X, em, (u32)wm, en, (u32)wn, (u32)p, false, 0, q));
}

void list_bivariate_stockham_dev(sfixn *X, sfixn q, sfixn em, const sfixn *Wm, sfixn en,
                                 const sfixn *Wn, sfixn p) {
    // the tables are [1, w, ..., w^(m/2-1)] (src/list_stockham.cu:1849-1862)
    list_bivariate_stockham_dev(X, q, em, (sfixn)read_root_k(Wm, em, (u32)p), en,
                                (sfixn)read_root_k(Wn, en, (u32)p), p);
}

void bi_stockham_poly_mul_dev(sfixn m, sfixn em, sfixn n, sfixn en, sfixn *F, sfixn *G,
                              sfixn p) {
    // src/list_stockham.cu:1317-1348: F <- DFT^-1(DFT(F) DFT(G)), G overwritten
    std::lock_guard<std::mutex> lk(api_mutex());
    const u32 P = (u32)p;
    const u32 wm = default_root(P, em), wn = default_root(P, en);
    cumodp_err ce = (wm && wn) ? CUMODP_SUCCESS
                               : fail(CUMODP_FOURIER_DEGREE_TOO_SMALL, "2^k does not divide p-1");
    if (ce == CUMODP_SUCCESS) ce = biv_dev_impl((u32 *)F, em, wm, en, wn, P, false, 0);
    if (ce == CUMODP_SUCCESS) ce = biv_dev_impl((u32 *)G, em, wm, en, wn, P, false, 0);
    if (ce == CUMODP_SUCCESS) {
        hipError_t he = launch_pointwise((u32 *)F, (const u32 *)G, (int64_t)m * n, P, 0);
        if (he != hipSuccess) ce = fail(CUMODP_FFT_ERROR, "pointwise", he);
    }
    if (ce == CUMODP_SUCCESS) ce = biv_dev_impl((u32 *)F, em, wm, en, wn, P, true, 0);
    report("bi_stockham_poly_mul_dev", ce);
}
