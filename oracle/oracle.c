/*
 * oracle.c -- CPU restatement of CUMODP's NTT polynomial-multiplication path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): imported by tests/, smoke() and
 * bench.py's cpu_baseline leg; never linked into the product.
 *
 * Written from the mathematics; each function cites the reference code whose
 * behaviour it restates.  All residues are canonical, in [0, p).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef __int128 i128;

/* ------------------------------------------------------------------------ */
/* canonical arithmetic (include/inlines.h:70-122)                          */
/* ------------------------------------------------------------------------ */
sfixn orc_add_mod(sfixn a, sfixn b, sfixn p)
{
    int64_t r = (int64_t)a + b;
    if (r >= p) r -= p;
    return (sfixn)r;
}

sfixn orc_sub_mod(sfixn a, sfixn b, sfixn p)
{
    int64_t r = (int64_t)a - b;
    if (r < 0) r += p;
    return (sfixn)r;
}

sfixn orc_neg_mod(sfixn a, sfixn p) { return a == 0 ? 0 : p - a; }

sfixn orc_mul_mod(sfixn a, sfixn b, sfixn p)
{
    /* the reference estimates the quotient in double precision and corrects
     * twice; the canonical result is the exact remainder */
    return (sfixn)(((uint64_t)(uint32_t)a * (uint32_t)b) % (uint32_t)p);
}

sfixn orc_inv_mod(sfixn a, sfixn p)
{
    /* extended Euclid on (p, a), include/inlines.h:135-167 */
    int64_t r0 = p, r1 = ((int64_t)a % p + p) % p;
    int64_t t0 = 0, t1 = 1;
    while (r1 != 0) {
        int64_t q = r0 / r1, tmp;
        tmp = r0 - q * r1; r0 = r1; r1 = tmp;
        tmp = t0 - q * t1; t0 = t1; t1 = tmp;
    }
    /* r0 == 1 when a is invertible */
    t0 %= p;
    if (t0 < 0) t0 += p;
    return (sfixn)t0;
}

sfixn orc_pow_mod(sfixn a, sfixn ee, sfixn p)
{
    /* include/inlines.h:181-200: e == 0 -> 1, negative e -> inverse */
    if (ee == 0) return 1;
    uint32_t e = ee < 0 ? (uint32_t)(-(int64_t)ee) : (uint32_t)ee;
    uint64_t x = 1, y = (uint32_t)a % (uint32_t)p;
    while (e) {
        if (e & 1) x = x * y % (uint32_t)p;
        y = y * y % (uint32_t)p;
        e >>= 1;
    }
    if (ee < 0) return orc_inv_mod((sfixn)x, p);
    return (sfixn)x;
}

sfixn orc_fourier_degree(sfixn p)
{
    /* include/inlines.h:408-413: largest e with 2^e | p-1 */
    uint32_t q = (uint32_t)p - 1;
    sfixn e = 0;
    while (q && !(q & 1)) { q >>= 1; ++e; }
    return e;
}

sfixn orc_primitive_root(sfixn e, sfixn p)
{
    /* include/inlines.h:460-485.  The reference reseeds libc's generator
     * with srand(1) and draws r = rand() % p until r >= 2, raises r to
     * (p-1)/2^e and accepts when the result has exact order 2^e.  The same
     * libc calls are made here so the root is the reference's root. */
    sfixn k = (sfixn)(((uint32_t)p - 1) >> e);
    sfixn half = (sfixn)1 << (e - 1);
    sfixn root = 0;
    srand(1);
    while (p > 2) {
        root = 1;
        while (root < 2) root = rand() % p;
        root = orc_pow_mod(root, k, p);
        if (orc_pow_mod(root, half, p) != 1) break;
    }
    return root;
}

/* ------------------------------------------------------------------------ */
/* transforms                                                               */
/* ------------------------------------------------------------------------ */
void orc_get_powers(sfixn n, sfixn *W, sfixn w, sfixn p)
{
    /* src/fft_aux.cu:49-59 (host part) and :70-132 (doubling on device) both
     * produce W[j] = w^j */
    uint64_t x = 1;
    for (sfixn j = 0; j < n; ++j) {
        W[j] = (sfixn)x;
        x = x * (uint32_t)w % (uint32_t)p;
    }
}

void orc_dft_naive(sfixn n, const sfixn *X, sfixn *Y, sfixn w, sfixn p)
{
    uint64_t P = (uint32_t)p;
    for (sfixn j = 0; j < n; ++j) {
        uint64_t wj = (uint64_t)orc_pow_mod(w, j, p), acc = 0, t = 1;
        for (sfixn i = 0; i < n; ++i) {
            acc = (acc + (uint64_t)(uint32_t)X[i] * t) % P;
            t = t * wj % P;
        }
        Y[j] = (sfixn)acc;
    }
}

static uint32_t bitrev(uint32_t x, int bits)
{
    uint32_t r = 0;
    for (int i = 0; i < bits; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

void orc_gsnn_fft(sfixn n, sfixn e, sfixn *A, sfixn p, sfixn wn)
{
    /* include/inlines.h:520-542: decimation in frequency, problem size
     * halves each stage, twiddle w^j within a problem, root squares */
    sfixn problems = 1, size = n;
    sfixn wk = wn;
    while (size > 1) {
        sfixn half = size >> 1;
        for (sfixn k = 0; k < problems; ++k) {
            sfixn first = k * size;
            sfixn w = 1;
            for (sfixn j = first; j < first + half; ++j) {
                sfixn t = A[j];
                A[j] = orc_add_mod(t, A[j + half], p);
                A[j + half] = orc_mul_mod(w, orc_sub_mod(t, A[j + half], p), p);
                w = orc_mul_mod(w, wk, p);
            }
        }
        problems <<= 1;
        size = half;
        wk = orc_mul_mod(wk, wk, p);
    }
    /* include/inlines.h:549-558: bit-reversal to natural order */
    for (sfixn i = 0; i < n; ++i) {
        sfixn t = (sfixn)bitrev((uint32_t)i, e);
        if (t > i) { sfixn v = A[t]; A[t] = A[i]; A[i] = v; }
    }
}

/* DFT_2 (x) I_{n/2}: Y[j] = X[j] + X[j+n/2], Y[j+n/2] = X[j] - X[j+n/2]
 * (src/stockham.cu:369-379) */
static void st_butterfly(sfixn *Y, const sfixn *X, sfixn n, sfixn p)
{
    sfixn h = n >> 1;
    for (sfixn j = 0; j < h; ++j) {
        Y[j] = orc_add_mod(X[j], X[j + h], p);
        Y[j + h] = orc_sub_mod(X[j], X[j + h], p);
    }
}

/* (L_2^{2^(k-i)} (x) I_{2^i}): input block q (2^i words) goes to output
 * block (q mod 2) * 2^(k-i-1) + floor(q/2)  (src/stockham.cu:75-161) */
static void st_transpose(sfixn *Y, const sfixn *X, sfixn k, sfixn i)
{
    sfixn s = (sfixn)1 << i, nblk = (sfixn)1 << (k - i);
    for (sfixn q = 0; q < nblk; ++q) {
        sfixn dst = ((q & 1) * (nblk >> 1) + (q >> 1)) * s;
        memcpy(Y + dst, X + q * s, sizeof(sfixn) * (size_t)s);
    }
}

/* (D_{2,2^(k-i-1)} (x) I_{2^i}): word n/2 + j is multiplied by
 * W[(j >> i) << i]  (src/stockham.cu:201-310) */
static void st_twiddle(sfixn *X, const sfixn *W, sfixn k, sfixn i, sfixn p)
{
    sfixn h = (sfixn)1 << (k - 1);
    for (sfixn j = 0; j < h; ++j)
        X[h + j] = orc_mul_mod(X[h + j], W[(j >> i) << i], p);
}

void orc_stockham(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p)
{
    /* driver sequence of src/stockham.cu:413-434 */
    if (n == 1) return;
    sfixn *W = (sfixn *)malloc(sizeof(sfixn) * (size_t)(n / 2));
    sfixn *Y = (sfixn *)malloc(sizeof(sfixn) * (size_t)n);
    orc_get_powers(n / 2, W, w, p);
    st_butterfly(Y, X, n, p);
    for (sfixn i = k - 2; i >= 0; --i) {
        st_transpose(X, Y, k, i);
        st_twiddle(X, W, k, i, p);
        st_butterfly(Y, X, n, p);
    }
    memcpy(X, Y, sizeof(sfixn) * (size_t)n);
    free(W);
    free(Y);
}

void orc_inverse_stockham(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p)
{
    orc_stockham(X, n, k, orc_inv_mod(w, p), p);
    sfixn ninv = orc_inv_mod(n, p);
    for (sfixn i = 0; i < n; ++i) X[i] = orc_mul_mod(X[i], ninv, p);
}

void orc_list_stockham(sfixn *X, sfixn m, sfixn k, sfixn w, sfixn p)
{
    sfixn n = (sfixn)1 << k;
    for (sfixn r = 0; r < m; ++r) orc_stockham(X + (size_t)r * n, n, k, w, p);
}

/* ------------------------------------------------------------------------ */
/* multiplication mod p                                                     */
/* ------------------------------------------------------------------------ */
void orc_plain_polymul(sfixn dH, sfixn *H, sfixn dF, const sfixn *F,
                       sfixn dG, const sfixn *G, sfixn p)
{
    (void)dH;
    uint64_t P = (uint32_t)p;
    for (sfixn i = 0; i <= dF; ++i) {
        uint64_t f = (uint32_t)F[i];
        if (!f) continue;
        for (sfixn j = 0; j <= dG; ++j)
            H[i + j] = (sfixn)(((uint64_t)(uint32_t)H[i + j] + f * (uint32_t)G[j]) % P);
    }
}

static sfixn ceil_log2(int64_t s)
{
    sfixn e = 0;
    while (((int64_t)1 << e) < s) ++e;
    return e;
}

void orc_ntt_polymul(sfixn *H, sfixn dF, const sfixn *F, sfixn dG,
                     const sfixn *G, sfixn p)
{
    /* src/fft_aux.cu:862-879: forward both, pointwise, inverse with w^-1,
     * scale by n^-1 */
    sfixn s = dF + dG + 1, e = ceil_log2(s);
    sfixn n = (sfixn)1 << e;
    sfixn *F0 = (sfixn *)calloc((size_t)n, sizeof(sfixn));
    sfixn *G0 = (sfixn *)calloc((size_t)n, sizeof(sfixn));
    memcpy(F0, F, sizeof(sfixn) * (size_t)(dF + 1));
    memcpy(G0, G, sizeof(sfixn) * (size_t)(dG + 1));
    if (n == 1) {
        H[0] = orc_mul_mod(F0[0], G0[0], p);
    } else {
        sfixn w = orc_primitive_root(e, p);
        sfixn winv = orc_inv_mod(w, p), ninv = orc_inv_mod(n, p);
        orc_gsnn_fft(n, e, F0, p, w);
        orc_gsnn_fft(n, e, G0, p, w);
        for (sfixn i = 0; i < n; ++i) F0[i] = orc_mul_mod(F0[i], G0[i], p);
        orc_gsnn_fft(n, e, F0, p, winv);
        for (sfixn i = 0; i < s; ++i) H[i] = orc_mul_mod(F0[i], ninv, p);
    }
    free(F0);
    free(G0);
}

void orc_polymul_items(int64_t count, const int64_t *desc, const sfixn *F, const sfixn *G,
                       sfixn *H, sfixn p, int plain_cut, int threads)
{
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int64_t r = 0; r < count; ++r) {
        const int64_t *d = desc + 5 * r;
        sfixn df = (sfixn)d[3], dg = (sfixn)d[4];
        sfixn *h = H + d[2];
        if ((df < dg ? df : dg) + 1 <= plain_cut) {
            memset(h, 0, sizeof(sfixn) * (size_t)(df + dg + 1));
            orc_plain_polymul(df + dg, h, df, F + d[0], dg, G + d[1], p);
        } else {
            orc_ntt_polymul(h, df, F + d[0], dg, G + d[1], p);
        }
    }
}

int orc_stockham_poly_mul_host(sfixn dh, sfixn *H, sfixn df, const sfixn *F,
                               sfixn dg, const sfixn *G, sfixn p)
{
    /* src/stockham.cu:562-611: s = df+dg+1, n = 2^ceil(log2 s); n < 256 is
     * refused (stderr print, H unwritten); dh >= s copies the whole n-word
     * buffer (which overruns H when dh+1 < n), otherwise dh+1 words.  The
     * documented contract kept here: min(dh+1, n) words, the product
     * followed by the zeros the NTT leaves above degree s-1. */
    sfixn s = df + dg + 1, e = ceil_log2(s);
    int64_t n = (int64_t)1 << e;
    if (n < 256) return 1;
    sfixn *T = (sfixn *)malloc(sizeof(sfixn) * (size_t)s);
    orc_ntt_polymul(T, df, F, dg, G, p);
    int64_t nw = (int64_t)dh + 1 < n ? (int64_t)dh + 1 : n;
    for (int64_t i = 0; i < nw; ++i) H[i] = i < s ? T[i] : 0;
    free(T);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* multiword two's-complement helpers                                       */
/* ------------------------------------------------------------------------ */
static int mw_is_neg(const uint64_t *x, int w) { return (int)(x[w - 1] >> 63); }

static void mw_negate(uint64_t *x, int w)
{
    uint64_t carry = 1;
    for (int i = 0; i < w; ++i) {
        uint64_t v = ~x[i] + carry;
        carry = (carry && v == 0);
        x[i] = v;
    }
}

/* dst (w words) += src (ws words, unsigned magnitude) << (64*off)  or -= */
static void mw_addsub_mag(uint64_t *dst, int w, const uint64_t *src, int ws,
                          int off, int subtract)
{
    if (!subtract) {
        u128 carry = 0;
        for (int i = off; i < w; ++i) {
            int si = i - off;
            u128 t = (u128)dst[i] + (si < ws ? src[si] : 0) + carry;
            dst[i] = (uint64_t)t;
            carry = t >> 64;
        }
    } else {
        uint64_t borrow = 0;
        for (int i = off; i < w; ++i) {
            int si = i - off;
            uint64_t s = si < ws ? src[si] : 0;
            uint64_t d = dst[i];
            uint64_t r = d - s - borrow;
            borrow = (d < s) || (d - s < borrow);
            dst[i] = r;
        }
    }
}

/* magnitude of a w-word two's-complement value into mag (w words) */
static int mw_abs(const uint64_t *x, int w, uint64_t *mag)
{
    memcpy(mag, x, sizeof(uint64_t) * (size_t)w);
    int neg = mw_is_neg(x, w);
    if (neg) mw_negate(mag, w);
    return neg;
}

void orc_zmul_mw_schoolbook(const uint64_t *a, int64_t na, int wa,
                            const uint64_t *b, int64_t nb, int wb,
                            uint64_t *c, int wc)
{
    int64_t nc = na + nb - 1;
    memset(c, 0, sizeof(uint64_t) * (size_t)(nc * wc));
    uint64_t *am = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(na * wa));
    uint64_t *bm = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(nb * wb));
    char *as = (char *)malloc((size_t)na), *bs = (char *)malloc((size_t)nb);
    int *al = (int *)malloc(sizeof(int) * (size_t)na);
    int *bl = (int *)malloc(sizeof(int) * (size_t)nb);
    for (int64_t i = 0; i < na; ++i) {
        as[i] = (char)mw_abs(a + i * wa, wa, am + i * wa);
        int l = wa; while (l > 0 && am[i * wa + l - 1] == 0) --l; al[i] = l;
    }
    for (int64_t j = 0; j < nb; ++j) {
        bs[j] = (char)mw_abs(b + j * wb, wb, bm + j * wb);
        int l = wb; while (l > 0 && bm[j * wb + l - 1] == 0) --l; bl[j] = l;
    }
    uint64_t *prod = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(wa + wb));
    for (int64_t i = 0; i < na; ++i) {
        if (!al[i]) continue;
        for (int64_t j = 0; j < nb; ++j) {
            if (!bl[j]) continue;
            int lp = al[i] + bl[j];
            memset(prod, 0, sizeof(uint64_t) * (size_t)lp);
            for (int u = 0; u < al[i]; ++u) {
                u128 carry = 0;
                for (int v = 0; v < bl[j]; ++v) {
                    u128 t = (u128)am[i * wa + u] * bm[j * wb + v] + prod[u + v] + carry;
                    prod[u + v] = (uint64_t)t;
                    carry = t >> 64;
                }
                prod[u + bl[j]] = (uint64_t)carry;
            }
            mw_addsub_mag(c + (i + j) * wc, wc, prod, lp, 0, as[i] != bs[j]);
        }
    }
    free(prod); free(am); free(bm); free(as); free(bs); free(al); free(bl);
}

void orc_zmul_i64_schoolbook(const int64_t *a, int64_t na, const int64_t *b,
                             int64_t nb, uint64_t *c)
{
    /* 192-bit accumulators: |a_i b_j| < 2^126, at most 2^62 terms */
    int64_t nc = na + nb - 1;
    for (int64_t k = 0; k < nc; ++k) {
        int64_t lo = k - (nb - 1) > 0 ? k - (nb - 1) : 0;
        int64_t hi = k < na - 1 ? k : na - 1;
        u128 acc_lo = 0;   /* low 128 bits (mod 2^128) */
        int64_t acc_hi = 0; /* bits 128..191, signed */
        for (int64_t i = lo; i <= hi; ++i) {
            i128 t = (i128)a[i] * b[k - i];
            u128 ut = (u128)t;
            u128 s = acc_lo + ut;
            acc_hi += (s < acc_lo) ? 1 : 0;
            acc_hi += (t < 0) ? -1 : 0;
            acc_lo = s;
        }
        c[3 * k + 0] = (uint64_t)acc_lo;
        c[3 * k + 1] = (uint64_t)(acc_lo >> 64);
        c[3 * k + 2] = (uint64_t)acc_hi;
    }
}

int orc_bit_width_mw(const uint64_t *a, int64_t na, int wa)
{
    /* mpz_sizeinbase(x, 2) of |x|, which is 1 for x == 0
     * (src/two_convolution_poly_mul.cu:58-70) */
    int best = 1;
    uint64_t *mag = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)wa);
    for (int64_t i = 0; i < na; ++i) {
        mw_abs(a + i * wa, wa, mag);
        int l = wa;
        while (l > 0 && mag[l - 1] == 0) --l;
        if (l) {
            int bits = 64 * (l - 1) + (64 - __builtin_clzll(mag[l - 1]));
            if (bits > best) best = bits;
        }
    }
    free(mag);
    return best;
}

/* bits [pos, pos+M) of a w-word magnitude */
static uint64_t mw_bits(const uint64_t *mag, int w, int pos, int M)
{
    int word = pos >> 6, off = pos & 63;
    if (word >= w) return 0;
    u128 v = mag[word];
    if (word + 1 < w) v |= (u128)mag[word + 1] << 64;
    v >>= off;
    return (uint64_t)v & ((M >= 64) ? ~(uint64_t)0 : (((uint64_t)1 << M) - 1));
}

void orc_limb_split(const uint64_t *a, int64_t na, int wa, int K, int M,
                    sfixn prime, sfixn *out)
{
    uint64_t *mag = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)wa);
    for (int64_t i = 0; i < na; ++i) {
        int neg = mw_abs(a + i * wa, wa, mag);
        for (int j = 0; j < K; ++j) {
            uint64_t limb = mw_bits(mag, wa, j * M, M) % (uint32_t)prime;
            out[i * K + j] = (sfixn)(neg ? (limb ? (uint32_t)prime - limb : 0) : limb);
        }
    }
    free(mag);
}

void orc_scale_x(sfixn *A, int64_t len, int K, sfixn theta, sfixn prime)
{
    for (int64_t i = 0; i < len; ++i)
        A[i] = orc_mul_mod(A[i], orc_pow_mod(theta, (sfixn)(i % K), prime), prime);
}

/* ------------------------------------------------------------------------ */
/* two-convolution integer multiplication                                   */
/* ------------------------------------------------------------------------ */
static const int64_t TC_P[3] = {962592769, 950009857, 940572673};
/* include/two_convolution_poly_mul.h:55-65 */
static const int64_t TC_P2P3 = 893553310574837761LL;
static const int64_t TC_P3P1 = 905388453748801537LL;
static const int64_t TC_P1P2 = 914472618826924033LL;
static const int64_t TC_U23 = -412536414;
static const int64_t TC_U31 = 316662352;
static const int64_t TC_U12 = 89582607;

/* signed CRT of (r1, r2, r3) with the reference's constants
 * (src/two_convolution_poly_mul.cu:305-309).  The reference applies one
 * correction by +-P; the sum can lie in (-P, 2P), so the value is reduced
 * into (-P/2, P/2] completely here. */
static i128 tc_crt(int64_t r1, int64_t r2, int64_t r3)
{
    const i128 P = (i128)TC_P1P2 * TC_P[2];
    i128 e = (i128)((r1 * TC_U23) % TC_P[0]) * TC_P2P3
           + (i128)((r2 * TC_U31) % TC_P[1]) * TC_P3P1
           + (i128)((r3 * TC_U12) % TC_P[2]) * TC_P1P2;
    e %= P;
    if (e < 0) e += P;
    if (e > P / 2) e -= P;
    return e;
}

/* 2-D convolution mod p: x cyclic of length K (negacyclic when theta != 0,
 * via the twist A[y][j] *= theta^j, src/two_convolution_poly_mul.cu:490-512),
 * y linear.  A is na x K, B is nb x K (y-major, x fastest); out gets
 * (na+nb-1) x K. */
static void tc_conv2d(const sfixn *A, int64_t na, const sfixn *B, int64_t nb,
                      int K, sfixn theta, sfixn p, sfixn *out)
{
    int64_t ny = na + nb - 1;
    sfixn ey = ceil_log2(ny);
    int64_t Ly = (int64_t)1 << ey;
    uint64_t P = (uint32_t)p;
    sfixn wK = 1, wKinv = 1, thinv = 0;
    int eK = ceil_log2(K);
    if (K > 1) { wK = orc_primitive_root(eK, p); wKinv = orc_inv_mod(wK, p); }
    if (theta) thinv = orc_inv_mod(theta, p);
    sfixn wy = ey ? orc_primitive_root(ey, p) : 1;
    sfixn wyinv = orc_inv_mod(wy, p);
    /* power tables: wK^j, wK^-j, theta^j, theta^-j for j < K */
    sfixn *pw = (sfixn *)malloc(sizeof(sfixn) * (size_t)(4 * K));
    for (int j = 0; j < K; ++j) {
        pw[j] = orc_pow_mod(wK, j, p);
        pw[K + j] = orc_pow_mod(wKinv, j, p);
        pw[2 * K + j] = theta ? orc_pow_mod(theta, j, p) : 1;
        pw[3 * K + j] = theta ? orc_pow_mod(thinv, j, p) : 1;
    }
    /* x-transformed rows: R[j][y] */
    sfixn *RA = (sfixn *)calloc((size_t)(K * Ly), sizeof(sfixn));
    sfixn *RB = (sfixn *)calloc((size_t)(K * Ly), sizeof(sfixn));
    sfixn *tmp = (sfixn *)malloc(sizeof(sfixn) * (size_t)K);
    const sfixn *srcs[2] = {A, B};
    int64_t lens[2] = {na, nb};
    sfixn *dsts[2] = {RA, RB};
    for (int s = 0; s < 2; ++s) {
        for (int64_t y = 0; y < lens[s]; ++y) {
            for (int j = 0; j < K; ++j) {
                sfixn v = srcs[s][y * K + j];
                if (theta) v = orc_mul_mod(v, pw[2 * K + j], p);
                tmp[j] = v;
            }
            for (int u = 0; u < K; ++u) {           /* DFT_K over x */
                uint64_t acc = 0;
                for (int j = 0; j < K; ++j)
                    acc = (acc + (uint64_t)(uint32_t)tmp[j] * (uint32_t)pw[(u * j) % K]) % P;
                dsts[s][(int64_t)u * Ly + y] = (sfixn)acc;
            }
        }
    }
    for (int u = 0; u < K; ++u) {                   /* y-linear convolutions */
        sfixn *ra = RA + (int64_t)u * Ly, *rb = RB + (int64_t)u * Ly;
        if (Ly > 1) { orc_gsnn_fft((sfixn)Ly, ey, ra, p, wy); orc_gsnn_fft((sfixn)Ly, ey, rb, p, wy); }
        for (int64_t y = 0; y < Ly; ++y) ra[y] = orc_mul_mod(ra[y], rb[y], p);
        if (Ly > 1) orc_gsnn_fft((sfixn)Ly, ey, ra, p, wyinv);
    }
    sfixn scale = orc_inv_mod((sfixn)((Ly * K) % p), p);
    for (int64_t y = 0; y < ny; ++y) {              /* inverse DFT_K, untwist */
        for (int j = 0; j < K; ++j) {
            uint64_t acc = 0;
            for (int u = 0; u < K; ++u)
                acc = (acc + (uint64_t)(uint32_t)RA[(int64_t)u * Ly + y] *
                             (uint32_t)pw[K + (u * j) % K]) % P;
            sfixn v = orc_mul_mod((sfixn)acc, scale, p);
            if (theta) v = orc_mul_mod(v, pw[3 * K + j], p);
            out[y * K + j] = v;
        }
    }
    free(RA); free(RB); free(tmp); free(pw);
}

/* dst (w words, two's complement) += v << shift, v a signed 128-bit value */
static void mw_add_shifted_i128(uint64_t *dst, int w, i128 v, int shift)
{
    int neg = v < 0;
    u128 m = neg ? (u128)(-v) : (u128)v;
    uint64_t part[4] = {0, 0, 0, 0};
    int off = shift & 63, word = shift >> 6;
    part[0] = (uint64_t)(m << off);
    part[1] = (uint64_t)((m << off) >> 64);
    part[2] = off ? (uint64_t)(m >> (128 - off)) : 0;
    if (word >= w) return;
    mw_addsub_mag(dst, w, part, 3, word, neg);
}

int orc_two_convolution_zmul(const uint64_t *a, int64_t na, int wa,
                             const uint64_t *b, int64_t nb, int wb,
                             uint64_t *c, int wc)
{
    int bits = orc_bit_width_mw(a, na, wa);
    int bb = orc_bit_width_mw(b, nb, wb);
    if (bb > bits) bits = bb;
    int64_t nmin = na < nb ? na : nb;
    /* CRT bound on the reconstructed sum/difference (2 * C):
     *   2 * min(K, L) * nmin * (2^M - 1)^2 < P/2,  P ~ 2^89.47. */
    int K = 0, M = 0;
    for (int k = 2; k <= 128; k *= 2) {
        int m = (bits + k - 1) / k;
        if (m > 29) continue;
        int L = (bits + m - 1) / m;
        if (L > k) L = k;
        double lg = 1.0 + __builtin_log2((double)L) + __builtin_log2((double)nmin) + 2.0 * m;
        if (lg < 88.4) { K = k; M = m; break; }
    }
    if (!K) return -1;
    int64_t ny = na + nb - 1;
    int N = K * M;
    sfixn *A = (sfixn *)malloc(sizeof(sfixn) * (size_t)(na * K));
    sfixn *B = (sfixn *)malloc(sizeof(sfixn) * (size_t)(nb * K));
    sfixn *cyc[3], *neg[3];
    for (int r = 0; r < 3; ++r) {
        sfixn p = (sfixn)TC_P[r];
        cyc[r] = (sfixn *)malloc(sizeof(sfixn) * (size_t)(ny * K));
        neg[r] = (sfixn *)malloc(sizeof(sfixn) * (size_t)(ny * K));
        orc_limb_split(a, na, wa, K, M, p, A);
        orc_limb_split(b, nb, wb, K, M, p, B);
        tc_conv2d(A, na, B, nb, K, 0, p, cyc[r]);
        /* theta: primitive 2K-th root, theta^K = -1 */
        sfixn theta = orc_primitive_root(ceil_log2(2 * K), p);
        tc_conv2d(A, na, B, nb, K, theta, p, neg[r]);
    }
    int wt = wc + 2;
    uint64_t *acc = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)wt);
    for (int64_t y = 0; y < ny; ++y) {
        memset(acc, 0, sizeof(uint64_t) * (size_t)wt);
        for (int j = 0; j < K; ++j) {
            int64_t s[3], d[3];
            for (int r = 0; r < 3; ++r) {
                sfixn p = (sfixn)TC_P[r];
                s[r] = orc_add_mod(cyc[r][y * K + j], neg[r][y * K + j], p);
                d[r] = orc_sub_mod(cyc[r][y * K + j], neg[r][y * K + j], p);
            }
            /* product = (sum(2^M) + 2^N diff(2^M)) / 2
             * (src/two_convolution_poly_mul.cu:421-423) */
            mw_add_shifted_i128(acc, wt, tc_crt(s[0], s[1], s[2]), j * M);
            mw_add_shifted_i128(acc, wt, tc_crt(d[0], d[1], d[2]), N + j * M);
        }
        /* arithmetic shift right by one bit */
        for (int i = 0; i < wt; ++i) {
            uint64_t hi = (i + 1 < wt) ? acc[i + 1] : (uint64_t)((int64_t)acc[i] >> 63);
            acc[i] = (acc[i] >> 1) | (hi << 63);
        }
        memcpy(c + y * wc, acc, sizeof(uint64_t) * (size_t)wc);
    }
    free(acc); free(A); free(B);
    for (int r = 0; r < 3; ++r) { free(cyc[r]); free(neg[r]); }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* subproduct tree                                                          */
/* ------------------------------------------------------------------------ */
void orc_subproduct_level(const sfixn *pts, int64_t npts, int lev, sfixn p,
                          sfixn *out)
{
    /* leaves are x - x_i (src/fastPolyEvaluation.cu:512-520); each level is
     * the product of adjacent pairs of the level below (:528-656) */
    int64_t cur_len = 2;   /* coefficients per node, leading one included */
    int64_t nodes = npts;
    sfixn *cur = (sfixn *)malloc(sizeof(sfixn) * (size_t)(npts * 2));
    for (int64_t i = 0; i < npts; ++i) {
        cur[2 * i] = orc_neg_mod(pts[i] % p, p);
        cur[2 * i + 1] = 1;
    }
    for (int l = 1; l <= lev; ++l) {
        int64_t nl = 2 * cur_len - 1, nn = nodes / 2;
        sfixn *nxt = (sfixn *)calloc((size_t)(nn * nl), sizeof(sfixn));
        for (int64_t j = 0; j < nn; ++j) {
            const sfixn *L = cur + (2 * j) * cur_len, *R = cur + (2 * j + 1) * cur_len;
            if (cur_len <= 64)
                orc_plain_polymul((sfixn)(nl - 1), nxt + j * nl, (sfixn)(cur_len - 1), L,
                                  (sfixn)(cur_len - 1), R, p);
            else
                orc_ntt_polymul(nxt + j * nl, (sfixn)(cur_len - 1), L, (sfixn)(cur_len - 1), R, p);
        }
        free(cur);
        cur = nxt; cur_len = nl; nodes = nn;
    }
    memcpy(out, cur, sizeof(sfixn) * (size_t)(nodes * cur_len));
    free(cur);
}

/* ------------------------------------------------------------------------ */
/* multipoint evaluation                                                    */
/* ------------------------------------------------------------------------ */
void orc_poly_eval(const sfixn *F, int64_t n, const sfixn *X, int64_t m, sfixn p,
                   sfixn *out)
{
    for (int64_t j = 0; j < m; ++j) {
        const uint64_t x = (uint64_t)(uint32_t)X[j] % (uint64_t)p;
        uint64_t acc = 0;
        for (int64_t t = n - 1; t >= 0; --t)
            acc = (acc * x + (uint64_t)(uint32_t)F[t]) % (uint64_t)p;
        out[j] = (sfixn)acc;
    }
}

uint64_t orc_splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
