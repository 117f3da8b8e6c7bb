/*
 * oracle.h -- CPU restatement of CUMODP's NTT polynomial-multiplication path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (cumodp_amd/, include/)
 * may call, link or load this code.  It is imported by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and only
 * as the checker (or the timed CPU baseline), never as the thing measured.
 *
 * Every function names the reference file:line whose semantics it follows.
 * The code is written from the mathematics, not copied.
 *
 * Parity pins (what checks each function, see tests/test_oracle.py):
 *   orc_*_mod, orc_pow_mod, orc_inv_mod  -- exhaustive small-prime checks vs
 *                                           Python big ints, Fermat inverse.
 *   orc_primitive_root                   -- order check (w^(2^e)=1, w^(2^(e-1))!=1).
 *   orc_dft_naive                        -- the definition Y_j = sum X_i w^(ij).
 *   orc_stockham / orc_gsnn_fft          -- equal to orc_dft_naive (n <= 2^10);
 *                                           equal to each other above.
 *   orc_plain_polymul / orc_ntt_polymul  -- F=G=1 convolution-count closed form,
 *                                           equal to each other, and the
 *                                           reference's test_poly_mul vectors.
 *   orc_polymul_items                    -- equal to those two per item.
 *   orc_zmul_*                           -- reference gtest KATs and the 27
 *                                           Maple a*b=c benchmark fixtures.
 *   orc_bit_width_mw, orc_limb_split,
 *   orc_scale_x                          -- reference gtest KATs.
 *   orc_two_convolution_zmul             -- equal to orc_zmul_* on KATs,
 *                                           fixtures and random inputs.
 *   orc_poly_eval                        -- the definition (Horner); checked
 *                                           against direct power sums and the
 *                                           subproduct tree's roots.
 *   orc_subproduct_level                 -- evaluation of prod (x - x_i) at
 *                                           random points.
 * No function is left "parity unpinned".
 */
#ifndef CUMODP_ORACLE_H
#define CUMODP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t sfixn;

/* ---- canonical arithmetic mod p (include/inlines.h:70-122) ---- */
sfixn orc_add_mod(sfixn a, sfixn b, sfixn p);
sfixn orc_sub_mod(sfixn a, sfixn b, sfixn p);
sfixn orc_neg_mod(sfixn a, sfixn p);
sfixn orc_mul_mod(sfixn a, sfixn b, sfixn p);
/* include/inlines.h:135-167 (extended Euclid) */
sfixn orc_inv_mod(sfixn a, sfixn p);
/* include/inlines.h:181-200 (negative exponent = inverse power) */
sfixn orc_pow_mod(sfixn a, sfixn e, sfixn p);
/* include/inlines.h:408-413 */
sfixn orc_fourier_degree(sfixn p);
/* include/inlines.h:460-485: srand(1), rand()-driven search (libc RNG) */
sfixn orc_primitive_root(sfixn e, sfixn p);

/* ---- transforms ---- */
/* src/fft_aux.cu:49-59, 87-132: W[j] = w^j, j < n */
void orc_get_powers(sfixn n, sfixn *W, sfixn w, sfixn p);
/* the definition: Y_j = sum_i X_i w^(i j) (no reference line; pins the rest) */
void orc_dft_naive(sfixn n, const sfixn *X, sfixn *Y, sfixn w, sfixn p);
/* include/inlines.h:520-558: Gentleman-Sande + bit reversal, natural order */
void orc_gsnn_fft(sfixn n, sfixn e, sfixn *A, sfixn p, sfixn w);
/* src/stockham.cu:40-46 (factorisation), :75-397 (kernels), :413-453
 * (driver): X <- DFT_n(X, w), applied as butterfly / stride transpose /
 * twiddle passes exactly as the reference sequences them. */
void orc_stockham(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p);
/* src/stockham.cu:462-467: DFT with w^-1 followed by scaling with n^-1 */
void orc_inverse_stockham(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p);
/* src/list_stockham.cu:595-630: I_m (x) DFT_n on m contiguous rows */
void orc_list_stockham(sfixn *X, sfixn m, sfixn k, sfixn w, sfixn p);

/* ---- multiplication mod p ---- */
/* src/fft_aux.cu:805-820: H += F*G (accumulates; caller zeroes H) */
void orc_plain_polymul(sfixn dH, sfixn *H, sfixn dF, const sfixn *F,
                       sfixn dG, const sfixn *G, sfixn p);
/* src/fft_aux.cu:862-879: NTT multiply with the libc-compatible root;
 * H gets dF+dG+1 coefficients */
void orc_ntt_polymul(sfixn *H, sfixn dF, const sfixn *F, sfixn dG,
                     const sfixn *G, sfixn p);
/* Independent products on up to `threads` OpenMP threads (the parallel leg of
 * the CPU baseline): item r is desc[5r .. 5r+4] = (f_off, g_off, h_off, df,
 * dg) and writes df+dg+1 words at H + h_off; the schoolbook when the shorter
 * operand has at most `plain_cut` coefficients, else orc_ntt_polymul (the
 * reference's own split: src/fft_aux.cu:805-820 below, :862-879 above). */
void orc_polymul_items(int64_t count, const int64_t *desc, const sfixn *F, const sfixn *G,
                       sfixn *H, sfixn p, int plain_cut, int threads);
/* src/stockham.cu:562-611 host-wrapper semantics with the overrun fixed:
 * returns 0 and writes min(dh+1, n) words (product, then zeros), or returns
 * 1 and leaves H untouched when n < 256. */
int orc_stockham_poly_mul_host(sfixn dh, sfixn *H, sfixn df, const sfixn *F,
                               sfixn dg, const sfixn *G, sfixn p);

/* ---- exact integer multiplication ---- */
/* schoolbook over Z, int64 inputs, 192-bit two's-complement outputs
 * (3 little-endian words per coefficient): the CPU-bignum oracle */
void orc_zmul_i64_schoolbook(const int64_t *a, int64_t na, const int64_t *b,
                             int64_t nb, uint64_t *c);
/* schoolbook over Z on wa/wb/wc-word two's-complement coefficients */
void orc_zmul_mw_schoolbook(const uint64_t *a, int64_t na, int wa,
                            const uint64_t *b, int64_t nb, int wb,
                            uint64_t *c, int wc);
/* src/two_convolution_poly_mul.cu:58-70 (mpz_sizeinbase(.,2); 0 -> 1) */
int orc_bit_width_mw(const uint64_t *a, int64_t na, int wa);
/* src/two_convolution_poly_mul.cu:84-203: magnitude in K limbs of M bits,
 * limb j of coefficient i at out[i*K+j]; negative coefficients give
 * (prime - limb) for nonzero limbs.  Requires M <= 62. */
void orc_limb_split(const uint64_t *a, int64_t na, int wa, int K, int M,
                    sfixn prime, sfixn *out);
/* src/two_convolution_poly_mul.cu:207-245: A[i] *= theta^(i mod K) */
void orc_scale_x(sfixn *A, int64_t len, int K, sfixn theta, sfixn prime);
/* src/two_convolution_poly_mul.cu:532-569 restated with the x dimension
 * kept at K (cyclic mod x^K-1, negacyclic mod x^K+1, linear in y), the
 * reference's three primes and CRT constants
 * (include/two_convolution_poly_mul.h:51-69), and the recombination
 * (sum + 2^N diff)/2.  Returns 0, or -1 if no K <= 64 satisfies the CRT
 * bound. */
int orc_two_convolution_zmul(const uint64_t *a, int64_t na, int wa,
                             const uint64_t *b, int64_t nb, int wb,
                             uint64_t *c, int wc);

/* ---- subproduct tree (src/fastPolyEvaluation.cu:512-656) ---- */
/* Level `lev` of the tree over `npts` points (npts a power of two):
 * out[j*(2^lev+1) + t] = coefficient t of prod_{i in block j} (x - x_i),
 * blocks of 2^lev consecutive points, leading coefficient included. */
void orc_subproduct_level(const sfixn *pts, int64_t npts, int lev, sfixn p,
                          sfixn *out);

/* ---- multipoint evaluation (src/fastPolyEvaluation.cu:1371-1400) ---- */
/* out[j] = F(X[j]) mod p by Horner's rule, F of n coefficients: the
 * definition the reference's fast evaluation computes (its own check
 * compares against Maple's values). */
void orc_poly_eval(const sfixn *F, int64_t n, const sfixn *X, int64_t m, sfixn p,
                   sfixn *out);

/* SplitMix64 (Steele, Lea, Flood 2014): next output, advancing *state */
uint64_t orc_splitmix64(uint64_t *state);

#ifdef __cplusplus
}
#endif
#endif
