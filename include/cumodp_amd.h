/* cumodp_amd.h -- C-ABI entry points of the MI355X library that have no
 * single counterpart in the reference.  They serve callers that keep data in
 * device memory (no per-call host copies), batch many products, or need the
 * integer multiplier without GMP/thrust types.  Plain pointers and sizes only;
 * `stream` is a hipStream_t passed as void* (NULL = the default stream). */
#ifndef CUMODP_AMD_H
#define CUMODP_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- device memory and synchronisation (for callers without HIP headers) */
CUMODP_API void *cumodp_dev_alloc(size_t bytes);
CUMODP_API cumodp_err cumodp_dev_free(void *ptr);
CUMODP_API cumodp_err cumodp_memcpy_h2d(void *dst, const void *src, size_t bytes);
CUMODP_API cumodp_err cumodp_memcpy_d2h(void *dst, const void *src, size_t bytes);
CUMODP_API cumodp_err cumodp_memcpy_d2d(void *dst, const void *src, size_t bytes);
CUMODP_API cumodp_err cumodp_synchronize(void);
CUMODP_API void *cumodp_stream_create(void);
CUMODP_API cumodp_err cumodp_stream_destroy(void *stream);
CUMODP_API cumodp_err cumodp_stream_synchronize(void *stream);
/* 1 when all work queued on the stream has finished, 0 while some is
 * pending, -1 on error (non-blocking) */
CUMODP_API int cumodp_stream_query(void *stream);
/* error string of the last failing device call in this thread ("" if none) */
CUMODP_API const char *cumodp_last_error(void);

/* ---- univariate multiply on device buffers --------------------------------
 * Device-resident form of stockham_poly_mul_host (src/stockham.cu:562-611):
 * Hd[0 .. min(dh+1, n)) = (Fd * Gd mod p), n = 2^ceil(log2(df+dg+1)), any
 * n >= 2 (no 256 cut-off).  Fd/Gd are read only and hold df+1 / dg+1 words;
 * Hd may alias Fd or Gd.  Scratch space is owned by the library.  Stream
 * ordered (returns before the product is complete). */
CUMODP_API cumodp_err cumodp_fftmul_uni_dev(sfixn dh, sfixn *Hd, sfixn df, const sfixn *Fd, sfixn dg,
                                 const sfixn *Gd, sfixn p, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same product as cumodp_fftmul_uni_dev, bit for bit,
 * with no hipMalloc, hipFree or other allocating call.
 * - Limit: df + dg + 1 <= 2^22 (transform size at most 2^22); larger sizes
 *   return CUMODP_FFT_SIZE_TOO_LARGE.
 * - Scratch: a 32 MiB file-scope device array per GPU, held from the first
 *   load of the library's code until the process ends.  Every call on one
 *   GPU shares it and it keeps the last call's contents (nothing in it has
 *   to start at zero).  Calls that overlap on one GPU (two streams, or two
 *   threads) race on it: order them.
 * - Twiddle tables are only looked up: the prime and transform size must
 *   have been used once before by an allocating call (e.g.
 *   cumodp_fftmul_uni_dev of the same size), else CUMODP_ASSUMPTION_ERROR.
 * Stream ordered, like cumodp_fftmul_uni_dev. */
CUMODP_API cumodp_err cumodp_fftmul_uni_dev_noalloc(sfixn dh, sfixn *Hd, sfixn df,
                                                    const sfixn *Fd, sfixn dg, const sfixn *Gd,
                                                    sfixn p, void *stream);

/* ---- list multiply of pairs of uneven degree ------------------------------
 * One call multiplies `count` independent pairs mod p.  Item i reads
 * F[f_off .. f_off+df], G[g_off .. g_off+dg] and writes H[h_off ..
 * h_off+df+dg] (df+dg+1 words).  Output ranges must not overlap each other
 * or any input.  Products are binned by transform size; sizes up to 2^13 run
 * many products per workgroup in one launch per size class, larger ones as
 * batched four-step transforms.  Replaces a loop of stockham_poly_mul_dev
 * calls; the reference's list path (src/list_stockham.cu:595-630) only
 * transforms rows of one length. */
typedef struct {
    int64_t f_off, g_off, h_off;
    int32_t df, dg;
} cumodp_mul_item;

CUMODP_API cumodp_err cumodp_list_fftmul_dev(const cumodp_mul_item *items, int64_t count, const sfixn *F,
                                  const sfixn *G, sfixn *H, sfixn p, void *stream);
/* Schoolbook products of the same items (no transform, no root of unity):
 * any odd modulus 3 <= p < 2^31, canonical operands, at most 2^30 output
 * words per item.  One wavefront computes 256 consecutive outputs of one
 * item.  Replaces the reference's plain multiply, mulCUDA
 * (src/opt_plain_mul.cu:258-384), and its tree-level list form
 * listPlainMulGpu (src/fastPolyEvaluation.cu:198-280), batched over items of
 * any lengths. */
CUMODP_API cumodp_err cumodp_list_plainmul_dev(const cumodp_mul_item *items, int64_t count,
                                               const sfixn *F, const sfixn *G, sfixn *H,
                                               sfixn p, void *stream);
/* C[0 .. n+m-2] = A[0 .. n-1] * B[0 .. m-1] mod p on device buffers. */
CUMODP_API cumodp_err cumodp_plain_mul_dev(sfixn *C, const sfixn *A, sfixn n, const sfixn *B,
                                           sfixn m, sfixn p, void *stream);
/* The same products as one cumodp_fftmul_uni_dev call per item (the plain
 * per-call form), kept for comparisons with the batched engine. */
CUMODP_API cumodp_err cumodp_list_fftmul_dev_percall(const cumodp_mul_item *items, int64_t count,
                                                     const sfixn *F, const sfixn *G, sfixn *H,
                                                     sfixn p, void *stream);

/* ---- exact integer polynomial multiplication (two convolutions + CRT) ----
 * Replaces two_convolution_poly_mul (src/two_convolution_poly_mul.cu:532-569),
 * whose std::vector<mpz_class> interface needs GMP.  Coefficients are w-word
 * little-endian two's-complement integers (w = 1 for int64).  c receives
 * na+nb-1 coefficients of wc words; wc*64 must cover the product's bits
 * (bits(a) + bits(b) + ceil(log2(min(na, nb))) + 1), else
 * CUMODP_ASSUMPTION_ERROR.  The result is the exact product. */
CUMODP_API cumodp_err cumodp_two_convolution_poly_mul(const uint64_t *a, int64_t na, int wa,
                                                      const uint64_t *b, int64_t nb, int wb,
                                                      uint64_t *c, int wc);
/* Same on device buffers, stream ordered.  max_bits: the largest coefficient
 * bit width of a and b (mpz_sizeinbase semantics); 0 makes the library read
 * the inputs back to compute it, which synchronises. */
CUMODP_API cumodp_err cumodp_two_convolution_poly_mul_dev(const uint64_t *a, int64_t na, int wa,
                                                          const uint64_t *b, int64_t nb, int wb,
                                                          uint64_t *c, int wc, int max_bits,
                                                          void *stream);
/* Many exact integer products in one call (the serving path).  Item i reads
 * na coefficients of wa words at A + a_off and nb of wb words at B + b_off
 * (two's complement, little-endian words) and writes the na+nb-1 product
 * coefficients of wc words at C + c_off (offsets in 64-bit words).  bits:
 * the largest coefficient bit width of the item's operands (>= 1); wc*64 must
 * cover 2*bits + ceil(log2(min(na, nb))) + 1.  Items are grouped by
 * transform plan and each group runs as one batch of launches.  Output ranges
 * must not overlap each other or the inputs.  Stream ordered.  No counterpart
 * in the reference (its two_convolution_poly_mul takes one product). */
typedef struct {
    int64_t a_off, b_off, c_off;
    int64_t na, nb;
    int32_t wa, wb, wc, bits;
} cumodp_zmul_item;
CUMODP_API cumodp_err cumodp_two_convolution_list_dev(const cumodp_zmul_item *items, int64_t count,
                                                      const uint64_t *A, const uint64_t *B,
                                                      uint64_t *C, void *stream);
/* pinned host memory for staging (hipHostMalloc); NULL on failure */
CUMODP_API void *cumodp_host_alloc(size_t bytes);
CUMODP_API cumodp_err cumodp_host_free(void *ptr);
/* asynchronous copies on a stream (pinned host memory for overlap) */
CUMODP_API cumodp_err cumodp_memcpy_h2d_async(void *dst, const void *src, size_t bytes,
                                              void *stream);
CUMODP_API cumodp_err cumodp_memcpy_d2h_async(void *dst, const void *src, size_t bytes,
                                              void *stream);
/* find_largest_bit_width_of_coefficients (src/two_convolution_poly_mul.cu:58-70):
 * the largest mpz_sizeinbase(|x|, 2) over a and b (1 for zero) */
CUMODP_API int cumodp_find_largest_bit_width(const uint64_t *a, int64_t na, int wa,
                                             const uint64_t *b, int64_t nb, int wb);
/* convert_to_modular_bivariate (src/two_convolution_poly_mul.cu:84-203): limb j
 * (bits [jM, jM+M) of |a_i|) of coefficient i at out[i*K + j], reduced mod
 * prime and negated mod prime when a_i < 0 (canonical residues) */
CUMODP_API cumodp_err cumodp_convert_to_modular_bivariate(const uint64_t *a, int64_t na, int wa,
                                                          int K, int M, sfixn prime, sfixn *out);
/* scale_x_argument_dev (src/two_convolution_poly_mul.cu:207-245) on a plain
 * device array: A[i] = A[i] * theta^(i mod K) mod prime */
CUMODP_API cumodp_err cumodp_scale_x_argument_dev(sfixn *A, int64_t len, int K, sfixn theta,
                                                  sfixn prime, void *stream);
/* the (primes, K, M) the multiplier uses for these sizes; 0 if none fits */
CUMODP_API int cumodp_two_convolution_plan(int bits, int64_t na, int64_t nb, int *nprimes,
                                           int *K, int *M);

/* ---- subproduct tree ------------------------------------------------------
 * onlySubproductTree / subProductTree (src/fastPolyEvaluation.cu:512-656,
 * :1411-1434) over 2^k points mod p.  Level i (0 <= i <= k) holds the 2^(k-i)
 * monic products of 2^i consecutive (x - x_j), each stored with its leading
 * coefficient as 2^i + 1 canonical words, at word offset
 * cumodp_subproduct_tree_offset(k, i); level k is the root (the reference does
 * not form it).  The whole tree takes cumodp_subproduct_tree_words(k) words. */
CUMODP_API int64_t cumodp_subproduct_tree_words(int k);
CUMODP_API int64_t cumodp_subproduct_tree_offset(int k, int i);
CUMODP_API cumodp_err cumodp_subproduct_tree_dev(const sfixn *points, int k, sfixn p, sfixn *tree,
                                                 void *stream);
CUMODP_API cumodp_err cumodp_subproduct_tree(const sfixn *points, int k, sfixn p, sfixn *tree);

/* ---- graphs of repeated shapes --------------------------------------------
 * Records the stream-ordered cumodp_*_dev calls made on `stream` (a stream
 * from cumodp_stream_create, not the NULL stream) between begin and end into
 * a HIP graph that replays them with one launch.  Run the same calls once
 * eagerly on that stream before capturing (scratch buffers and twiddle
 * tables cannot be allocated inside a capture: such a call then fails).  A
 * replay reads the device buffers named at capture time and the descriptors
 * as they were then; the graph owns the pinned copies of those descriptors.
 * The graph also holds the library's scratch buffers of that stream (and of
 * its auxiliary streams) until cumodp_graph_destroy: a later eager call on
 * the stream that needs more scratch gets a new buffer, and destroying the
 * stream leaves the held ones alone, so a replay never reads freed memory.
 * A replay still shares that scratch with eager calls on the same stream,
 * so run them in stream order (one stream), not concurrently.
 * No memset nodes are recorded (the library clears memory with kernels). */
CUMODP_API cumodp_err cumodp_graph_begin(void *stream);
CUMODP_API void *cumodp_graph_end(void *stream);
CUMODP_API cumodp_err cumodp_graph_launch(void *graph, void *stream);
CUMODP_API cumodp_err cumodp_graph_destroy(void *graph);
/* ---- fast multipoint evaluation --------------------------------------------
 * fastEvaluation (src/fastPolyEvaluation.cu:1371-1400, with subInvTree :666
 * and the remainder descent :1088-1370): values[j] = F(points[j]) mod p for
 * the 2^k points, F given by its 2^k coefficients (degree < 2^k).  Built on
 * the subproduct tree, a subinverse tree (inverses of the reversed nodes by
 * Newton iteration) and a remainder tree whose divisions are batched
 * products.  The reference returns its internal per-level buffers
 * (struct PolyEvalSteps); these return the values. */
CUMODP_API cumodp_err cumodp_fast_evaluation_dev(const sfixn *F, const sfixn *points, int k,
                                                 sfixn p, sfixn *values, void *stream);
CUMODP_API cumodp_err cumodp_fast_evaluation(const sfixn *F, const sfixn *points, int k, sfixn p,
                                             sfixn *values);
/* ---- tuning: four-step split ----------------------------------------------
 * Transforms of 2^k points above 2^13 run as 2^(k-k2) columns x 2^k2 rows;
 * this chooses k2 (12, 13 or 14) for one k.  The tuner (cumodp_amd/autotune.py)
 * measures the candidates and records its choice. */
CUMODP_API cumodp_err cumodp_tune_set_split(int k, int k2);
/* the current split k2 of 2^k-point transforms (0 when k needs none) */
CUMODP_API int cumodp_tune_get_split(int k);
/* words per workgroup tile of the list multiply's mixed-size kernel: 2048
 * (128 threads), 4096 (256) or 8192 (512, the default); products of up to
 * that many points share tiles, larger ones run one per workgroup or
 * four-step */
CUMODP_API cumodp_err cumodp_tune_set_list_tile(int words);
CUMODP_API int cumodp_tune_get_list_tile(void);
/* cumodp_list_fftmul_dev sends products whose shorter operand has at most
 * `cutoff` coefficients to the schoolbook kernel (same results; default
 * 32, the measured best of the one-launch form; 256 suits the per-class
 * form; 0 sends every product to the NTT; at most 1024) */
CUMODP_API cumodp_err cumodp_tune_set_list_plain(int cutoff);
CUMODP_API int cumodp_tune_get_list_plain(void);
/* cumodp_list_fftmul_dev runs every product of up to 2^15 points, schoolbook
 * and NTT, in one persistent launch (1, default; workgroups take tasks
 * largest first from a shared counter), or one launch per size class (0).
 * Same results.  CUMODP_LIST_PERSISTENT=0 in the environment sets 0. */
CUMODP_API cumodp_err cumodp_tune_set_list_persistent(int on);
CUMODP_API int cumodp_tune_get_list_persistent(void);
/* cumodp_two_convolution_list_dev groups products by plan: 0 (default) one
 * group per exact plan; 1 one group per transform length, with the plan of
 * the widest coefficients of that length (same results, fewer launches, more
 * transform work for the narrower products) */
CUMODP_API cumodp_err cumodp_tune_set_zmul_merge(int merge);
CUMODP_API int cumodp_tune_get_zmul_merge(void);
/* words per workgroup tile of the four-step multiply's column passes: 16384
 * (the default; 1024 threads) or 8192 (512 threads, more workgroups) */
CUMODP_API cumodp_err cumodp_tune_set_col_tile(int words);
CUMODP_API int cumodp_tune_get_col_tile(void);

/* ---- optional per-kernel timing ------------------------------------------
 * Off unless enabled here or with CUMODP_PROFILE=1.  When on, each hot kernel
 * launch is bracketed by HIP events on its stream; cumodp_profile_read
 * synchronises them and returns the total milliseconds and launch count of
 * kernel `id` (names from cumodp_profile_name, ids 0 .. cumodp_profile_count-1). */
CUMODP_API cumodp_err cumodp_profile_enable(int on);
CUMODP_API void cumodp_profile_reset(void);
CUMODP_API int cumodp_profile_count(void);
CUMODP_API const char *cumodp_profile_name(int id);
CUMODP_API cumodp_err cumodp_profile_read(int id, double *total_ms, int64_t *count);

#ifdef __cplusplus
}
#endif

#endif
