/* cumodp.h -- C-linkage API of the MI355X (gfx950) replacement for CUMODP's
 * NTT polynomial-multiplication path.  Drop-in for include/cumodp.h of the
 * reference (declarations :39-129, implementation src/modpn.cu:47-408) for
 * the device queries, the univariate/bivariate FFTs and the FFT multiply.
 *
 * Behaviour kept from the reference:
 *   - all calls are synchronous and use device 0;
 *   - *_uni functions take caller-owned HOST arrays; H may alias F;
 *   - on any device error the functions return CUMODP_FFT_ERROR and leave
 *     the output untouched (modpn falls back to its CPU FFT, FMUL.c:1945-1958).
 * Differences (documented contract instead of the reference's bugs):
 *   - cumodp_fftmul_uni writes min(dh+1, n) words (n = 2^ceil(log2(df+dg+1))),
 *     never n words into a shorter H (src/stockham.cu:603-604);
 *   - an FFT size below 256 returns CUMODP_FFT_SIZE_TOO_SMALL instead of a
 *     silent no-op that reports success (src/stockham.cu:572-576);
 *   - sizes the prime cannot support return CUMODP_FOURIER_DEGREE_TOO_SMALL,
 *     sizes above 2^26 CUMODP_FFT_SIZE_TOO_LARGE.
 * The subresultant/scube entry points are outside this library's scope; the
 * link-compatible stubs report "not available" exactly as the reference does
 * on a device without double precision (src/modpn.cu:321). */
#ifndef CUMODP_AMD_CUMODP_H
#define CUMODP_AMD_CUMODP_H

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* device queries, src/modpn.cu:53-123 */
CUMODP_API int is_cuda_enabled(void);
CUMODP_API int num_of_cuda_devices(void);
CUMODP_API int is_double_float_enabled(void);
/* total device memory in bytes, saturated at UINT_MAX (the reference
 * truncates the 64-bit size to 32 bits) */
CUMODP_API unsigned int global_memory_in_bytes(void);
CUMODP_API float global_memory_in_megabytes(void);
/* src/modpn.cu:111-123: nonzero when the product size fits the device */
CUMODP_API int can_call_fftmul_uni(sfixn df, sfixn dg);
CUMODP_API int can_call_subres2(sfixn dx, sfixn d1, sfixn d2);          /* stub: 0 */
CUMODP_API int can_call_subres3(sfixn dx, sfixn dy, sfixn d1, sfixn d2); /* stub: 0 */

/* src/modpn.cu:181-193: in-place natural-order DFT_n(X, w) of a host array,
 * n = 2^k, w a primitive n-th root of unity mod p */
CUMODP_API cumodp_err cumodp_fft_uni(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p);
/* src/modpn.cu:198-210: inverse (DFT with w^-1, scaled by n^-1) */
CUMODP_API cumodp_err cumodp_invfft_uni(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p);
/* src/modpn.cu:215-244: DFT_{m,n} = (DFT_m (x) I_n)(I_m (x) DFT_n) of the
 * row-major 2^em x 2^en host array X (x fastest), and its inverse */
CUMODP_API cumodp_err cumodp_fft_bivariate(sfixn *X, sfixn em, sfixn wm, sfixn en, sfixn wn, sfixn p);
CUMODP_API cumodp_err cumodp_invfft_bivariate(sfixn *X, sfixn em, sfixn wm, sfixn en, sfixn wn, sfixn p);
/* src/modpn.cu:249-261 -> src/stockham.cu:562-611: H = F * G mod p */
CUMODP_API cumodp_err cumodp_fftmul_uni(sfixn dh, sfixn *H, sfixn df, const sfixn *F, sfixn dg,
                             const sfixn *G, sfixn p);

/* link-compatible stubs for the reference's GPU subresultant chains and
 * scube (src/modpn.cu:269-408); not part of this library */
CUMODP_API sfixn cumodp_subres_chain2_fine(sfixn *S, sfixn B, sfixn w, sfixn npx, sfixn npy,
                                const sfixn *P, sfixn nqx, sfixn nqy, const sfixn *Q, sfixn p);
CUMODP_API sfixn cumodp_subres_chain2_coarse(sfixn *S, sfixn B, sfixn w, sfixn npx, sfixn npy,
                                  const sfixn *P, sfixn nqx, sfixn nqy, const sfixn *Q, sfixn p);
CUMODP_API sfixn cumodp_subres_chain3_fine(sfixn *S, sfixn Bx, sfixn By, sfixn wx, sfixn wy, sfixn npx,
                                sfixn npy, sfixn npz, const sfixn *P, sfixn nqx, sfixn nqy,
                                sfixn nqz, const sfixn *Q, sfixn p);
CUMODP_API sfixn cumodp_subres_chain3_coarse(sfixn *S, sfixn Bx, sfixn By, sfixn wx, sfixn wy, sfixn npx,
                                  sfixn npy, sfixn npz, const sfixn *P, sfixn nqx, sfixn nqy,
                                  sfixn nqz, const sfixn *Q, sfixn p);
CUMODP_API void *init_cuda_scube(sfixn N, const sfixn *sz_p, const sfixn *sz_q, sfixn fp);
CUMODP_API void free_cuda_scube(void *scb);
CUMODP_API void print_cuda_scube(const void *scb);
CUMODP_API cumodp_err build_cuda_scube(void *scb, const sfixn *sz_p, const sfixn *P, const sfixn *sz_q,
                            const sfixn *Q);
CUMODP_API const sfixn *interp_subres_coeff2(sfixn *nx, void *scb, sfixn i, sfixn j);
CUMODP_API const sfixn *interp_subres_coeff3(sfixn *nx, sfixn *ny, void *scb, sfixn i, sfixn j);

/* link-compatible stubs for the reference's GPU gcd and division over a prime
 * field (include/cumodp.h:105-106, src/cudaGcd.cu:401, src/cudaDiv.cu:294),
 * which modpn's solver links (modpn/src/solve2.c:305,854,1049).  Outside
 * this library's scope: they leave every output untouched, print one line to
 * stderr and return -1.0f (the reference returns the elapsed time, >= 0). */
CUMODP_API float gcdPrimeField(sfixn *Dgcd, sfixn *A, sfixn *B, sfixn *G, sfixn n, sfixn m, sfixn p,
                               sfixn optionNormalize);
CUMODP_API float divPrimeField(sfixn *A, sfixn *B, sfixn *R, sfixn *Q, sfixn n, sfixn m, sfixn p);

#ifdef __cplusplus
}
#endif

#endif
