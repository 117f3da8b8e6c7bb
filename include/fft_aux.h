/* fft_aux.h -- C++-linkage device helpers with the reference's signatures
 * (include/fft_aux.h:39-98).  Synchronous. */
#ifndef CUMODP_AMD_FFT_AUX_H
#define CUMODP_AMD_FFT_AUX_H

#include "types.h"

/* device W[j] = w^j, j < 2^e (src/fft_aux.cu:87-132) */
CUMODP_API void get_powers_binary(sfixn e, sfixn *W, sfixn w, sfixn p);
/* Y[i] = Y[i] * X[i] mod p, i < n (src/fft_aux.cu:214-225) */
CUMODP_API void pointwise_mul_dev(sfixn n, sfixn k, sfixn *Y, const sfixn *X, sfixn p);
/* X[i] = a * X[i] mod p, i < n (src/fft_aux.cu:238-248) */
CUMODP_API void scale_vector_dev(sfixn a, sfixn n, sfixn *X, sfixn p);
/* zero X[n .. m) (src/fft_aux.cu:296-311) */
CUMODP_API void expand_to_fft_dev(sfixn m, sfixn n, sfixn *X);
/* zero X[0 .. n) (src/fft_aux.cu:316-322) */
CUMODP_API void reset_vector_dev(sfixn n, sfixn *X);

#endif
