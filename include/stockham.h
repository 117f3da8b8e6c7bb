/* stockham.h -- C++-linkage entry points of the univariate NTT and NTT
 * multiply, same signatures (and so the same mangled symbols) as the
 * reference's include/stockham.h:38-92.  Device-pointer functions work in
 * place on caller-owned device buffers of n = 2^k words; *_host functions
 * copy in and out.  All calls are synchronous.  A failed call prints the
 * reason to stderr and leaves its output untouched. */
#ifndef CUMODP_AMD_STOCKHAM_H
#define CUMODP_AMD_STOCKHAM_H

#include "types.h"

/* X <- DFT_n(X, w), natural order in and out (src/stockham.cu:413-434) */
CUMODP_API void stockham_dev(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p);
/* same with the device table W = [1, w, ..., w^(n/2-1)] (src/stockham.cu:437-453) */
CUMODP_API void stockham_dev(sfixn *X, sfixn n, sfixn k, const sfixn *W, sfixn p);
/* X <- n^-1 DFT_n(X, w^-1) (src/stockham.cu:462-467) */
CUMODP_API void inverse_stockham_dev(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p);
/* W holds the inverse root's powers (src/stockham.cu:472-477) */
CUMODP_API void inverse_stockham_dev(sfixn *X, sfixn n, sfixn k, const sfixn *W, sfixn p);
/* host-array versions (src/stockham.cu:479-503) */
CUMODP_API void stockham_host(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p);
CUMODP_API void inverse_stockham_host(sfixn *X, sfixn n, sfixn k, sfixn w, sfixn p);
/* F <- F * G mod x^n - 1 on device buffers of n = 2^e words; G is
 * overwritten (it holds DFT(G) in the reference, src/stockham.cu:521-542;
 * here its contents are unspecified afterwards) */
CUMODP_API void stockham_poly_mul_dev(sfixn n, sfixn e, sfixn *F, sfixn *G, sfixn p);
/* H = F * G mod p from host arrays (src/stockham.cu:562-611): writes
 * min(dh+1, n) words; FFT sizes below 256 print to stderr and return with
 * H unwritten, as in the reference */
CUMODP_API void stockham_poly_mul_host(sfixn dh, sfixn *H, sfixn df, const sfixn *F, sfixn dg,
                            const sfixn *G, sfixn p);

#endif
