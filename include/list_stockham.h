/* list_stockham.h -- C++-linkage list and bivariate transforms, same
 * signatures as the reference's include/list_stockham.h:33-104.  X holds m
 * contiguous rows of n = 2^k words; one root for all rows; m need not be a
 * power of two.  Synchronous. */
#ifndef CUMODP_AMD_LIST_STOCKHAM_H
#define CUMODP_AMD_LIST_STOCKHAM_H

#include "types.h"

/* I_m (x) DFT_n (src/list_stockham.cu:595-665) */
CUMODP_API void list_stockham_host(sfixn *X, sfixn m, sfixn k, sfixn w, sfixn p);
CUMODP_API void list_stockham_dev(sfixn *X, sfixn m, sfixn k, const sfixn *W, sfixn p);
CUMODP_API void list_stockham_dev(sfixn *X, sfixn m, sfixn k, sfixn w, sfixn p);

/* DFT_{m,n}(X) = (DFT_m (x) I_n)(I_m (x) DFT_n) X on a row-major 2^em x 2^en
 * array (src/list_stockham.cu:1161-1301); Wm/Wn are device power tables */
CUMODP_API void bivariate_stockham_dev(sfixn *X, sfixn em, const sfixn *Wm, sfixn en, const sfixn *Wn,
                            sfixn p);
CUMODP_API void inverse_bivariate_stockham_dev(sfixn *X, sfixn em, const sfixn *Wm, sfixn en,
                                    const sfixn *Wn, sfixn p);
CUMODP_API void bivariate_stockham_host(sfixn *X, sfixn em, sfixn wm, sfixn en, sfixn wn, sfixn p);
CUMODP_API void inverse_bivariate_stockham_host(sfixn *X, sfixn em, sfixn wm, sfixn en, sfixn wn, sfixn p);
/* I_q (x) DFT_{m,n}: q row-major 2^em x 2^en device arrays one after the
 * other, each transformed as bivariate_stockham_dev does (natural order in
 * and out).  Wm/Wn: device power tables [1, w, ..., w^(m/2-1)]
 * (include/list_stockham.h:94-95, src/list_stockham.cu:1803-1846); the
 * overload with roots wm/wn replaces src/list_stockham.cu:1848-1864.  Any
 * q >= 1, 2 <= m <= 2^13, n >= 16 (the reference needs m, n >= 2 N_THD). */
CUMODP_API void list_bivariate_stockham_dev(sfixn *X, sfixn q, sfixn em, const sfixn *Wm, sfixn en,
                                            const sfixn *Wn, sfixn p);
CUMODP_API void list_bivariate_stockham_dev(sfixn *X, sfixn q, sfixn em, sfixn wm, sfixn en,
                                            sfixn wn, sfixn p);
/* F <- DFT^-1(DFT(F) DFT(G)) on 2^em x 2^en device arrays
 * (src/list_stockham.cu:1317-1348); G is overwritten */
CUMODP_API void bi_stockham_poly_mul_dev(sfixn m, sfixn em, sfixn n, sfixn en, sfixn *F, sfixn *G,
                              sfixn p);

#endif
