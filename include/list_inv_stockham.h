/* list_inv_stockham.h -- same signature as the reference's
 * include/list_inv_stockham.h:36 (src/list_inv_stockham.cu:71-88): forward
 * list DFT with w, forward with w^-1, scaling by n^-1, i.e. a round trip
 * that returns X unchanged when w is a primitive n-th root. */
#ifndef CUMODP_AMD_LIST_INV_STOCKHAM_H
#define CUMODP_AMD_LIST_INV_STOCKHAM_H

#include "types.h"

CUMODP_API void list_inv_stockham_host(sfixn *X, sfixn m, sfixn k, sfixn w, sfixn p);

#endif
