/* opt_plain_mul.h -- C++-linkage plain (schoolbook) multiply, the same
 * signature (and so the same mangled symbol) as the reference's
 * include/opt_plain_mul.h:66 (mulCUDA, src/opt_plain_mul.cu:258-384).
 * Synchronous.  A failed call prints the reason to stderr and leaves C
 * untouched. */
#ifndef CUMODP_AMD_OPT_PLAIN_MUL_H
#define CUMODP_AMD_OPT_PLAIN_MUL_H

#include "types.h"

/* C[0 .. n+m-2] = A * B mod Prime for host arrays of n and m coefficients
 * (the reference's "degrees" n and m are these lengths), canonical residues.
 * Prime: any odd modulus in [3, 2^31); no root of unity is needed.  The
 * reference requires n >= m; either order works here.  Products whose shorter
 * operand has at least 128 coefficients run through the NTT multiply when
 * Prime has the roots for it (same result); the rest run the schoolbook
 * kernel.  With CUMODP_VERBOSE set and verify != 1, prints "n m seconds" to
 * stdout as the reference does unconditionally (src/opt_plain_mul.cu:367-369). */
CUMODP_API void mulCUDA(sfixn *A, sfixn *B, sfixn *C, sfixn n, sfixn m, sfixn Prime,
                        sfixn verify);

#endif
