/* fastPolyEvaluation.h -- the reference's subproduct-tree and fast-evaluation
 * drivers, same names, signatures and result struct as its
 * include/fastPolyEvaluation.h:73-99 (src/fastPolyEvaluation.cu:1371-1470),
 * C++ linkage.
 *
 * Each function takes host arrays (F and points of 2^k words), runs on the
 * GPU and returns a struct PolyEvalSteps of DEVICE arrays in the reference's
 * layout.  Every non-null pointer in it is a separate hipMalloc allocation
 * that the caller frees with hipFree (cudaFree in the reference).  The two
 * halves of the tree are kept apart as in the reference: the left half holds
 * the first 2^(k-1) points, the right half the rest.
 *   Ml[i], Mr[i]     (0 <= i < k) subproduct-tree level i of each half,
 *                    2^(k-1-i) nodes.  Levels below plainMulLimit (8) store
 *                    2^i + 1 words per node, leading 1 included; level 0
 *                    holds (-x, 1) with -x the 32-bit negation of the point,
 *                    as the reference writes it.  Levels >= 8 store the 2^i
 *                    low coefficients of each node (the leading 1 implicit).
 *   InvMl[i], InvMr[i] (8 <= i < k) subinverse tree: rev(M)^-1 mod x^(2^i)
 *                    for each node of level i, 2^i words per node.
 *   fL[j], fR[j]     (1 <= j <= k) remainder tree: F mod each node of level
 *                    j - 1, 2^(j-1) words per node; fL[1] and fR[1] are the
 *                    values F(x) at the points of each half.
 *   timeSubTree, timeSinvTree, timeEva: 0 (the reference leaves them unset).
 * What each function fills:
 *   onlySubproductTree: Ml, Mr and, when k > 8, InvMl[8], InvMr[8] (the
 *                       reference computes that level with the tree).
 *   onlyTrees:          Ml, Mr, InvMl, InvMr.
 *   fastEvaluation:     fL[1], fR[1]; with flag == 1 (the reference's
 *                       "keep every step for checking" mode) also Ml, Mr,
 *                       InvMl, InvMr and fL[j], fR[j] for every j.
 * Unlike the reference these are reentrant (no global struct) and accept
 * 1 <= k <= 26 for any prime with 2^max(k,4) | p - 1.  On an error the
 * reason is printed to stderr and every pointer of the result is null. */
#ifndef CUMODP_AMD_FASTPOLYEVALUATION_H
#define CUMODP_AMD_FASTPOLYEVALUATION_H

#include "types.h"

#define MAX_LEVEL 27

/* levels >= plainMulLimit are stored without their leading coefficient */
const sfixn plainMulLimit = 8;

struct PolyEvalSteps {
    sfixn *Ml[MAX_LEVEL], *Mr[MAX_LEVEL];
    sfixn *InvMl[MAX_LEVEL], *InvMr[MAX_LEVEL];
    sfixn *fL[MAX_LEVEL], *fR[MAX_LEVEL];
    float timeSubTree, timeSinvTree, timeEva;
};

/* the tree, the subinverse tree and the values of F (src/fastPolyEvaluation.cu:1371-1400) */
CUMODP_API struct PolyEvalSteps fastEvaluation(sfixn *F, sfixn *points, sfixn k, sfixn p,
                                               sfixn flag);
/* the subproduct tree only; F and flag are unused (src/fastPolyEvaluation.cu:1411-1434) */
CUMODP_API struct PolyEvalSteps onlySubproductTree(sfixn *F, sfixn *points, sfixn k, sfixn p,
                                                   sfixn flag);
/* the subproduct and subinverse trees (src/fastPolyEvaluation.cu:1445-1470) */
CUMODP_API struct PolyEvalSteps onlyTrees(sfixn *points, sfixn k, sfixn p, sfixn flag);

#endif
