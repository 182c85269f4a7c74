/* Full-domain Grotto evaluation on MI355X.
 *
 * No counterpart in the reference, whose Grotto API evaluates one point per
 * element (include/FastFss/cuda/grotto.h).  This entry point evaluates every
 * key on every point of its domain [0, 2^bitWidthIn), and its output is defined
 * by the per-point entry points:
 *
 *   mode FAST_FSS_GROTTO_FULL_LT:  point x of key e = FastFss_cuda_grottoEval(equalBound = false) at x
 *   mode FAST_FSS_GROTTO_FULL_LE:  point x of key e = FastFss_cuda_grottoEval(equalBound = true)  at x
 *   mode FAST_FSS_GROTTO_FULL_EQ:  point x of key e = FastFss_cuda_grottoEqEval                   at x
 *
 * Output layout:
 *   packed == 0: elementNum x 2^bitWidthIn group elements of elementSize bytes,
 *                element e * 2^bitWidthIn + x holding the 0/1 share of point x;
 *   packed == 1: elementNum x 2^(bitWidthIn-6) little-endian u64 words, bit
 *                (x & 63) of word e * 2^(bitWidthIn-6) + (x >> 6) holding it.
 *
 * Each key's tree is expanded once: the 2^(n-6) leaves of a key are split
 * over lanes, each lane walks to the root of its subtree and expands that
 * subtree depth-first, so every inner node costs one key schedule and two
 * AES blocks instead of one walk per point.
 *
 * Keys, seeds and outputs are device pointers (key and seed 8-byte aligned,
 * output 16-byte aligned). The key layout is that of FastFss_cuda_grottoKeyGen.
 * bitWidthIn is limited to [6, min(8 * elementSize, 40)].  Returns FastFssError
 * codes; an invalid mode or packed flag returns FAST_FSS_RUNTIME_ERROR.
 */
#ifndef FAST_FSS_HIP_GROTTO_FULL_H
#define FAST_FSS_HIP_GROTTO_FULL_H

#include <FastFss/api.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum FastFssGrottoFullMode
{
    FAST_FSS_GROTTO_FULL_LT = 0,
    FAST_FSS_GROTTO_FULL_LE = 1,
    FAST_FSS_GROTTO_FULL_EQ = 2,
};

FAST_FSS_API int FastFss_hip_grottoEvalFullDomain(void *sharedOut, size_t sharedOutDataSize, const void *key,
                                                  size_t keyDataSize, const void *seed, size_t seedDataSize,
                                                  int partyId, int mode, int packed, size_t bitWidthIn,
                                                  size_t elementSize, size_t elementNum, void *hipStreamPtr);

/* Number of lanes the subtree split of the full-domain evaluations (this
 * entry point and FastFss_cuda_dpfEvalAll) aims for; 0 restores the default,
 * one 1024-lane workgroup per CU.  Fewer lanes means deeper subtrees per lane
 * (at most 8 levels); results do not depend on it.  Process-wide setting. */
FAST_FSS_API int FastFss_hip_grottoFullSetLaneTarget(size_t lanes);

#ifdef __cplusplus
}
#endif

#endif
