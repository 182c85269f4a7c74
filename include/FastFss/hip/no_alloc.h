/* DCF evaluation that allocates no device memory, on MI355X.
 *
 * No counterpart in the reference.  It is for callers that cannot allocate:
 * code under HIP graph capture that must not record allocations, tight
 * serving loops, and harnesses that forbid hipMalloc.
 *
 * FastFss_hip_dcfEvalNoAlloc takes the arguments of FastFss_cuda_dcfEval
 * (include/FastFss/cuda/dcf.h), validates them in the same order with the same
 * codes, and returns the same shares, bit for bit.  The only difference is
 * where the work queue's tile counter lives: FastFss_cuda_dcfEval allocates it
 * on the stream (hipMallocAsync/hipFreeAsync) for launches that use the queue,
 * and this entry point keeps it in a 256-byte array of static device memory,
 * one copy per GPU, held until the process ends.  It calls no hipMalloc,
 * hipFree or other allocating routine.
 *
 *  - Every call on one GPU shares that GPU's array.  Calls that overlap on one
 *    GPU (on different streams, or from different host threads without
 *    ordering) race on the counter and give wrong shares; calls on one stream
 *    are ordered and safe.
 *  - The counter is cleared by a one-thread kernel on the stream before each
 *    launch that uses it, so it holds the last call's value between calls.
 *  - elementNum is limited to FAST_FSS_HIP_NO_ALLOC_MAX_ELEMENTS (the tile
 *    counter is 32 bits wide, 64 elements a tile); a larger call returns
 *    FAST_FSS_RUNTIME_ERROR without launching.
 */
#ifndef FAST_FSS_HIP_NO_ALLOC_H
#define FAST_FSS_HIP_NO_ALLOC_H

#include <FastFss/api.h>
#include <stddef.h>

#define FAST_FSS_HIP_NO_ALLOC_MAX_ELEMENTS (((size_t)1) << 36)

#ifdef __cplusplus
extern "C" {
#endif

/* No counterpart in the reference; for callers that cannot allocate device memory (see above).
 * Arguments as FastFss_cuda_dcfEval; hipStreamPtr points at a hipStream_t (NULL: per-thread
 * default stream). */
FAST_FSS_API int FastFss_hip_dcfEvalNoAlloc(void *sharedOut, size_t sharedOutSize, const void *maskedX,
                                            size_t maskedXDataSize, const void *key, size_t keyDataSize,
                                            const void *seed, size_t seedDataSize, int partyId, size_t bitWidthIn,
                                            size_t bitWidthOut, size_t groupSize, size_t elementSize,
                                            size_t elementNum, void *cache, size_t cacheDataSize,
                                            void *hipStreamPtr);

#ifdef __cplusplus
}
#endif

#endif
