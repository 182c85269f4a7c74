/* FastFss_cpu_* grotto entry points: the reference's host-buffer API, served on MI355X (gfx950).
 * Pointers are host pointers.  Each call copies its input buffers to device memory on the calling
 * thread's default HIP stream, runs the FastFss_cuda_* entry point of the same name (same validation,
 * same error codes, same results), copies the outputs back and returns when they are in place.
 * Without a usable GPU the calls return FAST_FSS_RUNTIME_ERROR. */

#ifndef FAST_FSS_CPU_GROTTO_H
#define FAST_FSS_CPU_GROTTO_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/cpu/grotto.h:13-23. */
FAST_FSS_API int FastFss_cpu_grottoKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1, size_t seedDataSize1, size_t bitWidthIn, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/cpu/grotto.h:25-38. */
FAST_FSS_API int FastFss_cpu_grottoEqEval(void *sharedBooleanOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId, size_t bitWidthIn, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize);

/* Replaces include/FastFss/cpu/grotto.h:40-54. */
FAST_FSS_API int FastFss_cpu_grottoEval(void *sharedBooleanOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, bool equalBound, int partyId, size_t bitWidthIn, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize);

/* Replaces include/FastFss/cpu/grotto.h:56-73. */
FAST_FSS_API int FastFss_cpu_grottoMICEval(void *sharedBooleanOut, size_t sharedBooleanOutDataSize, const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId, const void *leftEndpoints, size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize, size_t bitWidthIn, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize);

/* Replaces include/FastFss/cpu/grotto.h:75-97. */
FAST_FSS_API int FastFss_cpu_grottoIntervalLutEval(void *sharedOutE, size_t sharedOutEDataSize, void *sharedOutT, size_t sharedOutTDataSize, const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId, const void *leftEndpoints, size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize, const void *lookUpTable, size_t lookUpTableDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize);

#ifdef __cplusplus
}
#endif

#endif
