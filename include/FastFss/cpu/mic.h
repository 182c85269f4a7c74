/* FastFss_cpu_* mic entry points: the reference's host-buffer API, served on MI355X (gfx950).
 * Pointers are host pointers.  Each call copies its input buffers to device memory on the calling
 * thread's default HIP stream, runs the FastFss_cuda_* entry point of the same name (same validation,
 * same error codes, same results), copies the outputs back and returns when they are in place.
 * Without a usable GPU the calls return FAST_FSS_RUNTIME_ERROR. */

#ifndef FAST_FSS_CPU_MIC_H
#define FAST_FSS_CPU_MIC_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/cpu/mic.h:12-29. */
FAST_FSS_API int FastFss_cpu_dcfMICKeyGen(void *key, size_t keyDataSize, void *z, size_t zDataSize, const void *alpha, size_t alphaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1, size_t seedDataSize1, const void *leftEndpoints, size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/cpu/mic.h:31-51. */
FAST_FSS_API int FastFss_cpu_dcfMICEval(void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *sharedZ, size_t sharedZDataSize, const void *seed, size_t seedDataSize, int partyId, const void *leftEndpoints, size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize);

#ifdef __cplusplus
}
#endif

#endif
