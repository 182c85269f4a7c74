/* FastFss_cpu_* spline entry points: the reference's host-buffer API, served on MI355X (gfx950).
 * Pointers are host pointers.  Each call copies its input buffers to device memory on the calling
 * thread's default HIP stream, runs the FastFss_cuda_* entry point of the same name (same validation,
 * same error codes, same results), copies the outputs back and returns when they are in place.
 * Without a usable GPU the calls return FAST_FSS_RUNTIME_ERROR. */

#ifndef FAST_FSS_CPU_SPLINE_H
#define FAST_FSS_CPU_SPLINE_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/cpu/spline.h:12-35. */
FAST_FSS_API int FastFss_cpu_dcfSplineKeyGen(void *key, size_t keyDataSize, void *e, size_t eDataSize, void *beta, size_t betaDataSize, const void *alpha, size_t alphaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1, size_t seedDataSize1, const void *coefficients, size_t coefficientsDataSize, size_t degree, const void *leftEndpoints, size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize, size_t intervalNum, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/cpu/spline.h:37-61. */
FAST_FSS_API int FastFss_cpu_dcfSplineEval(void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *sharedE, size_t sharedEDataSize, const void *sharedBeta, size_t sharedBetaDataSize, const void *seed, size_t seedDataSize, int partyId, const void *leftEndpoints, size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize, size_t intervalNum, size_t degree, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize);

#ifdef __cplusplus
}
#endif

#endif
