/* FastFss_cuda_* spline entry points, served by HIP kernels on MI355X (gfx950). Pointers are device pointers; cudaStreamPtr points at a hipStream_t (NULL: per-thread default stream). */
#ifndef FAST_FSS_CUDA_SPLINE_H
#define FAST_FSS_CUDA_SPLINE_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/cuda/spline.h:12-36. */
FAST_FSS_API int FastFss_cuda_dcfSplineKeyGen(void *key, size_t keyDataSize, void *e, size_t eDataSize, void *beta, size_t betaDataSize, const void *alpha, size_t alphaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1, size_t seedDataSize1, const void *coefficients, size_t coefficientsDataSize, size_t degree, const void *leftEndpoints, size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize, size_t intervalNum, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum, void *cudaStreamPtr);

/* Replaces include/FastFss/cuda/spline.h:38-63. */
FAST_FSS_API int FastFss_cuda_dcfSplineEval(void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *sharedE, size_t sharedEDataSize, const void *sharedBeta, size_t sharedBetaDataSize, const void *seed, size_t seedDataSize, int partyId, const void *leftEndpoints, size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize, size_t intervalNum, size_t degree, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize, void *cudaStreamPtr);

#ifdef __cplusplus
}
#endif

#endif
