/* FastFss_cuda_* dpf entry points, served by HIP kernels on MI355X (gfx950). Pointers are device pointers; cudaStreamPtr points at a hipStream_t (NULL: per-thread default stream). */
#ifndef FAST_FSS_CUDA_DPF_H
#define FAST_FSS_CUDA_DPF_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/cuda/dpf.h:13-28. */
FAST_FSS_API int FastFss_cuda_dpfKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize, const void *beta, size_t betaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1, size_t seedDataSize1, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum, void *cudaStreamPtr);

/* Replaces include/FastFss/cuda/dpf.h:30-46. */
FAST_FSS_API int FastFss_cuda_dpfEval(void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize, void *cudaStreamPtr);

/* Replaces include/FastFss/cuda/dpf.h:48-64. */
FAST_FSS_API int FastFss_cuda_dpfEvalAll(void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize, void *cudaStreamPtr);

/* Replaces include/FastFss/cuda/dpf.h:66-84. */
FAST_FSS_API int FastFss_cuda_dpfEvalMulti(void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId, const void *point, size_t pointDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize, void *cudaStreamPtr);

#ifdef __cplusplus
}
#endif

#endif
