/* FastFss_cuda_* dcf entry points, served by HIP kernels on MI355X (gfx950). Pointers are device pointers; cudaStreamPtr points at a hipStream_t (NULL: per-thread default stream). */
#ifndef FAST_FSS_CUDA_DCF_H
#define FAST_FSS_CUDA_DCF_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/cuda/dcf.h:85-100; src/cuda/dcf.cu:9-49. */
FAST_FSS_API int FastFss_cuda_dcfKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize, const void *beta, size_t betaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1, size_t seedDataSize1, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum, void *cudaStreamPtr);

/* Replaces include/FastFss/cuda/dcf.h:167-183; src/cuda/dcf.cu:51-93.
 * The hot path: one MI355X kernel walks each element's depth-n tree.  The cache is validated and
 * otherwise unused (it only memoises identical intermediate values in the reference). */
FAST_FSS_API int FastFss_cuda_dcfEval(void *sharedOut, size_t sharedOutSize, const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize, void *cudaStreamPtr);

#ifdef __cplusplus
}
#endif

#endif
