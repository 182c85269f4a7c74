/* FastFss_cuda_* prng entry points, served by HIP kernels on MI355X (gfx950). Pointers are device pointers; cudaStreamPtr points at a hipStream_t (NULL: per-thread default stream). */
#ifndef FAST_FSS_CUDA_PRNG_H
#define FAST_FSS_CUDA_PRNG_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/cuda/prng.h:12-12. */
FAST_FSS_API void *FastFss_cuda_prngInit(void);

/* Replaces include/FastFss/cuda/prng.h:14-14. */
FAST_FSS_API void FastFss_cuda_prngRelease(void *prng);

/* Replaces include/FastFss/cuda/prng.h:16-16. */
FAST_FSS_API int FastFss_cuda_prngSetCurrentSeed(void *prng, const void *hostSeed128bit, const void *hostCounter128bit);

/* Replaces include/FastFss/cuda/prng.h:18-18. */
FAST_FSS_API int FastFss_cuda_prngGetCurrentSeed(const void *prng, void *hostSeed128bit, void *hostCounter128bit);

/* Replaces include/FastFss/cuda/prng.h:20-25; src/cuda/prng.cu:152-186. */
FAST_FSS_API int FastFss_cuda_prngGen(void *prng, void *deviceDst, size_t bitWidth, size_t elementSize, size_t elementNum, void *cudaStreamPtr);

#ifdef __cplusplus
}
#endif

#endif
