/* FastFss_cuda_* config entry points, served by HIP kernels on MI355X (gfx950). Pointers are device pointers; cudaStreamPtr points at a hipStream_t (NULL: per-thread default stream). */
#ifndef FAST_FSS_CUDA_CONFIG_H
#define FAST_FSS_CUDA_CONFIG_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/cuda/config.h:18-18; src/cuda/config.cpp.
 * Stored for compatibility; launches size their grids from the CU count (one persistent 1024-thread
 * workgroup per CU). */
FAST_FSS_API int FastFss_cuda_setGridDim(int dim);

/* Replaces include/FastFss/cuda/config.h:26-26; src/cuda/config.cpp. */
FAST_FSS_API int FastFss_cuda_getGridDim(void);

/* Replaces include/FastFss/cuda/config.h:28-28; src/cuda/config.cpp. */
FAST_FSS_API int FastFss_cuda_getFineGrainParallelGridDimThreshold(void);

/* Replaces include/FastFss/cuda/config.h:30-30; src/cuda/config.cpp.
 * Stored for compatibility; small MIC launches always split work across intervals. */
FAST_FSS_API int FastFss_cuda_setFineGrainParallelGridDimThreshold(int dim);

#ifdef __cplusplus
}
#endif

#endif
