/* FastFss_cpu_* prng entry points: the reference's host-buffer API, served on MI355X (gfx950).
 * Pointers are host pointers.  Each call copies its input buffers to device memory on the calling
 * thread's default HIP stream, runs the FastFss_cuda_* entry point of the same name (same validation,
 * same error codes, same results), copies the outputs back and returns when they are in place.
 * Without a usable GPU the calls return FAST_FSS_RUNTIME_ERROR. */
#ifndef FAST_FSS_CPU_PRNG_H
#define FAST_FSS_CPU_PRNG_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/cpu/prng.h:12-12.  Seed and counter start at zero, as the reference's. */
FAST_FSS_API void *FastFss_cpu_prngInit(void);

/* Replaces include/FastFss/cpu/prng.h:14-14. */
FAST_FSS_API void FastFss_cpu_prngRelease(void *prng);

/* Replaces include/FastFss/cpu/prng.h:16-16.  A NULL counter sets the counter to zero. */
FAST_FSS_API int FastFss_cpu_prngSetCurrentSeed(void *prng, const void *seed128bit, const void *counter128bit);

/* Replaces include/FastFss/cpu/prng.h:18-18. */
FAST_FSS_API int FastFss_cpu_prngGetCurrentSeed(const void *prng, void *seed128bit, void *counter128bit);

/* Replaces include/FastFss/cpu/prng.h:20-20.  Fills elementSize * elementNum bytes of the host buffer dst
 * with the AES-128-CTR stream and advances the counter; bitWidth is validated, not applied. */
FAST_FSS_API int FastFss_cpu_prngGen(void *prng, void *dst, size_t bitWidth, size_t elementSize, size_t elementNum);

#ifdef __cplusplus
}
#endif

#endif
