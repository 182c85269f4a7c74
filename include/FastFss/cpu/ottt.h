/* FastFss_cpu_* ottt entry points: the reference's host-buffer API, served on MI355X (gfx950).
 * Pointers are host pointers.  Each call copies its input buffers to device memory on the calling
 * thread's default HIP stream, runs the FastFss_cuda_* entry point of the same name (same validation,
 * same error codes, same results), copies the outputs back and returns when they are in place.
 * Without a usable GPU the calls return FAST_FSS_RUNTIME_ERROR. */

#ifndef FAST_FSS_CPU_OTTT_H
#define FAST_FSS_CPU_OTTT_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/cpu/ottt.h:12-18. */
FAST_FSS_API int FastFss_cpu_otttKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize, size_t bitWidthIn, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/cpu/ottt.h:20-33. */
FAST_FSS_API int FastFss_cpu_otttLutEval(void *sharedOutE, size_t sharedOutEDataSize, void *sharedOutT, size_t sharedOutTDataSize, const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize, int partyId, const void *lookUpTable, size_t lookUpTableDataSize, size_t bitWidthIn, size_t elementSize, size_t elementNum);

#ifdef __cplusplus
}
#endif

#endif
