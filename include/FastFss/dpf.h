/* FastFss dpf size queries and key-format entry points (backend independent). */
#ifndef FAST_FSS_DPF_H
#define FAST_FSS_DPF_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/dpf.h:27-35; src/dpf.cpp.
 * Returns FAST_FSS_RUNTIME_ERROR (a stub in the reference). */
FAST_FSS_API int FastFss_dpfKeyZip(void *zippedKey, size_t zippedKeyDataSize, const void *key, size_t keyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/dpf.h:37-45; src/dpf.cpp.
 * Returns FAST_FSS_RUNTIME_ERROR (a stub in the reference). */
FAST_FSS_API int FastFss_dpfKeyUnzip(void *key, size_t keyDataSize, const void *zippedKey, size_t zippedKeyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/dpf.h:47-52; src/dpf.cpp. */
FAST_FSS_API int FastFss_dpfGetKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/dpf.h:54-59; src/dpf.cpp. */
FAST_FSS_API int FastFss_dpfGetZippedKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/dpf.h:61-64; src/dpf.cpp. */
FAST_FSS_API int FastFss_dpfGetCacheDataSize(size_t *cacheDataSize, size_t bitWidthIn, size_t elementSize, size_t elementNum);

#ifdef __cplusplus
}
#endif

#endif
