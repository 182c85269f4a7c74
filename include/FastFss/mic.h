/* FastFss mic size queries and key-format entry points (backend independent). */
#ifndef FAST_FSS_MIC_H
#define FAST_FSS_MIC_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/mic.h:29-36; src/mic.cpp.
 * Returns FAST_FSS_RUNTIME_ERROR (a stub in the reference). */
FAST_FSS_API int FastFss_dcfMICKeyZip(void *zippedKey, size_t zippedKeyDataSize, const void *key, size_t keyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/mic.h:38-45; src/mic.cpp.
 * Returns FAST_FSS_RUNTIME_ERROR (a stub in the reference). */
FAST_FSS_API int FastFss_dcfMICKeyUnzip(void *key, size_t keyDataSize, const void *zippedKey, size_t zippedKeyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/mic.h:47-51; src/mic.cpp. */
FAST_FSS_API int FastFss_dcfMICGetCacheDataSize(size_t *cacheDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/mic.h:53-57; src/mic.cpp. */
FAST_FSS_API int FastFss_dcfMICGetKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/mic.h:59-63; src/mic.cpp. */
FAST_FSS_API int FastFss_dcfMICGetZippedKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum);

#ifdef __cplusplus
}
#endif

#endif
