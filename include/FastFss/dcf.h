/* FastFss dcf size queries and key-format entry points (backend independent). */
#ifndef FAST_FSS_DCF_H
#define FAST_FSS_DCF_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/dcf.h:62-70; src/dcf.cpp:10-21.
 * Not implemented by the reference either: returns FAST_FSS_RUNTIME_ERROR. */
FAST_FSS_API int FastFss_dcfKeyZip(void *zippedKey, size_t zippedKeyDataSize, const void *key, size_t keyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/dcf.h:105-113; src/dcf.cpp:23-34.
 * Not implemented by the reference either: returns FAST_FSS_RUNTIME_ERROR. */
FAST_FSS_API int FastFss_dcfKeyUnzip(void *key, size_t keyDataSize, const void *zippedKey, size_t zippedKeyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/dcf.h:139-144; src/dcf.cpp:36-63. */
FAST_FSS_API int FastFss_dcfGetKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/dcf.h:170-175; src/dcf.cpp:65-92. */
FAST_FSS_API int FastFss_dcfGetZippedKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/dcf.h:201-206; src/dcf.cpp:94-124. */
FAST_FSS_API int FastFss_dcfGetCacheDataSize(size_t *cacheDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum);

#ifdef __cplusplus
}
#endif

#endif
