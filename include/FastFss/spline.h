/* FastFss spline size queries and key-format entry points (backend independent). */
#ifndef FAST_FSS_SPLINE_H
#define FAST_FSS_SPLINE_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/spline.h:29-38; src/spline.cpp.
 * Returns FAST_FSS_RUNTIME_ERROR (a stub in the reference). */
FAST_FSS_API int FastFss_dcfSplineKeyZip(void *zippedKey, size_t zippedKeyDataSize, const void *key, size_t keyDataSize, size_t degree, size_t intervalNum, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/spline.h:40-49; src/spline.cpp.
 * Returns FAST_FSS_RUNTIME_ERROR (a stub in the reference). */
FAST_FSS_API int FastFss_dcfSplineKeyUnzip(void *key, size_t keyDataSize, const void *zippedKey, size_t zippedKeyDataSize, size_t degree, size_t intervalNum, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/spline.h:51-57; src/spline.cpp. */
FAST_FSS_API int FastFss_dcfSplineGetCacheDataSize(size_t *cacheDataSize, size_t degree, size_t intervalNum, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/spline.h:59-65; src/spline.cpp. */
FAST_FSS_API int FastFss_dcfSplineGetKeyDataSize(size_t *keyDataSize, size_t degree, size_t intervalNum, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/spline.h:67-73; src/spline.cpp. */
FAST_FSS_API int FastFss_dcfSplineGetZippedKeyDataSize(size_t *keyDataSize, size_t degree, size_t intervalNum, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum);

#ifdef __cplusplus
}
#endif

#endif
