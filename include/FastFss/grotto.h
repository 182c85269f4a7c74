/* FastFss grotto size queries and key-format entry points (backend independent). */
#ifndef FAST_FSS_GROTTO_H
#define FAST_FSS_GROTTO_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/grotto.h:27-33; src/grotto.cpp.
 * Returns FAST_FSS_RUNTIME_ERROR (a stub in the reference). */
FAST_FSS_API int FastFss_grottoKeyZip(void *zippedKey, size_t zippedKeyDataSize, const void *key, size_t keyDataSize, size_t bitWidthIn, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/grotto.h:35-41; src/grotto.cpp.
 * Returns FAST_FSS_RUNTIME_ERROR (a stub in the reference). */
FAST_FSS_API int FastFss_grottoKeyUnzip(void *key, size_t keyDataSize, const void *zippedKey, size_t zippedKeyDataSize, size_t bitWidthIn, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/grotto.h:43-46; src/grotto.cpp. */
FAST_FSS_API int FastFss_grottoGetKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/grotto.h:48-51; src/grotto.cpp. */
FAST_FSS_API int FastFss_grottoGetZippedKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t elementSize, size_t elementNum);

/* Replaces include/FastFss/grotto.h:53-56; src/grotto.cpp. */
FAST_FSS_API int FastFss_grottoGetCacheDataSize(size_t *cacheDataSize, size_t bitWidthIn, size_t elementSize, size_t elementNum);

#ifdef __cplusplus
}
#endif

#endif
