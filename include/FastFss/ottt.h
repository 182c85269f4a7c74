/* FastFss ottt size queries and key-format entry points (backend independent). */
#ifndef FAST_FSS_OTTT_H
#define FAST_FSS_OTTT_H

#include <FastFss/api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/ottt.h:28-28; src/ottt.cpp:10-20. */
FAST_FSS_API int FastFss_otttGetKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t elementNum);

#ifdef __cplusplus
}
#endif

#endif
