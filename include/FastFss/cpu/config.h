/* FastFss_cpu_* config entry points (reference include/FastFss/cpu/config.h). */
#ifndef FAST_FSS_CPU_CONFIG_H
#define FAST_FSS_CPU_CONFIG_H

#include <FastFss/api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/FastFss/cpu/config.h:18-18; src/cpu/config.cpp:9-24.  Returns -2 for num <= 0.  The
 * count is stored and returned; the host-buffer entry points evaluate on the GPU, not on host threads. */
FAST_FSS_API int FastFss_cpu_setNumThreads(int num);

/* Replaces include/FastFss/cpu/config.h:26-26; src/cpu/config.cpp:26-37.  Starts at the host's
 * hardware thread count. */
FAST_FSS_API int FastFss_cpu_getNumThreads(void);

#ifdef __cplusplus
}
#endif

#endif
