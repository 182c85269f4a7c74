/* Export macro of the FastFss C ABI (reference include/FastFss/api.h:4-14). */
#ifndef FAST_FSS_API_H
#define FAST_FSS_API_H

#if defined(FAST_FSS_BUILD)
#define FAST_FSS_API __attribute__((visibility("default")))
#else
#define FAST_FSS_API
#endif

#endif
