/* Deprecated aliases of the OTTT entry points (reference include/FastFss/cpu/onehot.h:13-18). */
#ifndef FAST_FSS_CPU_ONEHOT_H
#define FAST_FSS_CPU_ONEHOT_H

#include <FastFss/cpu/ottt.h>

#define FastFss_cpu_onehotKeyGen  FastFss_cpu_otttKeyGen
#define FastFss_cpu_onehotLutEval FastFss_cpu_otttLutEval

#endif
