// Internal launchers shared between translation units (not part of the C ABI).
// `grid` workgroups of `block` lanes (a multiple of 64, at most 1024); the
// kernels loop grid-stride, so any shape is correct.  The public entry points
// pass shape_for(work); the serving runtime passes one 1024-lane workgroup
// per 1024 elements so that concurrent requests' workgroups interleave on the
// CUs.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "fss_host.h"

struct FastFssBatchGroup;

namespace fss {

// Zeroed 32-bit work-queue counter in stream-ordered memory on `st`
// (nullptr on failure); release it on the same stream after the launch.
uint32_t *queue_counter_acquire(hipStream_t st);
void      queue_counter_release(hipStream_t st, uint32_t *p);

int launch_dcf_eval_g1_rt(size_t es, int grid, hipStream_t st, void *out, const void *x, const void *key,
                          const void *seed, int party, int n, int m, size_t N, int block = kBlock);

int launch_dpf_eval_g1_rt(size_t es, int grid, hipStream_t st, void *out, const void *x, const void *key,
                          const void *seed, int party, int n, int m, size_t N, int block = kBlock);

// MIC calls of at least this many elements run on the work queue, smaller
// ones on the static grid split.  Measured on MI355X (16 intervals, same
// process): the queue is 10% faster at 2^19-2^20 elements for n = 64 (2^20:
// 23.9 against 26.7 ms) and 0-2% for n = 16 and 32; the static split is 1-12%
// faster at 2^16-2^18.
constexpr size_t kMicQueueMinElems = (size_t)1 << 19;

// counter: zeroed work-queue counter, or nullptr for the static grid split
int launch_mic_eval_rt(size_t es, int grid, hipStream_t st, void *out, const void *x, const void *key, const void *z,
                       const void *seed, int party, const void *left, const void *right, size_t I, int n, int m,
                       size_t N, uint32_t *counter, int block = kBlock);

// Interval blocks per element of the spline walk (1 when the call fills the
// GPU); launchers that size their grid from the work multiply N by it.
uint32_t spline_blocks_per_elem(size_t N, size_t I);

// degree <= 7 (the path that needs no per-element workspace)
int launch_spline_eval_rt(size_t es, int grid, hipStream_t st, void *out, const void *x, const void *key,
                          const void *e, const void *beta, const void *seed, int party, const void *left,
                          const void *right, size_t I, int degree, int n, int m, size_t N, uint32_t *counter,
                          int block = kBlock);

// Elements of 16, 24 and 32 bytes (wide.hip): one lane per element, for
// arguments the entry point has already validated.  DPF mode 0 is eval, 1
// evalAll, 2 evalMulti (point, P); Grotto op 0 is eval (mode 0 lt, 1 le, 2 eq),
// 1 MIC and 2 interval LUT, whose `bits` is N * I ints of scratch.
int wide_dcf_keygen(hipStream_t st, void *key, const void *alpha, const void *beta, const void *seed0,
                    const void *seed1, int n, int m, size_t g, size_t es, size_t N);
int wide_dcf_eval(hipStream_t st, void *out, const void *x, const void *key, const void *seed, int party, int n, int m,
                  size_t g, size_t es, size_t N);
int wide_dpf_keygen(hipStream_t st, void *key, const void *alpha, const void *beta, const void *seed0,
                    const void *seed1, int n, int m, size_t g, size_t es, size_t N);
int wide_dpf_eval(hipStream_t st, int mode, void *out, const void *x, const void *key, const void *seed, int party,
                  int n, int m, size_t g, size_t es, size_t N, const void *point, size_t P);
int wide_grotto_keygen(hipStream_t st, void *key, const void *alpha, const void *seed0, const void *seed1, int n,
                       size_t es, size_t N);
int wide_grotto_eval(hipStream_t st, int op, int mode, void *out, void *outT, void *bits, const void *x,
                     const void *key, const void *seed, int party, int n, size_t es, size_t N, const void *left,
                     const void *right, size_t I, const void *lut, size_t lutNum);
int wide_mic_keygen(hipStream_t st, void *key, void *z, const void *alpha, const void *seed0, const void *seed1,
                    const void *left, const void *right, size_t I, int n, int m, size_t es, size_t N);
int wide_mic_eval(hipStream_t st, void *out, const void *x, const void *key, const void *z, const void *seed,
                  int party, const void *left, const void *right, size_t I, int n, int m, size_t es, size_t N);
int wide_spline_keygen(hipStream_t st, void *key, void *e, void *beta, const void *alpha, const void *seed0,
                       const void *seed1, const void *coef, int d, const void *left, const void *right, size_t I,
                       int n, int m, size_t es, size_t N);
int wide_spline_eval(hipStream_t st, void *out, const void *x, const void *key, const void *e, const void *beta,
                     const void *seed, int party, const void *left, const void *right, size_t I, int d, int n, int m,
                     size_t es, size_t N, void *cache);
int wide_ottt_keygen(hipStream_t st, void *key, const void *alpha, int n, size_t es, size_t N);
int wide_ottt_eval(hipStream_t st, void *outE, void *outT, const void *x, const void *key, int party, const void *lut,
                   size_t lutNum, int n, size_t es, size_t N);

// Bytes of one entry of the device group table of a fused batch launch.
size_t batch_table_entry_bytes();

// One persistent batch launch over `groupNum` groups (validated like
// FastFss_hip_batchEval) with tables the caller owns: the group table is
// built in `hostTab` (pinned host memory, `cap` entries), copied to `devTab`
// on `st`, and `devCounter` (4 bytes) is reset on `st` before the launch.
// The caller may reuse the tables once work recorded after the launch on
// `st` has completed.  Allocates nothing and does not synchronize.
int batch_enqueue_owned(const FastFssBatchGroup *groups, size_t groupNum, void *hostTab, void *devTab,
                        uint32_t *devCounter, size_t cap, hipStream_t st);

} // namespace fss
