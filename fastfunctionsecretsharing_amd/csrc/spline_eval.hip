// Launchers of the degree <= 7 DCF-spline eval kernels (spline_device.h).
// This file is compiled on its own with LLVM's iterative ILP scheduler
// (-mllvm -amdgpu-sched-strategy=iterative-ilp, set in build.py): with the
// default scheduler the walk's AES rounds wait for all outstanding LDS
// lookups over a thousand times per kernel, keeping 1-3 lookups in flight;
// with it a wave keeps up to 16 in flight, and the 2^22-element eval of 16
// intervals ran 7% faster (90.3 -> 83.7 ms, 1024 elements 0.45 -> 0.29 ms) in
// a same-process A/B with equal outputs.  The keyGen and generic-degree
// kernels stay in spline.hip, whose compile the same option crashes.
#include "fss_host.h"
#include "fss_launch.h"
#include "spline_device.h"

namespace fss {

// The smallest split of an element's intervals: 2 from 16 intervals, else 1.
// With every lane busy, the 2-way split of 16 adjacent intervals costs 18
// lane-walks per element instead of 17 (9 steps on two lanes) and still ran
// 2.5-5% faster at 2^20 and 2^22 elements in same-process A/B runs; with 4
// intervals (6 lane-walks instead of 5) it ran 7-10% slower.
static uint32_t spline_min_split(size_t I) { return I >= 16 ? 2u : 1u; }

// Interval blocks per element for the spline walk: a power of two dividing
// 64 (so an element's lanes share one wave), at most I, and enough lanes to
// give every CU a full workgroup, or spline_min_split(I) if that is more.
uint32_t spline_blocks_per_elem(size_t N, size_t I)
{
    const size_t lanes = (size_t)device_cu_count() * kBlock;
    const size_t want  = (lanes + N - 1) / N;
    uint32_t     b     = 1;
    while (b * 2 <= 64 && b * 2 <= want && b * 2 <= I) b *= 2;
    return b < spline_min_split(I) ? spline_min_split(I) : b;
}

template <typename T, int OB, int BY = 32>
static void spline_launch_ob(int grid, int block, hipStream_t st, void *out, const void *x, const void *key, const void *e,
                             const void *beta, const void *seed, int party, const void *left, const void *right,
                             size_t I, int degree, int n, int m, size_t N, uint32_t *counter)
{
    // split only as far as the launch has lanes for (the serving runtime sizes
    // its grid by elements; a wider split it has no lanes for only adds walks)
    uint32_t bpe = spline_blocks_per_elem(N, I);
    while (bpe > spline_min_split(I) && N * bpe > (size_t)grid * (size_t)block) bpe >>= 1;
    if (bpe > 1)
        spline_eval_split_kernel<T, OB, BY><<<grid, block, 0, st>>>(
            (T *)out, (const T *)x, (const uint8_t *)key, (const T *)e, (const T *)beta, (const uint8_t *)seed, party,
            (const T *)left, (const T *)right, I, degree, n, m, N, bpe, counter);
    else
        spline_eval_kernel<T, OB, BY><<<grid, block, 0, st>>>((T *)out, (const T *)x, (const uint8_t *)key, (const T *)e,
                                                           (const T *)beta, (const uint8_t *)seed, party,
                                                           (const T *)left, (const T *)right, I, degree, n, m, N,
                                                           counter);
}

// degree <= 7: the slot-conversion kernel for OB = ceil(m/8) when a linked
// walk's 2(degree+1) slots fit 32 bytes, or 48 bytes for 8-byte slots
// (degree <= 2 at bitWidthOut > 56), else the generic conversion, which pays
// a key schedule per 16-byte window.
template <typename T>
static void spline_launch(int grid, int block, hipStream_t st, void *out, const void *x, const void *key, const void *e,
                          const void *beta, const void *seed, int party, const void *left, const void *right, size_t I,
                          int degree, int n, int m, size_t N, uint32_t *counter)
{
    const int ob = (m + 7) / 8;
    // (64-bit elements with 2-byte outputs keep the generic kernel: its register footprint is smaller)
    const int fast = (2 * (degree + 1) * ob <= 32 && !(sizeof(T) == 8 && ob == 2)) ? ob : 0;
    // small calls with 4-byte slots and AES conversion: each walk on eight
    // lanes, while 8 lanes per item and at least two items per element leave
    // at most two waves per SIMD and an item holds at most 4 intervals (the
    // per-call shape of a small call only; one workgroup per 1024 elements
    // keeps its launch).  Measured against the split kernel (same process):
    // 16 intervals, n = 24: 1024 elements 176 against 295 us, 4096 286
    // against 381; 8192 (8 intervals per item) 510 against 413, so not
    // used; 4 intervals, n = 32: 4096 elements 180 against 385 us.
    if constexpr (sizeof(T) >= 4)
    {
        if (fast == 4 && block < kBlock && 4 * I * (size_t)(degree + 1) > 16)
        {
            const size_t lanes = (size_t)device_cu_count() * 512;
            uint32_t     b     = 1;
            while (b * 2 <= 8 && b * 2 <= I && 8 * (size_t)(b * 2) * N <= lanes) b *= 2;
            if (b >= 2 && (I + b - 1) / b <= 4)
            {
                const LaunchShape os = shape_for(8 * (size_t)b * N);
                spline_eval_oct_kernel<T><<<os.grid, os.block, 0, st>>>(
                    (T *)out, (const T *)x, (const uint8_t *)key, (const T *)e, (const T *)beta, (const uint8_t *)seed,
                    party, (const T *)left, (const T *)right, I, degree, n, m, N, b);
                return;
            }
        }
    }
    if constexpr (sizeof(T) == 8)
    {
        if (fast == 0 && ob == 8 && 2 * (degree + 1) * ob <= 48)
        {
            spline_launch_ob<T, 8, 48>(grid, block, st, out, x, key, e, beta, seed, party, left, right, I, degree, n,
                                       m, N, counter);
            return;
        }
    }
#define FSS_SPLINE_OB(K)                                                                                            \
    case K:                                                                                                         \
        if constexpr (K <= (int)sizeof(T))                                                                          \
        {                                                                                                           \
            spline_launch_ob<T, K>(grid, block, st, out, x, key, e, beta, seed, party, left, right, I, degree, n, m, N,   \
                                   counter);                                                                        \
            return;                                                                                                 \
        }                                                                                                           \
        break;
    switch (fast)
    {
        FSS_SPLINE_OB(1)
        FSS_SPLINE_OB(2)
        FSS_SPLINE_OB(3)
        FSS_SPLINE_OB(4)
        FSS_SPLINE_OB(5)
        FSS_SPLINE_OB(6)
        FSS_SPLINE_OB(7)
        FSS_SPLINE_OB(8)
        default: break;
    }
#undef FSS_SPLINE_OB
    spline_launch_ob<T, 0>(grid, block, st, out, x, key, e, beta, seed, party, left, right, I, degree, n, m, N, counter);
}

int launch_spline_eval_rt(size_t es, int grid, hipStream_t st, void *out, const void *x, const void *key,
                          const void *e, const void *beta, const void *seed, int party, const void *left,
                          const void *right, size_t I, int degree, int n, int m, size_t N, uint32_t *counter,
                          int block)
{
    if (degree > 7) return FAST_FSS_RUNTIME_ERROR;
    FSS_DISPATCH(es, spline_launch<T>(grid, block, st, out, x, key, e, beta, seed, party, left, right, I, degree, n, m, N,
                                      counter));
    return launch_status();
}

} // namespace fss
