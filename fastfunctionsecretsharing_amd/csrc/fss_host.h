// Host-side helpers shared by the C-ABI entry points: error codes, element
// type dispatch, stream resolution, persistent-grid sizing and size formulas.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "FastFss/errors.h"

namespace fss {

// At most one workgroup fits on a CU: the AES tables take 128 KB of its
// 160 KB LDS.  Launches that fill the GPU use 1024-thread workgroups, one
// per CU (shape_for gives smaller calls smaller ones).
constexpr int kBlock = 1024;
// groupSize > 1 keyGen workgroup size: 768 threads (three waves per SIMD
// with the one table-holding workgroup per CU, 168 registers per lane; the
// few values spilled are live only across the outer slot-chunk loop), 512
// (256 registers) for 64-bit elements with conversions narrower than the
// element, whose chunks hold 16 or more 64-bit slots
constexpr int group_keygen_block(size_t es, int ob)
{
    return (es == 8 && ob > 0 && ob < 8) ? 512 : 768;
}

int device_cu_count(); // cached per device

inline hipStream_t resolve_stream(void *streamPtr)
{
    // nullptr selects the per-thread default stream, as the reference does
    // (src/kernel/parallel_execute.h:164-168)
    return streamPtr ? *(hipStream_t *)streamPtr : hipStreamPerThread;
}

inline int grid_for(size_t work, int wavesPerItemBlock = 1)
{
    size_t blocks = (work + kBlock - 1) / kBlock;
    size_t cap    = (size_t)device_cu_count() * (size_t)wavesPerItemBlock;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    return (int)blocks;
}

// Workgroup count and size for `work` items of one lane each.  Work that
// fills every CU gets the persistent shape, one maxBlock-lane workgroup per
// CU.  Smaller work is spread over all CUs in smaller workgroups (whole
// waves) rather than packed into full workgroups on a few CUs: a small call
// then takes about the latency of one element's walk instead of waiting on
// the LDS throughput of the few CUs it would occupy.
struct LaunchShape
{
    int grid, block;
};

inline LaunchShape shape_for(size_t work, int maxBlock = kBlock)
{
    const size_t cus = (size_t)device_cu_count();
    size_t       b   = (work + cus - 1) / cus;
    b                = (b + 63) / 64 * 64;
    if (b < 64) b = 64;
    if (b > (size_t)maxBlock) b = (size_t)maxBlock;
    size_t g = (work + b - 1) / b;
    if (g > cus) g = cus;
    if (g == 0) g = 1;
    return {(int)g, (int)b};
}

// kernel<<<FSS_SHAPE(work), 0, stream>>>: the grid and block of shape_for(work)
#define FSS_SHAPE(work) ::fss::shape_for(work).grid, ::fss::shape_for(work).block

inline int launch_status()
{
    return hipPeekAtLastError() == hipSuccess ? FAST_FSS_SUCCESS : FAST_FSS_RUNTIME_ERROR;
}

// element sizes of the reference's API: the built-in widths and the 16-,
// 24- and 32-byte wide integers (src/impl/def.h:37-44)
inline bool elem_size_ok(size_t es)
{
    return es == 1 || es == 2 || es == 4 || es == 8 || es == 16 || es == 24 || es == 32;
}

// the widths of the fixed-width kernels (everything but wide.hip)
inline bool elem_size_narrow(size_t es)
{
    return es == 1 || es == 2 || es == 4 || es == 8;
}

inline bool aligned(const void *p, size_t a)
{
    return ((uintptr_t)p % a) == 0;
}

// ---- key / cache sizes (reference src/impl/<prim>.h) ----

// src/impl/dcf.h:38-50
__host__ __device__ inline size_t dcf_key_size(size_t n, size_t g, size_t es, size_t N)
{
    return N * (16 * n + es * n * g + es + es + es * g);
}
// src/impl/dcf.h:61-67
__host__ __device__ inline size_t dcf_cache_size(size_t n, size_t g, size_t es, size_t N)
{
    return N * (16 * n) + N * es * n * g;
}
// src/impl/dpf.h:35-61
__host__ __device__ inline size_t dpf_key_size(size_t n, size_t g, size_t es, size_t N)
{
    return N * (16 * n + es + es + es * g);
}
__host__ __device__ inline size_t dpf_cache_size(size_t n, size_t N)
{
    return N * (16 * n);
}
// src/impl/grotto.h:37-44,81-85
__host__ __device__ inline size_t grotto_key_size(size_t n, size_t es, size_t N)
{
    return N * (16 * (n - 6) + 8 + es + es);
}
__host__ __device__ inline size_t grotto_cache_size(size_t n, size_t N)
{
    return N * (16 * (n - 6) + (n - 6));
}
// src/impl/spline.h:27-48
__host__ __device__ inline size_t spline_cache_size(size_t n, size_t intervals, size_t degree, size_t es, size_t N)
{
    size_t g = intervals * (degree + 1);
    return dcf_cache_size(n, g, es, N) + N * es * (2 * g + degree + 1);
}

} // namespace fss

// Dispatch a runtime element size to a fixed-width unsigned type `T`.
#define FSS_DISPATCH(ES, ...)                \
    switch (ES)                              \
    {                                        \
        case 1: {                            \
            using T = uint8_t;               \
            __VA_ARGS__;                     \
        }                                    \
        break;                               \
        case 2: {                            \
            using T = uint16_t;              \
            __VA_ARGS__;                     \
        }                                    \
        break;                               \
        case 4: {                            \
            using T = uint32_t;              \
            __VA_ARGS__;                     \
        }                                    \
        break;                               \
        case 8: {                            \
            using T = uint64_t;              \
            __VA_ARGS__;                     \
        }                                    \
        break;                               \
        default: return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR; \
    }
