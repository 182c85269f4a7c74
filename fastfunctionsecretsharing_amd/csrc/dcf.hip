// FastFss_cuda_dcfKeyGen / FastFss_cuda_dcfEval on MI355X.
//
// Launch shape (shape_for): one 1024-thread workgroup per CU once the call
// fills every CU (the 128 KB AES tables take most of the CU's LDS), smaller
// workgroups spread over all CUs below that; one element per lane per
// iteration of a grid-stride loop or work-queue tile.  Argument validation
// follows the reference Task::check order (src/kernel/dcf.h:31-69, 115-163) so
// callers see the same error codes.
#include <atomic>
#include <mutex>

#include "FastFss/cuda/dcf.h"
#include "FastFss/hip/no_alloc.h"
#include "dcf_device.h"
#include "fss_host.h"
#include "fss_launch.h"

namespace fss {

template <typename T>
__global__ __launch_bounds__(kBlock) void dcf_keygen_g1_kernel(uint8_t *key, const T *alpha, const T *beta,
                                                               const uint8_t *seed0, const uint8_t *seed1, int n,
                                                               int m, size_t N, uint32_t *counter)
{
    __shared__ uint32_t lds[kLdsTableWords];
    aes_fill_lds(lds);
    __syncthreads();
    const AesCtx c      = aes_ctx(lds);
    const size_t stride = (size_t)n + 3;
    uint8_t     *tails  = key + N * 16 * (size_t)n;
    for_each_item(N, counter, [&](size_t e) {
        uint32_t a[4], b[4];
        load16(a, seed0 + 16 * e);
        load16(b, seed1 + 16 * e);
        dcf_keygen_g1<T, 8>(c, key + e * 16 * (size_t)n, (T *)tails + e * stride, alpha[e], beta ? beta[e] : (T)1,
                            a, b, n, m);
    });
}

template <typename T, int OB>
__global__ __launch_bounds__(group_keygen_block(sizeof(T), OB)) void dcf_keygen_kernel(uint8_t *key, const T *alpha, const T *beta,
                                                            const uint8_t *seed0, const uint8_t *seed1, int n, int m,
                                                            int g, size_t N)
{
    __shared__ uint32_t lds[kLdsTableWords];
    aes_fill_lds(lds);
    __syncthreads();
    const AesCtx c      = aes_ctx(lds);
    const size_t stride = (size_t)n * g + 2 + g;
    uint8_t     *tails  = key + N * 16 * (size_t)n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (size_t)gridDim.x * blockDim.x)
    {
        uint32_t a[4], b[4];
        load16(a, seed0 + 16 * e);
        load16(b, seed1 + 16 * e);
        dcf_keygen_group<T, OB>(c, key + e * 16 * (size_t)n, (T *)tails + e * stride, alpha[e],
                                beta ? beta + e * (size_t)g : nullptr, a, b, n, m, g);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void dcf_eval_g1_kernel(T *out, const T *x, const uint8_t *key,
                                                             const uint8_t *seed, int party, int n, int outBytes,
                                                             size_t N)
{
    __shared__ __align__(16) uint32_t lds[kLdsTableWords];
    aes_fill_lds_staged(lds);
    __syncthreads();
    const AesCtx   c     = aes_ctx(lds);
    const size_t   sw    = 16 * (size_t)n;
    const T       *tails = (const T *)(key + N * sw);
    const size_t   ts    = (size_t)n + 3;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (size_t)gridDim.x * blockDim.x)
    {
        uint32_t s[4];
        load16(s, seed + 16 * e);
        out[e] = dcf_eval_g1<T>(c, key + e * sw, tails + e * ts, x[e], s, party, n, outBytes);
    }
}

// Two lanes per element (dcf_eval_g1_pair); the grid-stride step is even, so
// both lanes of a pair always hold the same element.
template <typename T>
__global__ __launch_bounds__(kBlock) void dcf_eval_pair_kernel(T *out, const T *x, const uint8_t *key,
                                                               const uint8_t *seed, int party, int n, int outBytes,
                                                               size_t N)
{
    __shared__ __align__(16) uint32_t lds[kLdsTableWords];
    aes_fill_lds_staged(lds);
    __syncthreads();
    const AesCtx   c     = aes_ctx(lds);
    const size_t   sw    = 16 * (size_t)n;
    const T       *tails = (const T *)(key + N * sw);
    const size_t   ts    = (size_t)n + 3;
    const uint32_t odd   = threadIdx.x & 1u;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; (g >> 1) < N; g += (size_t)gridDim.x * blockDim.x)
    {
        const size_t e = g >> 1;
        uint32_t     s[4];
        load16(s, seed + 16 * e);
        const T v = dcf_eval_g1_pair<T>(c, key + e * sw, tails + e * ts, x[e], s, party, n, outBytes, odd);
        if (odd) out[e] = v;
    }
}

// Eight lanes per element (dcf_eval_g1_oct); the grid-stride step is a
// multiple of 8, so a group of eight lanes always holds one element.
template <typename T>
__global__ __launch_bounds__(kBlock) void dcf_eval_oct_kernel(T *out, const T *x, const uint8_t *key,
                                                              const uint8_t *seed, int party, int n, int outBytes,
                                                              size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    const uint32_t q = threadIdx.x & 3u, h = (threadIdx.x >> 2) & 1u;
    QuadAes        qa;
    qa.init(c, q);
    const size_t sw    = 16 * (size_t)n;
    const T     *tails = (const T *)(key + N * sw);
    const size_t ts    = (size_t)n + 3;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; (g >> 3) < N; g += (size_t)gridDim.x * blockDim.x)
    {
        const size_t   e = g >> 3;
        const uint32_t s = *(const uint32_t *)(seed + 16 * e + 4 * q);
        const T        v = dcf_eval_g1_oct<T>(c, qa, key + e * sw, tails + e * ts, x[e], s, party, n, outBytes, h, q);
        if ((threadIdx.x & 7u) == 4u) out[e] = v;
    }
}

// E elements per lane: lane `gid` takes elements gid + k*S (k < E, S = total
// lanes) per step of a grid-stride loop over E*S; surplus slots redo the
// first element and are not stored.
template <typename T, int CH, int E>
__global__ __launch_bounds__(kBlock) void dcf_eval_burst_kernel(T *out, const T *x, const uint8_t *key,
                                                                const uint8_t *seed, int party, int n, int outBytes,
                                                                size_t N)
{
    __shared__ __align__(16) uint32_t lds[kLdsTableWords];
    aes_fill_lds_staged(lds);
    __syncthreads();
    const AesCtx c     = aes_ctx(lds);
    const size_t sw    = 16 * (size_t)n;
    const T     *tails = (const T *)(key + N * sw);
    const size_t ts    = (size_t)n + 3;
    const size_t S     = (size_t)gridDim.x * blockDim.x;
    for (size_t e0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e0 < N; e0 += S * E)
    {
        size_t         idx[E];
        const uint8_t *sc[E];
        const T       *tl[E];
        T              xv[E], o[E];
        uint32_t       s[E][4];
#pragma unroll
        for (int k = 0; k < E; k++)
        {
            idx[k] = (e0 + k * S < N) ? e0 + k * S : e0;
            sc[k]  = key + idx[k] * sw;
            tl[k]  = tails + idx[k] * ts;
            xv[k]  = x[idx[k]];
            load16(s[k], seed + 16 * idx[k]);
        }
        dcf_eval_g1_burst<T, CH, E>(c, o, sc, tl, xv, s, party, n, outBytes);
#pragma unroll
        for (int k = 0; k < E; k++)
            if (e0 + k * S < N) out[idx[k]] = o[k];
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void dcf_eval_group_kernel(T *out, const T *x, const uint8_t *key,
                                                                const uint8_t *seed, int party, int n, int m, int g,
                                                                size_t N)
{
    __shared__ __align__(16) uint32_t lds[kLdsTableWords];
    aes_fill_lds_staged(lds);
    __syncthreads();
    const AesCtx c     = aes_ctx(lds);
    const size_t sw    = 16 * (size_t)n;
    const T     *tails = (const T *)(key + N * sw);
    const size_t ts    = (size_t)n * g + 2 + g;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (size_t)gridDim.x * blockDim.x)
    {
        uint32_t s[4];
        load16(s, seed + 16 * e);
        if (g <= 8)
        {
            T acc[8];
            dcf_eval_range<T, 8>(c, acc, key + e * sw, tails + e * ts, x[e], s, party, n, m, g, 0, g);
            for (int j = 0; j < g; j++) out[e * g + j] = acc[j];
        }
        else
        {
            dcf_eval_global<T>(c, out + e * (size_t)g, key + e * sw, tails + e * ts, x[e], s, party, n, m, g);
        }
    }
}

// Work-queue form of the burst kernel: each wave takes the next 64-element
// tile from a device-wide counter (lane L evaluates element 64*tile + L)
// until the tiles run out, so waves that run fast take more tiles and the
// launch ends when the work does, not when the slowest fixed share does.
template <typename T, int CH, bool PRE>
__device__ __forceinline__ void dcf_eval_queue_body(const AesCtx &c, T *out, const T *x, const uint8_t *key,
                                                    const uint8_t *seed, int party, int n, int outBytes, size_t N,
                                                    uint32_t *counter)
{
    const size_t   sw    = 16 * (size_t)n;
    const T       *tails = (const T *)(key + N * sw);
    const size_t   ts    = (size_t)n + 3;
    const uint32_t lane  = threadIdx.x & 63u;
    const uint32_t tiles = (uint32_t)((N + 63) / 64);
    for (;;)
    {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(counter, 1u);
        tile = __builtin_amdgcn_readfirstlane(tile);
        if (tile >= tiles) break;
        const size_t e = (size_t)tile * 64 + lane;
        if (e < N)
        {
            if constexpr (PRE)
            {
                uint32_t s[4];
                load16(s, seed + 16 * e);
                out[e] = dcf_eval_g1_pre<T, CH>(c, key + e * sw, tails + e * ts, x[e], s, party, n, outBytes);
            }
            else
            {
                const uint8_t *sc[1] = {key + e * sw};
                const T       *tl[1] = {tails + e * ts};
                const T        xv[1] = {x[e]};
                uint32_t       s[1][4];
                T              o[1];
                load16(s[0], seed + 16 * e);
                dcf_eval_g1_burst<T, CH, 1>(c, o, sc, tl, xv, s, party, n, outBytes);
                out[e] = o[0];
            }
        }
    }
}

template <typename T, int CH, bool PRE>
__global__ __launch_bounds__(kBlock) void dcf_eval_queue_kernel(T *out, const T *x, const uint8_t *key,
                                                                const uint8_t *seed, int party, int n, int outBytes,
                                                                size_t N, uint32_t *counter)
{
    __shared__ uint32_t lds[kLdsTableWords];
    aes_fill_lds(lds);
    __syncthreads();
    const AesCtx c = aes_ctx(lds);
    dcf_eval_queue_body<T, CH, PRE>(c, out, x, key, seed, party, n, outBytes, N, counter);
}

// The same for workgroups below 1024 lanes, with aes_fill_lds_staged.
template <typename T, int CH, bool PRE>
__global__ __launch_bounds__(kBlock) void dcf_eval_queue_small_kernel(T *out, const T *x, const uint8_t *key,
                                                                      const uint8_t *seed, int party, int n,
                                                                      int outBytes, size_t N, uint32_t *counter)
{
    __shared__ __align__(16) uint32_t lds[kLdsTableWords];
    aes_fill_lds_staged(lds);
    __syncthreads();
    const AesCtx c = aes_ctx(lds);
    dcf_eval_queue_body<T, CH, PRE>(c, out, x, key, seed, party, n, outBytes, N, counter);
}

__global__ void queue_reset_kernel(uint32_t *counter)
{
    *counter = 0;
}

// A zeroed tile counter on `st` (stream-ordered allocation, cleared by a
// kernel so the sequence can be captured into a graph); nullptr on failure.
uint32_t *queue_counter_acquire(hipStream_t st)
{
    uint32_t *p = nullptr;
    if (hipMallocAsync((void **)&p, sizeof(uint32_t), st) != hipSuccess) return nullptr;
    queue_reset_kernel<<<1, 1, 0, st>>>(p);
    return p;
}

void queue_counter_release(hipStream_t st, uint32_t *p)
{
    (void)hipFreeAsync(p, st);
}

// Groupsize-1 evaluation variants: 0 = per-level loads, 1..4 = bursts of
// CH levels with E elements per lane (see dcf_eval_g1_burst) over a static
// grid-stride split, 5 = CH 8 bursts over the work queue, 6 = variant 5 with
// each element's value corrections loaded up front (dcf_eval_g1_pre; elements
// of at most 32 bits, 64-bit elements run variant 5), 7 = two lanes per
// element (dcf_eval_g1_pair; its own launch shape over 2N lanes), 8 = eight
// lanes per element (dcf_eval_g1_oct, column-parallel AES; over 8N lanes).
struct EvalVariant
{
    int ch, e, queue;
};
constexpr EvalVariant kEvalVariants[] = {{1, 1, 0}, {4, 1, 0}, {8, 1, 0}, {4, 2, 0},
                                         {2, 2, 0}, {8, 1, 1}, {8, 1, 2}, {1, 1, 0}, {1, 1, 0}};
constexpr int         kNumEvalVariants = sizeof(kEvalVariants) / sizeof(kEvalVariants[0]);
constexpr size_t      kQueueMinElems   = (size_t)1 << 20; // below this the static split is used

// set from any host thread while others launch: atomics, relaxed (each value
// is independent and only steers the choice of launch)
static std::atomic<int> g_variant_override{-1};                             // FastFss_hip_setEvalVariant
static std::atomic<int> g_variant_table[9] = {-1, -1, -1, -1, -1, -1, -1, -1, -1}; // per element size (-1: default)

int dcf_eval_variant(size_t es, size_t n, size_t N)
{
    (void)n;
    const int o = g_variant_override.load(std::memory_order_relaxed);
    if (o >= 0) return o;
    int v = (es < 9) ? g_variant_table[es].load(std::memory_order_relaxed) : -1;
    if (v >= 0) return v;
    // measured on MI355X (n = 16/32/64): the queue form with preloaded value
    // corrections overtakes the static split at 2^20-2^21 elements and is
    // 9-16% faster at 2^22-2^24, where it also reads each key byte once
    // (n = 32, 2^24: 11.3 instead of 19.0 GB fetched); below that its counter
    // setup and per-tile atomics cost more than they balance.  Lane pairs win
    // while they leave at most two waves per SIMD (2^16 elements: 72 against
    // 88 us for the static split at n = 32), and eight lanes per element
    // (column-parallel AES) while those do (n = 32: 1024 elements 36 against
    // 49 us for lane pairs, 2^14 41 against 52; 2^15 64 against 54).
    if (N >= kQueueMinElems) return 6;
    const size_t lanes = (size_t)device_cu_count() * 512;
    if (8 * N <= lanes) return 8;
    return 2 * N <= lanes ? 7 : 2;
}

// `fixed_counter`: a caller-owned tile counter for the work-queue variants
// (cleared here by a kernel before use), or nullptr for a stream-ordered
// allocation made and freed around the launch.
template <typename T>
int launch_dcf_eval_g1(int v, int grid, int block, hipStream_t st, T *out, const T *x, const uint8_t *key,
                       const uint8_t *seed, int party, int n, int outBytes, size_t N, uint32_t *fixed_counter = nullptr)
{
    if (v == 5 || v == 6)
    {
        uint32_t *counter = fixed_counter;
        if (counter)
            queue_reset_kernel<<<1, 1, 0, st>>>(counter);
        else
            counter = queue_counter_acquire(st);
        if (!counter) return FAST_FSS_RUNTIME_ERROR;
        const bool small = block < kBlock;
        if (v == 6 && sizeof(T) <= 4 && small)
            dcf_eval_queue_small_kernel<T, 8, (sizeof(T) <= 4)>
                <<<grid, block, 0, st>>>(out, x, key, seed, party, n, outBytes, N, counter);
        else if (v == 6 && sizeof(T) <= 4)
            dcf_eval_queue_kernel<T, 8, (sizeof(T) <= 4)>
                <<<grid, block, 0, st>>>(out, x, key, seed, party, n, outBytes, N, counter);
        else if (small)
            dcf_eval_queue_small_kernel<T, 8, false>
                <<<grid, block, 0, st>>>(out, x, key, seed, party, n, outBytes, N, counter);
        else
            dcf_eval_queue_kernel<T, 8, false>
                <<<grid, block, 0, st>>>(out, x, key, seed, party, n, outBytes, N, counter);
        if (!fixed_counter) queue_counter_release(st, counter);
        return FAST_FSS_SUCCESS;
    }
    if (v == 7)
    {
        const LaunchShape ps = shape_for(2 * N);
        dcf_eval_pair_kernel<T><<<ps.grid, ps.block, 0, st>>>(out, x, key, seed, party, n, outBytes, N);
        return FAST_FSS_SUCCESS;
    }
    if (v == 8)
    {
        const LaunchShape ps = shape_for(8 * N);
        dcf_eval_oct_kernel<T><<<ps.grid, ps.block, 0, st>>>(out, x, key, seed, party, n, outBytes, N);
        return FAST_FSS_SUCCESS;
    }
    switch (v)
    {
        case 0: dcf_eval_g1_kernel<T><<<grid, block, 0, st>>>(out, x, key, seed, party, n, outBytes, N); break;
        case 2: dcf_eval_burst_kernel<T, 8, 1><<<grid, block, 0, st>>>(out, x, key, seed, party, n, outBytes, N); break;
        case 3: dcf_eval_burst_kernel<T, 4, 2><<<grid, block, 0, st>>>(out, x, key, seed, party, n, outBytes, N); break;
        case 4: dcf_eval_burst_kernel<T, 2, 2><<<grid, block, 0, st>>>(out, x, key, seed, party, n, outBytes, N); break;
        default: dcf_eval_burst_kernel<T, 4, 1><<<grid, block, 0, st>>>(out, x, key, seed, party, n, outBytes, N); break;
    }
    return FAST_FSS_SUCCESS;
}

// Serving launches choose their own grid shape, so they keep the static split
// (variants 5 to 8 fall back to variant 2 there), except that the per-call
// shape of a small call (block < kBlock) takes the lane-pair and eight-lane
// variants with their own shapes, as the per-call entry point does.
int launch_dcf_eval_g1_rt(size_t es, int grid, hipStream_t st, void *out, const void *x, const void *key,
                          const void *seed, int party, int n, int m, size_t N, int block)
{
    int v = dcf_eval_variant(es, (size_t)n, N);
    if (v == 5 || v == 6 || ((v == 7 || v == 8) && block >= kBlock)) v = 2;
    int rc = FAST_FSS_SUCCESS;
    FSS_DISPATCH(es, rc = launch_dcf_eval_g1<T>(v, grid, block, st, (T *)out, (const T *)x, (const uint8_t *)key,
                                                (const uint8_t *)seed, party, n, (m + 7) / 8, N));
    return rc != FAST_FSS_SUCCESS ? rc : launch_status();
}

// ------------------------------------------ no-allocation DCF evaluation ----
//
// FastFss_hip_dcfEvalNoAlloc keeps the work queue's tile counter in this
// file-scope array instead of a stream-ordered allocation.  Each GPU holds its
// own copy at its own address until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns for that device, kept
// in g_scratch_ptr.  Every call on one GPU shares the array, so calls that
// overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = 256;
// one 32-bit tile counter: the tiles (64 elements each) plus one final fetch
// per wave must fit in it, which holds up to 2^36 elements
constexpr size_t kNoAllocMaxElems = FAST_FSS_HIP_NO_ALLOC_MAX_ELEMENTS;
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchBytes / sizeof(uint32_t)];

constexpr int                     kMaxDevices = 64;
static std::atomic<uint32_t *>    g_scratch_ptr[kMaxDevices];
static std::mutex                 g_scratch_mu;

// The scratch address on the device that owns `st`; nullptr on failure.
static uint32_t *noalloc_scratch(hipStream_t st)
{
    int dev = -1;
    if (hipStreamGetDevice(st, &dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    uint32_t *p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (p) return p;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
    if (p) return p;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return nullptr;
    if (cur != dev && hipSetDevice(dev) != hipSuccess) return nullptr;
    void      *addr = nullptr;
    hipError_t e    = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !addr) return nullptr;
    p = (uint32_t *)addr;
    g_scratch_ptr[dev].store(p, std::memory_order_release);
    return p;
}

} // namespace fss

using namespace fss;

extern "C" FAST_FSS_API int FastFss_hip_setEvalVariant(int variant)
{
    if (variant < -1 || variant >= kNumEvalVariants) return FAST_FSS_RUNTIME_ERROR;
    g_variant_override.store(variant, std::memory_order_relaxed);
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_hip_getEvalVariantCount(void)
{
    return kNumEvalVariants;
}

extern "C" FAST_FSS_API int FastFss_hip_setEvalVariantForElementSize(size_t elementSize, int variant)
{
    if (elementSize >= 9 || !elem_size_ok(elementSize) || variant < -1 || variant >= kNumEvalVariants)
        return FAST_FSS_RUNTIME_ERROR;
    g_variant_table[elementSize].store(variant, std::memory_order_relaxed);
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_cuda_dcfKeyGen(void *key, size_t keyDataSize, const void *alpha,
                                                   size_t alphaDataSize, const void *beta, size_t betaDataSize,
                                                   const void *seed0, size_t seedDataSize0, const void *seed1,
                                                   size_t seedDataSize1, size_t bitWidthIn, size_t bitWidthOut,
                                                   size_t groupSize, size_t elementSize, size_t elementNum,
                                                   void *cudaStreamPtr)
{
    const size_t n = bitWidthIn, m = bitWidthOut, g = groupSize, es = elementSize, N = elementNum;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(0 < n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!(0 < m && m <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (keyDataSize != dcf_key_size(n, g, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (g == 0) return FAST_FSS_INVALID_GROUP_SIZE_ERROR;
    if (alphaDataSize != N * es) return FAST_FSS_INVALID_ALPHA_DATA_SIZE_ERROR;
    if (betaDataSize != 0 && betaDataSize != N * es * g) return FAST_FSS_INVALID_BETA_DATA_SIZE_ERROR;
    if (seedDataSize0 != N * 16 || seedDataSize1 != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed0, 8) || !aligned(seed1, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st = resolve_stream(cudaStreamPtr);
    const void *b  = betaDataSize ? beta : nullptr;
    if (es > 8)
    {
        if (!aligned(alpha, 8) || (b && !aligned(b, 8))) return FAST_FSS_RUNTIME_ERROR;
        return wide_dcf_keygen(st, key, alpha, b, seed0, seed1, (int)n, (int)m, g, es, N);
    }
    if (g == 1)
    {
        uint32_t *counter = queue_counter_acquire(st);
        if (!counter) return FAST_FSS_RUNTIME_ERROR;
        const LaunchShape ls = shape_for(N);
        FSS_DISPATCH(es, dcf_keygen_g1_kernel<T><<<ls.grid, ls.block, 0, st>>>(
                             (uint8_t *)key, (const T *)alpha, (const T *)b, (const uint8_t *)seed0,
                             (const uint8_t *)seed1, (int)n, (int)m, N, counter));
        queue_counter_release(st, counter);
    }
    else
    {
        const int ob = keygen_chunk_ob((int)m, es);
        FSS_DISPATCH(es, FSS_KEYGEN_OB_DISPATCH(T, ob,
                                                (dcf_keygen_kernel<T, OBV><<<grid_for(N), group_keygen_block(sizeof(T), OBV), 0, st>>>(
                                                    (uint8_t *)key, (const T *)alpha, (const T *)b,
                                                    (const uint8_t *)seed0, (const uint8_t *)seed1, (int)n, (int)m,
                                                    (int)g, N))));
    }
    return launch_status();
}

extern "C" FAST_FSS_API int FastFss_cuda_dcfEval(void *sharedOut, size_t sharedOutSize, const void *maskedX,
                                                 size_t maskedXDataSize, const void *key, size_t keyDataSize,
                                                 const void *seed, size_t seedDataSize, int partyId,
                                                 size_t bitWidthIn, size_t bitWidthOut, size_t groupSize,
                                                 size_t elementSize, size_t elementNum, void *cache,
                                                 size_t cacheDataSize, void *cudaStreamPtr)
{
    (void)cache;
    const size_t n = bitWidthIn, m = bitWidthOut, g = groupSize, es = elementSize, N = elementNum;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(0 < n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!(0 < m && m <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (partyId != 0 && partyId != 1) return FAST_FSS_INVALID_PARTY_ID_ERROR;
    if (g == 0) return FAST_FSS_INVALID_GROUP_SIZE_ERROR;
    if (keyDataSize != dcf_key_size(n, g, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (sharedOutSize != N * es * g) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (maskedXDataSize != N * es) return FAST_FSS_INVALID_MASKED_X_DATA_SIZE_ERROR;
    if (seedDataSize != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    if (cacheDataSize != 0 && cacheDataSize != dcf_cache_size(n, g, es, N)) return FAST_FSS_INVALID_CACHE_DATA_SIZE_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t       st = resolve_stream(cudaStreamPtr);
    if (es > 8)
    {
        if (!aligned(sharedOut, 8) || !aligned(maskedX, 8)) return FAST_FSS_RUNTIME_ERROR;
        return wide_dcf_eval(st, sharedOut, maskedX, key, seed, partyId, (int)n, (int)m, g, es, N);
    }
    const LaunchShape ls = shape_for(N);
    if (g == 1)
    {
        const int v  = dcf_eval_variant(es, n, N);
        int       rc = FAST_FSS_SUCCESS;
        FSS_DISPATCH(es, rc = launch_dcf_eval_g1<T>(v, ls.grid, ls.block, st, (T *)sharedOut, (const T *)maskedX,
                                                    (const uint8_t *)key, (const uint8_t *)seed, partyId, (int)n,
                                                    (int)((m + 7) / 8), N));
        if (rc != FAST_FSS_SUCCESS) return rc;
    }
    else
    {
        FSS_DISPATCH(es, dcf_eval_group_kernel<T><<<ls.grid, ls.block, 0, st>>>(
                             (T *)sharedOut, (const T *)maskedX, (const uint8_t *)key, (const uint8_t *)seed, partyId,
                             (int)n, (int)m, (int)g, N));
    }
    return launch_status();
}

extern "C" FAST_FSS_API int FastFss_hip_dcfEvalNoAlloc(void *sharedOut, size_t sharedOutSize, const void *maskedX,
                                                       size_t maskedXDataSize, const void *key, size_t keyDataSize,
                                                       const void *seed, size_t seedDataSize, int partyId,
                                                       size_t bitWidthIn, size_t bitWidthOut, size_t groupSize,
                                                       size_t elementSize, size_t elementNum, void *cache,
                                                       size_t cacheDataSize, void *hipStreamPtr)
{
    (void)cache;
    const size_t n = bitWidthIn, m = bitWidthOut, g = groupSize, es = elementSize, N = elementNum;
    if (!elem_size_narrow(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(0 < n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!(0 < m && m <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (partyId != 0 && partyId != 1) return FAST_FSS_INVALID_PARTY_ID_ERROR;
    if (g == 0) return FAST_FSS_INVALID_GROUP_SIZE_ERROR;
    if (keyDataSize != dcf_key_size(n, g, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (sharedOutSize != N * es * g) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (maskedXDataSize != N * es) return FAST_FSS_INVALID_MASKED_X_DATA_SIZE_ERROR;
    if (seedDataSize != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    if (cacheDataSize != 0 && cacheDataSize != dcf_cache_size(n, g, es, N)) return FAST_FSS_INVALID_CACHE_DATA_SIZE_ERROR;
    // beyond this the scratch's 32-bit tile counter could wrap
    if (N > kNoAllocMaxElems) return FAST_FSS_RUNTIME_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t       st = resolve_stream(hipStreamPtr);
    const LaunchShape ls = shape_for(N);
    if (g == 1)
    {
        uint32_t *counter = noalloc_scratch(st);
        if (!counter) return FAST_FSS_RUNTIME_ERROR;
        const int v  = dcf_eval_variant(es, n, N);
        int       rc = FAST_FSS_SUCCESS;
        FSS_DISPATCH(es, rc = launch_dcf_eval_g1<T>(v, ls.grid, ls.block, st, (T *)sharedOut, (const T *)maskedX,
                                                    (const uint8_t *)key, (const uint8_t *)seed, partyId, (int)n,
                                                    (int)((m + 7) / 8), N, counter));
        if (rc != FAST_FSS_SUCCESS) return rc;
    }
    else
    {
        FSS_DISPATCH(es, dcf_eval_group_kernel<T><<<ls.grid, ls.block, 0, st>>>(
                             (T *)sharedOut, (const T *)maskedX, (const uint8_t *)key, (const uint8_t *)seed, partyId,
                             (int)n, (int)m, (int)g, N));
    }
    return launch_status();
}
