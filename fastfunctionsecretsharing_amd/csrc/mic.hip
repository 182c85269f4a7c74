// DCF-based multiple-interval containment (MIC) on MI355X (reference
// src/impl/mic.h, src/kernel/mic.h).  Each element's single DCF key (on
// alpha - 1, beta = 1) is evaluated at the interval end points; adjacent
// intervals share the walk at their common point.
#include "FastFss/cuda/mic.h"
#include "dcf_device.h"
#include "fss_host.h"
#include "fss_launch.h"

namespace fss {

// ------------------------------------------------------------------ MIC ----

template <typename T>
__global__ __launch_bounds__(kBlock) void mic_keygen_kernel(uint8_t *key, T *z, const T *alpha, const uint8_t *seed0,
                                                            const uint8_t *seed1, const T *left, const T *right,
                                                            size_t I, int n, int m, size_t N, uint32_t *counter)
{
    FSS_KERNEL_PROLOGUE
    T *tails = (T *)(key + N * 16 * (size_t)n);
    const T MAX = mod_bits<T>((T)~(T)0, n);
    for_each_item(N, counter, [&](size_t e) {
        uint32_t a[4], b[4];
        load16(a, seed0 + 16 * e);
        load16(b, seed1 + 16 * e);
        const T al = alpha[e];
        // DCF on alpha - 1 with beta = 1 (src/impl/mic.h:22)
        dcf_keygen_g1<T, 8>(c, key + e * 16 * (size_t)n, tails + e * (size_t)(n + 3), (T)(al - 1), (T)1, a, b, n, m);
        for (size_t i = 0; i < I; i++)
        {
            const T qP = mod_bits<T>((T)(right[i] + 1), n);
            const T aP = mod_bits<T>((T)(left[i] + al), n);
            const T aQ = mod_bits<T>((T)(right[i] + al), n);
            const T aQp = mod_bits<T>((T)(right[i] + 1 + al), n);
            const T p   = mod_bits<T>(left[i], n);
            const int zv = (int)(aP > aQ) - (int)(aP > p) + (int)(aQp > qP) + (int)(aQ == MAX);
            z[e * I + i] = (T)zv;
        }
    });
}

// Interval shares [first, first+cnt) of one element (src/impl/mic.h:43-109).
template <typename T>
__device__ __forceinline__ void mic_eval_range(const AesCtx &c, T *out, const T *zs, const uint8_t *sCW,
                                               const T *tail, T x, const uint32_t seed[4], int party, const T *left,
                                               const T *right, size_t first, size_t cnt, int n, int outBytes)
{
    x    = mod_bits<T>(x, n);
    T sp = 0, sq = 0;
    for (size_t i = first; i < first + cnt; i++)
    {
        const T qP  = mod_bits<T>((T)(right[i] + 1), n);
        const T xP  = mod_bits<T>((T)(x - 1 - left[i]), n);
        const T xQp = mod_bits<T>((T)(x - 1 - (T)(right[i] + 1)), n);
        const T p   = mod_bits<T>(left[i], n);
        uint32_t       s[1][4];
        const uint8_t *sc[1] = {sCW};
        const T       *tl[1] = {tail};
        if (i > first && left[i] == mod_bits<T>((T)(right[i - 1] + 1), n))
            sp = sq;
        else
        {
            s[0][0] = seed[0], s[0][1] = seed[1], s[0][2] = seed[2], s[0][3] = seed[3];
            const T xin[1] = {xP};
            T       r[1];
            dcf_eval_g1_burst<T, 8, 1>(c, r, sc, tl, xin, s, party, n, outBytes);
            sp = r[0];
        }
        s[0][0] = seed[0], s[0][1] = seed[1], s[0][2] = seed[2], s[0][3] = seed[3];
        const T xin[1] = {xQp};
        T       r[1];
        dcf_eval_g1_burst<T, 8, 1>(c, r, sc, tl, xin, s, party, n, outBytes);
        sq = r[0];
        T o = (T)(sq - sp + zs[i]);
        if (party == 1) o = (T)(o + (T)((int)(x > p) - (int)(x > qP)));
        out[i] = o;
    }
}

// Work item = (element, block of intervals): small element counts still fill
// the GPU; the first interval of a block is evaluated without the reuse of
// its predecessor, which gives the same value (src/kernel/mic.h:189-272).
template <typename T>
__device__ __forceinline__ void mic_eval_body(const AesCtx &c, T *out, const T *x, const uint8_t *key, const T *z,
                                              const uint8_t *seed, int party, const T *left, const T *right, size_t I,
                                              size_t blocksPerElem, int n, int outBytes, size_t N, uint32_t *counter)
{
    const T     *tails = (const T *)(key + N * 16 * (size_t)n);
    const size_t span  = (I + blocksPerElem - 1) / blocksPerElem;
    for_each_item(N * blocksPerElem, counter, [&](size_t w) {
        const size_t e = w / blocksPerElem, blk = w % blocksPerElem;
        const size_t first = blk * span;
        if (first >= I) return;
        const size_t cnt = (first + span < I) ? span : (I - first);
        uint32_t     s[4];
        load16(s, seed + 16 * e);
        mic_eval_range<T>(c, out + e * I, z + e * I, key + e * 16 * (size_t)n, tails + e * (size_t)(n + 3), x[e], s,
                          party, left, right, first, cnt, n, outBytes);
    });
}

template <typename T>
__global__ __launch_bounds__(kBlock) void mic_eval_kernel(T *out, const T *x, const uint8_t *key, const T *z,
                                                          const uint8_t *seed, int party, const T *left,
                                                          const T *right, size_t I, size_t blocksPerElem, int n,
                                                          int outBytes, size_t N, uint32_t *counter)
{
    FSS_KERNEL_PROLOGUE
    mic_eval_body<T>(c, out, x, key, z, seed, party, left, right, I, blocksPerElem, n, outBytes, N, counter);
}

// The same for workgroups below 1024 lanes, with aes_fill_lds_staged.
template <typename T>
__global__ __launch_bounds__(kBlock) void mic_eval_small_kernel(T *out, const T *x, const uint8_t *key, const T *z,
                                                                const uint8_t *seed, int party, const T *left,
                                                                const T *right, size_t I, size_t blocksPerElem, int n,
                                                                int outBytes, size_t N, uint32_t *counter)
{
    FSS_KERNEL_PROLOGUE_STAGED
    mic_eval_body<T>(c, out, x, key, z, seed, party, left, right, I, blocksPerElem, n, outBytes, N, counter);
}

// mic_eval_range with each DCF walk on eight lanes (dcf_eval_g1_oct), for
// calls small enough that eight lanes per (element, block of intervals) still
// leave at most two waves per SIMD.  Lane 4 of each group of eight writes the
// outputs.  Same results as mic_eval_kernel.
template <typename T>
__global__ __launch_bounds__(kBlock) void mic_eval_oct_kernel(T *out, const T *x, const uint8_t *key, const T *z,
                                                              const uint8_t *seed, int party, const T *left,
                                                              const T *right, size_t I, size_t blocksPerElem, int n,
                                                              int outBytes, size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    const uint32_t q = threadIdx.x & 3u, h = (threadIdx.x >> 2) & 1u;
    const bool     writer = (threadIdx.x & 7u) == 4u;
    QuadAes        qa;
    qa.init(c, q);
    const T     *tails = (const T *)(key + N * 16 * (size_t)n);
    const size_t span  = (I + blocksPerElem - 1) / blocksPerElem;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; (g >> 3) < N * blocksPerElem;
         g += (size_t)gridDim.x * blockDim.x)
    {
        const size_t w = g >> 3, e = w / blocksPerElem, first = (w % blocksPerElem) * span;
        if (first >= I) continue;
        const size_t   cnt = (first + span < I) ? span : (I - first);
        const uint32_t s   = *(const uint32_t *)(seed + 16 * e + 4 * q);
        const uint8_t *sc  = key + e * 16 * (size_t)n;
        const T       *tl  = tails + e * (size_t)(n + 3);
        const T        xe  = mod_bits<T>(x[e], n);
        T              sp = 0, sq = 0;
        for (size_t i = first; i < first + cnt; i++)
        {
            const T qP  = mod_bits<T>((T)(right[i] + 1), n);
            const T xP  = mod_bits<T>((T)(xe - 1 - left[i]), n);
            const T xQp = mod_bits<T>((T)(xe - 1 - (T)(right[i] + 1)), n);
            const T p   = mod_bits<T>(left[i], n);
            if (i > first && left[i] == mod_bits<T>((T)(right[i - 1] + 1), n))
                sp = sq;
            else
                sp = dcf_eval_g1_oct<T>(c, qa, sc, tl, xP, s, party, n, outBytes, h, q);
            sq  = dcf_eval_g1_oct<T>(c, qa, sc, tl, xQp, s, party, n, outBytes, h, q);
            T o = (T)(sq - sp + z[e * I + i]);
            if (party == 1) o = (T)(o + (T)((int)(xe > p) - (int)(xe > qP)));
            if (writer) out[e * I + i] = o;
        }
    }
}

// eight lanes per work item while they leave at most two waves per SIMD
static bool mic_oct_lanes(size_t items)
{
    return 8 * items <= (size_t)device_cu_count() * 512;
}

static size_t mic_blocks_per_elem(size_t N, size_t I)
{
    // split each element's intervals until the launch has a lane per CU slot
    const size_t lanes = (size_t)device_cu_count() * kBlock;
    size_t       bpe   = (lanes + N - 1) / N;
    if (bpe < 1) bpe = 1;
    if (bpe > I) bpe = I;
    return bpe;
}

int launch_mic_eval_rt(size_t es, int grid, hipStream_t st, void *out, const void *x, const void *key, const void *z,
                       const void *seed, int party, const void *left, const void *right, size_t I, int n, int m,
                       size_t N, uint32_t *counter, int block)
{
    const size_t bpe = mic_blocks_per_elem(N, I);
    // (block < kBlock: the per-call shape of a small call; the serving
    // runtime's one-workgroup-per-1024-elements shape keeps its launch)
    if (!counter && block < kBlock && mic_oct_lanes(N * bpe))
    {
        // its own launch shape: eight lanes per work item
        const LaunchShape os = shape_for(8 * N * bpe);
        FSS_DISPATCH(es, mic_eval_oct_kernel<T><<<os.grid, os.block, 0, st>>>(
                             (T *)out, (const T *)x, (const uint8_t *)key, (const T *)z, (const uint8_t *)seed, party,
                             (const T *)left, (const T *)right, I, bpe, n, (m + 7) / 8, N));
    }
    else if (block < kBlock)
    {
        FSS_DISPATCH(es, mic_eval_small_kernel<T><<<grid, block, 0, st>>>(
                             (T *)out, (const T *)x, (const uint8_t *)key, (const T *)z, (const uint8_t *)seed, party,
                             (const T *)left, (const T *)right, I, bpe, n, (m + 7) / 8, N, counter));
    }
    else
    {
        FSS_DISPATCH(es, mic_eval_kernel<T><<<grid, block, 0, st>>>(
                             (T *)out, (const T *)x, (const uint8_t *)key, (const T *)z, (const uint8_t *)seed, party,
                             (const T *)left, (const T *)right, I, bpe, n, (m + 7) / 8, N, counter));
    }
    return launch_status();
}

} // namespace fss

using namespace fss;

extern "C" FAST_FSS_API int FastFss_cuda_dcfMICKeyGen(void *key, size_t keyDataSize, void *z, size_t zDataSize,
                                                      const void *alpha, size_t alphaDataSize, const void *seed0,
                                                      size_t seedDataSize0, const void *seed1, size_t seedDataSize1,
                                                      const void *leftEndpoints, size_t leftEndpointsDataSize,
                                                      const void *rightEndpoints, size_t rightEndpointsDataSize,
                                                      size_t bitWidthIn, size_t bitWidthOut, size_t elementSize,
                                                      size_t elementNum, void *cudaStreamPtr)
{
    const size_t n = bitWidthIn, m = bitWidthOut, es = elementSize, N = elementNum;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    const size_t I = leftEndpointsDataSize / es;
    if (!(0 < n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!(0 < m && m <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (keyDataSize != dcf_key_size(n, 1, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (zDataSize != N * es * I) return FAST_FSS_INVALID_Z_DATA_SIZE_ERROR;
    if (alphaDataSize != N * es) return FAST_FSS_INVALID_ALPHA_DATA_SIZE_ERROR;
    if (seedDataSize0 != N * 16 || seedDataSize1 != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    if (leftEndpointsDataSize != I * es || rightEndpointsDataSize != I * es)
        return FAST_FSS_INVALID_BOUNDARY_DATA_SIZE_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed0, 8) || !aligned(seed1, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st = resolve_stream(cudaStreamPtr);
    if (es > 8)
    {
        if (!aligned(z, 8) || !aligned(alpha, 8) || !aligned(leftEndpoints, 8) || !aligned(rightEndpoints, 8))
            return FAST_FSS_RUNTIME_ERROR;
        return wide_mic_keygen(st, key, z, alpha, seed0, seed1, leftEndpoints, rightEndpoints, I, (int)n, (int)m, es,
                               N);
    }
    uint32_t *counter = queue_counter_acquire(st);
    if (!counter) return FAST_FSS_RUNTIME_ERROR;
    FSS_DISPATCH(es, mic_keygen_kernel<T><<<FSS_SHAPE(N), 0, st>>>(
                         (uint8_t *)key, (T *)z, (const T *)alpha, (const uint8_t *)seed0, (const uint8_t *)seed1,
                         (const T *)leftEndpoints, (const T *)rightEndpoints, I, (int)n, (int)m, N, counter));
    queue_counter_release(st, counter);
    return launch_status();
}

extern "C" FAST_FSS_API int FastFss_cuda_dcfMICEval(void *sharedOut, size_t sharedOutDataSize, const void *maskedX,
                                                    size_t maskedXDataSize, const void *key, size_t keyDataSize,
                                                    const void *sharedZ, size_t sharedZDataSize, const void *seed,
                                                    size_t seedDataSize, int partyId, const void *leftEndpoints,
                                                    size_t leftEndpointsDataSize, const void *rightEndpoints,
                                                    size_t rightEndpointsDataSize, size_t bitWidthIn,
                                                    size_t bitWidthOut, size_t elementSize, size_t elementNum,
                                                    void *cache, size_t cacheDataSize, void *cudaStreamPtr)
{
    (void)cache;
    const size_t n = bitWidthIn, m = bitWidthOut, es = elementSize, N = elementNum;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    const size_t I = leftEndpointsDataSize / es;
    if (!(0 < n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!(0 < m && m <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (partyId != 0 && partyId != 1) return FAST_FSS_INVALID_PARTY_ID_ERROR;
    if (keyDataSize != dcf_key_size(n, 1, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (sharedOutDataSize != N * es * I) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (maskedXDataSize != N * es) return FAST_FSS_INVALID_MASKED_X_DATA_SIZE_ERROR;
    if (sharedZDataSize != N * es * I) return FAST_FSS_INVALID_Z_DATA_SIZE_ERROR;
    if (seedDataSize != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    if (leftEndpointsDataSize != I * es || rightEndpointsDataSize != I * es)
        return FAST_FSS_INVALID_BOUNDARY_DATA_SIZE_ERROR;
    if (cacheDataSize != 0 && cacheDataSize != dcf_cache_size(n, 1, es, N)) return FAST_FSS_INVALID_CACHE_DATA_SIZE_ERROR;
    if (N == 0 || I == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st = resolve_stream(cudaStreamPtr);
    if (es > 8)
    {
        if (!aligned(sharedOut, 8) || !aligned(maskedX, 8) || !aligned(sharedZ, 8) || !aligned(leftEndpoints, 8) ||
            !aligned(rightEndpoints, 8))
            return FAST_FSS_RUNTIME_ERROR;
        return wide_mic_eval(st, sharedOut, maskedX, key, sharedZ, seed, partyId, leftEndpoints, rightEndpoints, I,
                             (int)n, (int)m, es, N);
    }
    const size_t bpe     = mic_blocks_per_elem(N, I);
    const LaunchShape ls = shape_for(N * bpe);
    // the static split below kMicQueueMinElems (2^14 elements, n = 24: 282
    // against 336 us), the work queue from there (2^22: 32.3 against 36.7 ms)
    if (N < kMicQueueMinElems)
        return launch_mic_eval_rt(es, ls.grid, st, sharedOut, maskedX, key, sharedZ, seed, partyId, leftEndpoints,
                                  rightEndpoints, I, (int)n, (int)m, N, nullptr, ls.block);
    uint32_t *counter = queue_counter_acquire(st);
    if (!counter) return FAST_FSS_RUNTIME_ERROR;
    const int rc = launch_mic_eval_rt(es, ls.grid, st, sharedOut, maskedX, key, sharedZ, seed, partyId, leftEndpoints,
                                      rightEndpoints, I, (int)n, (int)m, N, counter, ls.block);
    queue_counter_release(st, counter);
    return rc;
}
