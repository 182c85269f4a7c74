// FastFss_cuda_dpf* and FastFss_cuda_grotto* on MI355X.
//
// Validation follows the reference Task::check order (src/kernel/dpf.h,
// src/kernel/grotto.h).  Every kernel uses the persistent 1024-thread
// workgroup-per-CU shape of the AES tables and a grid-stride loop over work
// items (elements, or element x point pairs for the multi-point forms).
#include "FastFss/cuda/dpf.h"
#include "FastFss/cuda/grotto.h"
#include "dpf_device.h"
#include "fss_host.h"

namespace fss {

#define FSS_KERNEL_PROLOGUE                 \
    __shared__ uint32_t lds[kLdsTableWords]; \
    aes_fill_lds(lds);                      \
    __syncthreads();                        \
    const AesCtx c = aes_ctx(lds);

#define FSS_GRID_LOOP(i, count)                                                             \
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (count); \
         i += (size_t)gridDim.x * blockDim.x)

// ------------------------------------------------------------------ DPF ----

template <typename T>
__global__ __launch_bounds__(kBlock) void dpf_keygen_kernel(uint8_t *key, const T *alpha, const T *beta,
                                                            const uint8_t *seed0, const uint8_t *seed1, int n, int m,
                                                            int g, size_t N)
{
    FSS_KERNEL_PROLOGUE
    T *tails = (T *)(key + N * 16 * (size_t)n);
    FSS_GRID_LOOP(e, N)
    {
        uint32_t a[4], b[4];
        load16(a, seed0 + 16 * e);
        load16(b, seed1 + 16 * e);
        dpf_keygen_one<T>(c, key + e * 16 * (size_t)n, tails + e * (size_t)(2 + g), alpha[e],
                          beta ? beta + e * (size_t)g : nullptr, a, b, n, m, g);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void dpf_eval_kernel(T *out, const T *x, const uint8_t *key, const uint8_t *seed,
                                                          int party, int n, int m, int g, size_t N)
{
    FSS_KERNEL_PROLOGUE
    const T *tails = (const T *)(key + N * 16 * (size_t)n);
    FSS_GRID_LOOP(e, N)
    {
        uint32_t s[4];
        load16(s, seed + 16 * e);
        dpf_eval_one<T>(c, out + e * (size_t)g, key + e * 16 * (size_t)n, tails + e * (size_t)(2 + g), x[e], s,
                        party, n, m, g);
    }
}

// One work item per (element, point): point j of element e lands at
// out[e][(x_e - j) mod 2^n] (src/kernel/dpf.h:253-276).
template <typename T>
__global__ __launch_bounds__(kBlock) void dpf_eval_all_kernel(T *out, const T *x, const uint8_t *key,
                                                              const uint8_t *seed, int party, int n, int m, int g,
                                                              size_t N)
{
    FSS_KERNEL_PROLOGUE
    const T     *tails = (const T *)(key + N * 16 * (size_t)n);
    const size_t size  = (size_t)1 << n;
    FSS_GRID_LOOP(w, N * size)
    {
        const size_t e = w >> n, j = w & (size - 1);
        uint32_t     s[4];
        load16(s, seed + 16 * e);
        const size_t idx = ((size_t)x[e] - j) & (size - 1);
        dpf_eval_one<T>(c, out + (e * size + idx) * g, key + e * 16 * (size_t)n, tails + e * (size_t)(2 + g), (T)j,
                        s, party, n, m, g);
    }
}

// One work item per (element, point): out[e][j] = Eval(x_e - point_j)
// (src/kernel/dpf.h:481-534).
template <typename T>
__global__ __launch_bounds__(kBlock) void dpf_eval_multi_kernel(T *out, const T *x, const uint8_t *key,
                                                                const uint8_t *seed, int party, const T *points,
                                                                size_t P, int n, int m, int g, size_t N)
{
    FSS_KERNEL_PROLOGUE
    const T *tails = (const T *)(key + N * 16 * (size_t)n);
    FSS_GRID_LOOP(w, N * P)
    {
        const size_t e = w / P, j = w % P;
        uint32_t     s[4];
        load16(s, seed + 16 * e);
        dpf_eval_one<T>(c, out + (e * P + j) * g, key + e * 16 * (size_t)n, tails + e * (size_t)(2 + g),
                        (T)(x[e] - points[j]), s, party, n, m, g);
    }
}

// --------------------------------------------------------------- Grotto ----

struct GrottoKeyPtrs
{
    const uint8_t  *sCW;
    const uint64_t *lastCW;
    const uint8_t  *tlr;
};

__device__ __forceinline__ GrottoKeyPtrs grotto_key(const uint8_t *key, int n, size_t es, size_t N)
{
    GrottoKeyPtrs k;
    k.sCW    = key;
    k.lastCW = (const uint64_t *)(key + N * 16 * (size_t)(n - 6));
    k.tlr    = key + N * 16 * (size_t)(n - 6) + N * 8;
    return k;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void grotto_keygen_kernel(uint8_t *key, const T *alpha, const uint8_t *seed0,
                                                               const uint8_t *seed1, int n, size_t N)
{
    FSS_KERNEL_PROLOGUE
    uint64_t *last = (uint64_t *)(key + N * 16 * (size_t)(n - 6));
    T        *tlr  = (T *)(key + N * 16 * (size_t)(n - 6) + N * 8);
    FSS_GRID_LOOP(e, N)
    {
        uint32_t a[4], b[4];
        load16(a, seed0 + 16 * e);
        load16(b, seed1 + 16 * e);
        grotto_keygen_one<T>(c, key + e * 16 * (size_t)(n - 6), last + e, tlr + 2 * e, alpha[e], a, b, n);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void grotto_eval_kernel(T *out, const T *x, const uint8_t *key,
                                                             const uint8_t *seed, int party, int n, int mode,
                                                             size_t N)
{
    // mode 0: [x < alpha] share, 1: [x <= alpha] share, 2: [x == alpha] share
    FSS_KERNEL_PROLOGUE
    const GrottoKeyPtrs k   = grotto_key(key, n, sizeof(T), N);
    const T            *tlr = (const T *)k.tlr;
    FSS_GRID_LOOP(e, N)
    {
        uint32_t s[4];
        load16(s, seed + 16 * e);
        const uint8_t *sc = k.sCW + e * 16 * (size_t)(n - 6);
        if (mode == 2)
            out[e] = grotto_eq_eval_one<T>(c, sc, k.lastCW[e], tlr[2 * e], tlr[2 * e + 1], x[e], s, party, n);
        else
            out[e] = grotto_eval_one<T>(c, sc, k.lastCW[e], tlr[2 * e], tlr[2 * e + 1], x[e], s, party, n, mode == 1);
    }
}

// Interval shares of one element (src/impl/grotto.h:393-447); writes `I`
// boolean shares and, for the LUT form, also accumulates E and T.
template <typename T>
__device__ __forceinline__ void grotto_mic_one(const AesCtx &c, T *share, const uint8_t *sc, uint64_t lastCW, T tl,
                                               T tr, T x, const uint32_t s[4], int party, const T *left,
                                               const T *right, size_t I, int n)
{
    T sp = 0, sq = 0;
    for (size_t i = 0; i < I; i++)
    {
        const T L = left[i], R = right[i];
        const T xP = mod_bits<T>((T)(x - (T)(L - 1)), n);
        const T xQ = mod_bits<T>((T)(x - R), n);
        if (i > 0 && L == mod_bits<T>((T)(right[i - 1] + 1), n))
            sp = sq;
        else
            sp = grotto_eval_one<T>(c, sc, lastCW, tl, tr, xP, s, party, n, true);
        sq       = grotto_eval_one<T>(c, sc, lastCW, tl, tr, xQ, s, party, n, true);
        share[i] = (T)((sp ^ sq) ^ ((xQ >= xP) ? (T)party : (T)0));
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void grotto_mic_kernel(T *out, const T *x, const uint8_t *key, const uint8_t *seed,
                                                            int party, const T *left, const T *right, size_t I, int n,
                                                            size_t N)
{
    FSS_KERNEL_PROLOGUE
    const GrottoKeyPtrs k   = grotto_key(key, n, sizeof(T), N);
    const T            *tlr = (const T *)k.tlr;
    FSS_GRID_LOOP(e, N)
    {
        uint32_t s[4];
        load16(s, seed + 16 * e);
        grotto_mic_one<T>(c, out + e * I, k.sCW + e * 16 * (size_t)(n - 6), k.lastCW[e], tlr[2 * e], tlr[2 * e + 1],
                          x[e], s, party, left, right, I, n);
    }
}

// src/impl/grotto.h:449-545: E = sum of interval bits, T_j = sum bit_i*LUT[i + j*I],
// both negated for party 1.  `scratch` holds the I bits of each element.
template <typename T>
__global__ __launch_bounds__(kBlock) void grotto_lut_kernel(T *outE, T *outT, T *bits, const T *x, const uint8_t *key,
                                                            const uint8_t *seed, int party, const T *left,
                                                            const T *right, const T *lut, size_t lutNum, size_t I,
                                                            int n, size_t N)
{
    FSS_KERNEL_PROLOGUE
    const GrottoKeyPtrs k   = grotto_key(key, n, sizeof(T), N);
    const T            *tlr = (const T *)k.tlr;
    FSS_GRID_LOOP(e, N)
    {
        uint32_t s[4];
        load16(s, seed + 16 * e);
        T *b = bits + e * I;
        grotto_mic_one<T>(c, b, k.sCW + e * 16 * (size_t)(n - 6), k.lastCW[e], tlr[2 * e], tlr[2 * e + 1], x[e], s,
                          party, left, right, I, n);
        T E = 0;
        for (size_t i = 0; i < I; i++) E = (T)(E + (T)(b[i] & 1));
        outE[e] = party ? (T)(0 - E) : E;
        for (size_t j = 0; j < lutNum; j++)
        {
            T acc = 0;
            for (size_t i = 0; i < I; i++) acc = (T)(acc + (T)((b[i] & 1) * lut[i + j * I]));
            outT[e * lutNum + j] = party ? (T)(0 - acc) : acc;
        }
    }
}

} // namespace fss

using namespace fss;

// ------------------------------------------------------------- C ABI: DPF ----

extern "C" FAST_FSS_API int FastFss_cuda_dpfKeyGen(void *key, size_t keyDataSize, const void *alpha,
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
    if (g == 0) return FAST_FSS_INVALID_GROUP_SIZE_ERROR;
    if (keyDataSize != dpf_key_size(n, g, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (alphaDataSize != N * es) return FAST_FSS_INVALID_ALPHA_DATA_SIZE_ERROR;
    if (betaDataSize != 0)
    {
        if (betaDataSize != N * es * g) return FAST_FSS_INVALID_BETA_DATA_SIZE_ERROR;
    }
    else if (g != 1)
        return FAST_FSS_INVALID_BETA_DATA_SIZE_ERROR;
    if (seedDataSize0 != N * 16 || seedDataSize1 != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed0, 8) || !aligned(seed1, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st = resolve_stream(cudaStreamPtr);
    const void *b  = betaDataSize ? beta : nullptr;
    FSS_DISPATCH(es, dpf_keygen_kernel<T><<<grid_for(N), kBlock, 0, st>>>((uint8_t *)key, (const T *)alpha,
                                                                           (const T *)b, (const uint8_t *)seed0,
                                                                           (const uint8_t *)seed1, (int)n, (int)m,
                                                                           (int)g, N));
    return launch_status();
}

static int dpf_eval_head(int party, size_t n, size_t m, size_t g, size_t es)
{
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(0 < n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!(0 < m && m <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (g == 0) return FAST_FSS_INVALID_GROUP_SIZE_ERROR;
    if (party != 0 && party != 1) return FAST_FSS_INVALID_PARTY_ID_ERROR;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_cuda_dpfEval(void *sharedOut, size_t sharedOutDataSize, const void *maskedX,
                                                 size_t maskedXDataSize, const void *key, size_t keyDataSize,
                                                 const void *seed, size_t seedDataSize, int partyId,
                                                 size_t bitWidthIn, size_t bitWidthOut, size_t groupSize,
                                                 size_t elementSize, size_t elementNum, void *cache,
                                                 size_t cacheDataSize, void *cudaStreamPtr)
{
    (void)cache;
    const size_t n = bitWidthIn, m = bitWidthOut, g = groupSize, es = elementSize, N = elementNum;
    int          r = dpf_eval_head(partyId, n, m, g, es);
    if (r) return r;
    if (keyDataSize != dpf_key_size(n, g, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (seedDataSize != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    if (maskedXDataSize != N * es) return FAST_FSS_INVALID_MASKED_X_DATA_SIZE_ERROR;
    if (sharedOutDataSize != N * es * g) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (cacheDataSize != 0 && cacheDataSize != dpf_cache_size(n, N)) return FAST_FSS_INVALID_CACHE_DATA_SIZE_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st = resolve_stream(cudaStreamPtr);
    FSS_DISPATCH(es, dpf_eval_kernel<T><<<grid_for(N), kBlock, 0, st>>>((T *)sharedOut, (const T *)maskedX,
                                                                         (const uint8_t *)key, (const uint8_t *)seed,
                                                                         partyId, (int)n, (int)m, (int)g, N));
    return launch_status();
}

extern "C" FAST_FSS_API int FastFss_cuda_dpfEvalAll(void *sharedOut, size_t sharedOutDataSize, const void *maskedX,
                                                    size_t maskedXDataSize, const void *key, size_t keyDataSize,
                                                    const void *seed, size_t seedDataSize, int partyId,
                                                    size_t bitWidthIn, size_t bitWidthOut, size_t groupSize,
                                                    size_t elementSize, size_t elementNum, void *cache,
                                                    size_t cacheDataSize, void *cudaStreamPtr)
{
    (void)cache;
    const size_t n = bitWidthIn, m = bitWidthOut, g = groupSize, es = elementSize, N = elementNum;
    int          r = dpf_eval_head(partyId, n, m, g, es);
    if (r) return r;
    if (keyDataSize != dpf_key_size(n, g, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (n >= 48 || sharedOutDataSize != N * es * ((size_t)1 << n) * g)
        return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (maskedXDataSize != N * es) return FAST_FSS_INVALID_MASKED_X_DATA_SIZE_ERROR;
    if (seedDataSize != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    if (cacheDataSize != 0 && cacheDataSize != dpf_cache_size(n, N)) return FAST_FSS_INVALID_CACHE_DATA_SIZE_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t  st   = resolve_stream(cudaStreamPtr);
    const size_t work = N << n;
    FSS_DISPATCH(es, dpf_eval_all_kernel<T><<<grid_for(work), kBlock, 0, st>>>(
                         (T *)sharedOut, (const T *)maskedX, (const uint8_t *)key, (const uint8_t *)seed, partyId,
                         (int)n, (int)m, (int)g, N));
    return launch_status();
}

extern "C" FAST_FSS_API int FastFss_cuda_dpfEvalMulti(void *sharedOut, size_t sharedOutDataSize, const void *maskedX,
                                                      size_t maskedXDataSize, const void *key, size_t keyDataSize,
                                                      const void *seed, size_t seedDataSize, int partyId,
                                                      const void *point, size_t pointDataSize, size_t bitWidthIn,
                                                      size_t bitWidthOut, size_t groupSize, size_t elementSize,
                                                      size_t elementNum, void *cache, size_t cacheDataSize,
                                                      void *cudaStreamPtr)
{
    (void)cache;
    const size_t n = bitWidthIn, m = bitWidthOut, g = groupSize, es = elementSize, N = elementNum;
    int          r = dpf_eval_head(partyId, n, m, g, es);
    if (r) return r;
    if (pointDataSize % es != 0) return FAST_FSS_INVALID_POINT_DATA_SIZE_ERROR;
    if (keyDataSize != dpf_key_size(n, g, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    const size_t P = pointDataSize / es;
    if (sharedOutDataSize != N * es * P * g) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (maskedXDataSize != N * es) return FAST_FSS_INVALID_MASKED_X_DATA_SIZE_ERROR;
    if (seedDataSize != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    if (cacheDataSize != 0 && cacheDataSize != dpf_cache_size(n, N)) return FAST_FSS_INVALID_CACHE_DATA_SIZE_ERROR;
    if (N == 0 || P == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st = resolve_stream(cudaStreamPtr);
    FSS_DISPATCH(es, dpf_eval_multi_kernel<T><<<grid_for(N * P), kBlock, 0, st>>>(
                         (T *)sharedOut, (const T *)maskedX, (const uint8_t *)key, (const uint8_t *)seed, partyId,
                         (const T *)point, P, (int)n, (int)m, (int)g, N));
    return launch_status();
}

// ---------------------------------------------------------- C ABI: Grotto ----

extern "C" FAST_FSS_API int FastFss_cuda_grottoKeyGen(void *key, size_t keyDataSize, const void *alpha,
                                                      size_t alphaDataSize, const void *seed0, size_t seedDataSize0,
                                                      const void *seed1, size_t seedDataSize1, size_t bitWidthIn,
                                                      size_t elementSize, size_t elementNum, void *cudaStreamPtr)
{
    const size_t n = bitWidthIn, es = elementSize, N = elementNum;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(6 <= n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (alphaDataSize != N * es) return FAST_FSS_INVALID_ALPHA_DATA_SIZE_ERROR;
    if (seedDataSize0 != N * 16 || seedDataSize1 != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    if (keyDataSize != grotto_key_size(n, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed0, 8) || !aligned(seed1, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st = resolve_stream(cudaStreamPtr);
    FSS_DISPATCH(es, grotto_keygen_kernel<T><<<grid_for(N), kBlock, 0, st>>>(
                         (uint8_t *)key, (const T *)alpha, (const uint8_t *)seed0, (const uint8_t *)seed1, (int)n, N));
    return launch_status();
}

static int grotto_tail_checks(int party, size_t n, size_t es, size_t N, size_t keyDataSize, size_t cacheDataSize)
{
    if (party != 0 && party != 1) return FAST_FSS_INVALID_PARTY_ID_ERROR;
    if (keyDataSize != grotto_key_size(n, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (cacheDataSize != 0 && cacheDataSize != grotto_cache_size(n, N)) return FAST_FSS_INVALID_CACHE_DATA_SIZE_ERROR;
    return FAST_FSS_SUCCESS;
}

static int grotto_eval_launch(void *out, size_t outSize, const void *x, size_t xSize, const void *key,
                              size_t keyDataSize, const void *seed, size_t seedSize, int mode, int party, size_t n,
                              size_t es, size_t N, size_t cacheDataSize, void *sp)
{
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(6 <= n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (outSize != N * es) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (xSize != N * es) return FAST_FSS_INVALID_MASKED_X_DATA_SIZE_ERROR;
    if (seedSize != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    int r = grotto_tail_checks(party, n, es, N, keyDataSize, cacheDataSize);
    if (r) return r;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st = resolve_stream(sp);
    FSS_DISPATCH(es, grotto_eval_kernel<T><<<grid_for(N), kBlock, 0, st>>>(
                         (T *)out, (const T *)x, (const uint8_t *)key, (const uint8_t *)seed, party, (int)n, mode, N));
    return launch_status();
}

extern "C" FAST_FSS_API int FastFss_cuda_grottoEqEval(void *sharedBooleanOut, size_t sharedOutDataSize,
                                                      const void *maskedX, size_t maskedXDataSize, const void *key,
                                                      size_t keyDataSize, const void *seed, size_t seedDataSize,
                                                      int partyId, size_t bitWidthIn, size_t elementSize,
                                                      size_t elementNum, void *cache, size_t cacheDataSize,
                                                      void *cudaStreamPtr)
{
    (void)cache;
    return grotto_eval_launch(sharedBooleanOut, sharedOutDataSize, maskedX, maskedXDataSize, key, keyDataSize, seed,
                              seedDataSize, 2, partyId, bitWidthIn, elementSize, elementNum, cacheDataSize,
                              cudaStreamPtr);
}

extern "C" FAST_FSS_API int FastFss_cuda_grottoEval(void *sharedBooleanOut, size_t sharedOutDataSize,
                                                    const void *maskedX, size_t maskedXDataSize, const void *key,
                                                    size_t keyDataSize, const void *seed, size_t seedDataSize,
                                                    bool equalBound, int partyId, size_t bitWidthIn,
                                                    size_t elementSize, size_t elementNum, void *cache,
                                                    size_t cacheDataSize, void *cudaStreamPtr)
{
    (void)cache;
    return grotto_eval_launch(sharedBooleanOut, sharedOutDataSize, maskedX, maskedXDataSize, key, keyDataSize, seed,
                              seedDataSize, equalBound ? 1 : 0, partyId, bitWidthIn, elementSize, elementNum,
                              cacheDataSize, cudaStreamPtr);
}

extern "C" FAST_FSS_API int FastFss_cuda_grottoMICEval(void *sharedBooleanOut, size_t sharedBooleanOutDataSize,
                                                       const void *maskedX, size_t maskedXDataSize, const void *key,
                                                       size_t keyDataSize, const void *seed, size_t seedDataSize,
                                                       int partyId, const void *leftEndpoints,
                                                       size_t leftEndpointsDataSize, const void *rightEndpoints,
                                                       size_t rightEndpointsDataSize, size_t bitWidthIn,
                                                       size_t elementSize, size_t elementNum, void *cache,
                                                       size_t cacheDataSize, void *cudaStreamPtr)
{
    (void)cache;
    const size_t n = bitWidthIn, es = elementSize, N = elementNum;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(6 <= n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    const size_t I = leftEndpointsDataSize / es;
    if (I * es != leftEndpointsDataSize || I * es != rightEndpointsDataSize)
        return FAST_FSS_INVALID_BOUNDARY_DATA_SIZE_ERROR;
    if (sharedBooleanOutDataSize != N * es * I) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (maskedXDataSize != N * es) return FAST_FSS_INVALID_MASKED_X_DATA_SIZE_ERROR;
    if (seedDataSize != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    int r = grotto_tail_checks(partyId, n, es, N, keyDataSize, cacheDataSize);
    if (r) return r;
    if (N == 0 || I == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st = resolve_stream(cudaStreamPtr);
    FSS_DISPATCH(es, grotto_mic_kernel<T><<<grid_for(N), kBlock, 0, st>>>(
                         (T *)sharedBooleanOut, (const T *)maskedX, (const uint8_t *)key, (const uint8_t *)seed,
                         partyId, (const T *)leftEndpoints, (const T *)rightEndpoints, I, (int)n, N));
    return launch_status();
}

extern "C" FAST_FSS_API int FastFss_cuda_grottoIntervalLutEval(
    void *sharedOutE, size_t sharedOutEDataSize, void *sharedOutT, size_t sharedOutTDataSize, const void *maskedX,
    size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId,
    const void *leftEndpoints, size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize,
    const void *lookUpTable, size_t lookUpTableDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize,
    size_t elementNum, void *cache, size_t cacheDataSize, void *cudaStreamPtr)
{
    (void)cache, (void)bitWidthOut;
    const size_t n = bitWidthIn, es = elementSize, N = elementNum;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(6 <= n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!(leftEndpointsDataSize == rightEndpointsDataSize && leftEndpointsDataSize % es == 0))
        return FAST_FSS_INVALID_BOUNDARY_DATA_SIZE_ERROR;
    const size_t I = leftEndpointsDataSize / es;
    if (I == 0 || lookUpTableDataSize % (es * I) != 0) return FAST_FSS_INVALID_LOOKUP_TABLE_DATA_SIZE_ERROR;
    const size_t lutNum = lookUpTableDataSize / (es * I);
    if (sharedOutEDataSize != N * es) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (sharedOutTDataSize != N * es * lutNum) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (maskedXDataSize != N * es) return FAST_FSS_INVALID_MASKED_X_DATA_SIZE_ERROR;
    if (seedDataSize != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    int r = grotto_tail_checks(partyId, n, es, N, keyDataSize, cacheDataSize);
    if (r) return r;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st   = resolve_stream(cudaStreamPtr);
    void       *bits = nullptr;
    if (hipMallocAsync(&bits, N * I * es, st) != hipSuccess) return FAST_FSS_RUNTIME_ERROR;
    FSS_DISPATCH(es, grotto_lut_kernel<T><<<grid_for(N), kBlock, 0, st>>>(
                         (T *)sharedOutE, (T *)sharedOutT, (T *)bits, (const T *)maskedX, (const uint8_t *)key,
                         (const uint8_t *)seed, partyId, (const T *)leftEndpoints, (const T *)rightEndpoints,
                         (const T *)lookUpTable, lutNum, I, (int)n, N));
    int rc = launch_status();
    hipFreeAsync(bits, st);
    return rc;
}
