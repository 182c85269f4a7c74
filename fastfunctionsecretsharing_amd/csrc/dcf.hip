// FastFss_cuda_dcfKeyGen / FastFss_cuda_dcfEval on MI355X.
//
// Launch shape: persistent grid of one 1024-thread workgroup per CU (the
// 128 KB AES tables fill most of the CU's LDS), one element per lane per
// iteration of a grid-stride loop.  Argument validation follows the
// reference Task::check order (src/kernel/dcf.h:31-69, 115-163) so callers see
// the same error codes.
#include "FastFss/cuda/dcf.h"
#include "dcf_device.h"
#include "fss_host.h"

namespace fss {

template <typename T>
__global__ __launch_bounds__(kBlock) void dcf_keygen_kernel(uint8_t *key, const T *alpha, const T *beta,
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
        dcf_keygen_one<T>(c, key + e * 16 * (size_t)n, (T *)tails + e * stride, alpha[e],
                          beta ? beta + e * (size_t)g : nullptr, a, b, n, m, g);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void dcf_eval_g1_kernel(T *out, const T *x, const uint8_t *key,
                                                             const uint8_t *seed, int party, int n, int outBytes,
                                                             size_t N)
{
    __shared__ uint32_t lds[kLdsTableWords];
    aes_fill_lds(lds);
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

template <typename T>
__global__ __launch_bounds__(kBlock) void dcf_eval_group_kernel(T *out, const T *x, const uint8_t *key,
                                                                const uint8_t *seed, int party, int n, int m, int g,
                                                                size_t N)
{
    __shared__ uint32_t lds[kLdsTableWords];
    aes_fill_lds(lds);
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

} // namespace fss

using namespace fss;

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
    FSS_DISPATCH(es, dcf_keygen_kernel<T><<<grid_for(N), kBlock, 0, st>>>(
                         (uint8_t *)key, (const T *)alpha, (const T *)b, (const uint8_t *)seed0,
                         (const uint8_t *)seed1, (int)n, (int)m, (int)g, N));
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
    hipStream_t st = resolve_stream(cudaStreamPtr);
    if (g == 1)
    {
        FSS_DISPATCH(es, dcf_eval_g1_kernel<T><<<grid_for(N), kBlock, 0, st>>>(
                             (T *)sharedOut, (const T *)maskedX, (const uint8_t *)key, (const uint8_t *)seed, partyId,
                             (int)n, (int)((m + 7) / 8), N));
    }
    else
    {
        FSS_DISPATCH(es, dcf_eval_group_kernel<T><<<grid_for(N), kBlock, 0, st>>>(
                             (T *)sharedOut, (const T *)maskedX, (const uint8_t *)key, (const uint8_t *)seed, partyId,
                             (int)n, (int)m, (int)g, N));
    }
    return launch_status();
}
