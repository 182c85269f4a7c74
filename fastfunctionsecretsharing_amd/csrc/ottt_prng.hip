// One-time truth tables (reference src/impl/ottt.h, src/kernel/ottt.h) and
// the AES-128-CTR PRNG (reference src/cuda/prng.cu) on MI355X.
#include <string.h>

#include "FastFss/cuda/ottt.h"
#include "FastFss/cuda/prng.h"
#include "fss_device.h"
#include "fss_host.h"

namespace fss {

// ----------------------------------------------------------------- OTTT ----

// keyGen XORs bit alpha into the caller's (random) key (src/impl/ottt.h:15-24);
// each element owns distinct key bytes, so plain byte stores do not race.
template <typename T>
__global__ void ottt_keygen_kernel(uint8_t *key, const T *alpha, int n, size_t N)
{
    const size_t per = ((size_t)1 << n) / 8;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (size_t)gridDim.x * blockDim.x)
    {
        const size_t a = (size_t)mod_bits<T>(alpha[e], n);
        key[e * per + a / 8] ^= (uint8_t)(1u << (a % 8));
    }
}

// E = popcount(key), T_j = sum over set key bits b of LUT_j[(x - b) mod 2^n];
// party 0 negates both (src/impl/ottt.h:26-192).  One thread per element,
// walking the set bits of 32-bit key words; the shared LUT stays in L2.
template <typename T>
__global__ void ottt_eval_kernel(T *outE, T *outT, const T *x, const uint8_t *key, int party, const T *lut,
                                 size_t lutNum, int n, size_t N)
{
    const size_t total = (size_t)1 << n, per = total / 8, mask = total - 1;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (size_t)gridDim.x * blockDim.x)
    {
        const uint8_t *kb = key + e * per;
        const size_t   xv = (size_t)x[e];
        T              E  = 0;
        for (size_t j = 0; j < lutNum; j++) outT[e * lutNum + j] = 0;
        for (size_t w = 0; w < total; w += 32)
        {
            uint32_t bits;
            if (total >= 32)
                bits = *(const uint32_t *)(kb + w / 8);
            else
                bits = kb[0]; // n = 3, 4: one byte (n=4: two bytes below)
            if (total == 16) bits = (uint32_t)kb[0] | ((uint32_t)kb[1] << 8);
            E = (T)(E + (T)__popc(bits));
            while (bits)
            {
                const int    b   = __ffs(bits) - 1;
                const size_t idx = (xv - (w + (size_t)b)) & mask;
                bits &= bits - 1;
                for (size_t j = 0; j < lutNum; j++) outT[e * lutNum + j] = (T)(outT[e * lutNum + j] + lut[idx + j * total]);
            }
        }
        outE[e] = party == 0 ? (T)(0 - E) : E;
        if (party == 0)
            for (size_t j = 0; j < lutNum; j++) outT[e * lutNum + j] = (T)(0 - outT[e * lutNum + j]);
    }
}

// ----------------------------------------------------------------- PRNG ----

struct PrngState
{
    uint32_t seed[4];
    uint32_t counter[4]; // 128-bit little-endian
};

// Stream block k = AES_seed(counter + k); writes `bytes` bytes (the last block
// may be partial) and leaves the counter untouched (see prng_advance_kernel).
__global__ __launch_bounds__(kBlock) void prng_gen_kernel(const PrngState *st, uint8_t *dst, size_t bytes)
{
    __shared__ uint32_t lds[kLdsTableWords];
    aes_fill_lds(lds);
    __syncthreads();
    const AesCtx   c    = aes_ctx(lds);
    const uint32_t k[4] = {st->seed[0], st->seed[1], st->seed[2], st->seed[3]};
    const uint64_t c0   = (uint64_t)st->counter[0] | ((uint64_t)st->counter[1] << 32);
    const uint64_t c1   = (uint64_t)st->counter[2] | ((uint64_t)st->counter[3] << 32);
    const size_t   nblk = (bytes + 15) / 16;
    const bool     vec  = ((uintptr_t)dst % 16) == 0;
    // four consecutive blocks per lane share one on-the-fly key schedule
    for (size_t q = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; q < nblk;
         q += (size_t)gridDim.x * blockDim.x * 4)
    {
        uint32_t b[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
        {
            const uint64_t lo = c0 + q + i;
            const uint64_t hi = c1 + (lo < c0 ? 1 : 0);
            b[i][0] = (uint32_t)lo, b[i][1] = (uint32_t)(lo >> 32), b[i][2] = (uint32_t)hi, b[i][3] = (uint32_t)(hi >> 32);
        }
        aes_encrypt<4>(c, k, b);
#pragma unroll
        for (int i = 0; i < 4; i++)
        {
            const size_t blk = q + i;
            if (blk >= nblk) break;
            const size_t off = 16 * blk;
            if (vec && off + 16 <= bytes)
                *(uint4 *)(dst + off) = make_uint4(b[i][0], b[i][1], b[i][2], b[i][3]);
            else
                for (size_t t = 0; t < 16 && off + t < bytes; t++) dst[off + t] = (uint8_t)byte_of(b[i], (int)t);
        }
    }
}

__global__ void prng_advance_kernel(PrngState *st, size_t nblk)
{
    const uint64_t c0 = (uint64_t)st->counter[0] | ((uint64_t)st->counter[1] << 32);
    const uint64_t c1 = (uint64_t)st->counter[2] | ((uint64_t)st->counter[3] << 32);
    const uint64_t lo = c0 + nblk;
    const uint64_t hi = c1 + (lo < c0 ? 1 : 0);
    st->counter[0] = (uint32_t)lo, st->counter[1] = (uint32_t)(lo >> 32);
    st->counter[2] = (uint32_t)hi, st->counter[3] = (uint32_t)(hi >> 32);
}

} // namespace fss

using namespace fss;

extern "C" FAST_FSS_API int FastFss_cuda_otttKeyGen(void *key, size_t keyDataSize, const void *alpha,
                                                    size_t alphaDataSize, size_t bitWidthIn, size_t elementSize,
                                                    size_t elementNum, void *cudaStreamPtr)
{
    const size_t n = bitWidthIn, es = elementSize, N = elementNum;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(3 <= n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (n >= 48 || keyDataSize != N * ((size_t)1 << n) / 8) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (alphaDataSize != N * es) return FAST_FSS_INVALID_ALPHA_DATA_SIZE_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    hipStream_t st = resolve_stream(cudaStreamPtr);
    FSS_DISPATCH(es, ottt_keygen_kernel<T><<<FSS_SHAPE(N), 0, st>>>((uint8_t *)key, (const T *)alpha, (int)n,
                                                                            N));
    return launch_status();
}

extern "C" FAST_FSS_API int FastFss_cuda_otttLutEval(void *sharedOutE, size_t sharedOutEDataSize, void *sharedOutT,
                                                     size_t sharedOutTDataSize, const void *maskedX,
                                                     size_t maskedXDataSize, const void *key, size_t keyDataSize,
                                                     int partyId, const void *lookUpTable, size_t lookUpTableDataSize,
                                                     size_t bitWidthIn, size_t elementSize, size_t elementNum,
                                                     void *cudaStreamPtr)
{
    const size_t n = bitWidthIn, es = elementSize, N = elementNum;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(3 <= n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (partyId != 0 && partyId != 1) return FAST_FSS_INVALID_PARTY_ID_ERROR;
    if (n >= 48 || keyDataSize != N * ((size_t)1 << n) / 8) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    const size_t each = es * ((size_t)1 << n);
    // a malformed table reports the seed-size code, as the reference does (src/kernel/ottt.h:75-78)
    if (lookUpTableDataSize % each != 0) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    const size_t lutNum = lookUpTableDataSize / each;
    if (sharedOutEDataSize != N * es) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (sharedOutTDataSize != N * es * lutNum) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (maskedXDataSize != N * es) return FAST_FSS_INVALID_MASKED_X_DATA_SIZE_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (n >= 5 && !aligned(key, 4)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st = resolve_stream(cudaStreamPtr);
    FSS_DISPATCH(es, ottt_eval_kernel<T><<<grid_for(N), 256, 0, st>>>(
                         (T *)sharedOutE, (T *)sharedOutT, (const T *)maskedX, (const uint8_t *)key, partyId,
                         (const T *)lookUpTable, lutNum, (int)n, N));
    return launch_status();
}

extern "C" FAST_FSS_API void *FastFss_cuda_prngInit(void)
{
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(PrngState)) != hipSuccess) return nullptr;
    if (hipMemset(p, 0, sizeof(PrngState)) != hipSuccess)
    {
        (void)hipFree(p);
        return nullptr;
    }
    return p;
}

extern "C" FAST_FSS_API void FastFss_cuda_prngRelease(void *prng)
{
    if (prng) (void)hipFree(prng);
}

extern "C" FAST_FSS_API int FastFss_cuda_prngSetCurrentSeed(void *prng, const void *hostSeed128bit,
                                                            const void *hostCounter128bit)
{
    if (prng == nullptr || hostSeed128bit == nullptr) return FAST_FSS_PRNG_INPUT_INVALID_ARGUMENT;
    PrngState s;
    memcpy(s.seed, hostSeed128bit, 16);
    if (hostCounter128bit)
        memcpy(s.counter, hostCounter128bit, 16);
    else
        memset(s.counter, 0, 16);
    if (hipMemcpy(prng, &s, sizeof(s), hipMemcpyHostToDevice) != hipSuccess) return FAST_FSS_RUNTIME_ERROR;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_cuda_prngGetCurrentSeed(const void *prng, void *hostSeed128bit,
                                                            void *hostCounter128bit)
{
    if (prng == nullptr || hostSeed128bit == nullptr) return FAST_FSS_PRNG_INPUT_INVALID_ARGUMENT;
    PrngState s;
    if (hipMemcpy(&s, prng, sizeof(s), hipMemcpyDeviceToHost) != hipSuccess) return FAST_FSS_RUNTIME_ERROR;
    memcpy(hostSeed128bit, s.seed, 16);
    if (hostCounter128bit) memcpy(hostCounter128bit, s.counter, 16);
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_cuda_prngGen(void *prng, void *deviceDst, size_t bitWidth, size_t elementSize,
                                                 size_t elementNum, void *cudaStreamPtr)
{
    if (prng == nullptr || deviceDst == nullptr) return FAST_FSS_PRNG_INPUT_INVALID_ARGUMENT;
    if (bitWidth == 0 || bitWidth > elementSize * 8) return FAST_FSS_PRNG_INVALID_BIT_WIDTH;
    const size_t bytes = elementSize * elementNum;
    const size_t nblk  = (bytes + 15) / 16;
    hipStream_t  st    = resolve_stream(cudaStreamPtr);
    if (nblk)
    {
        prng_gen_kernel<<<FSS_SHAPE((nblk + 3) / 4), 0, st>>>((const PrngState *)prng, (uint8_t *)deviceDst,
                                                                     bytes);
        if (hipPeekAtLastError() != hipSuccess) return FAST_FSS_RUNTIME_ERROR;
    }
    prng_advance_kernel<<<1, 1, 0, st>>>((PrngState *)prng, nblk);
    return launch_status();
}
