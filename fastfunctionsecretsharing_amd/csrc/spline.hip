// DCF-spline (piecewise polynomial) keyGen and eval on MI355X (reference
// src/impl/spline.h, src/kernel/spline.h).  The degree <= 7 eval kernels are in
// spline_device.h and launched from spline_eval.hip; any other degree keeps the
// reference's two group vectors in the caller's cache (spline_eval_wide_kernel).
#include "FastFss/cuda/spline.h"
#include "dcf_device.h"
#include "fss_host.h"
#include "fss_launch.h"

namespace fss {

// --------------------------------------------------------------- spline ----

// coefficients of f(x - shift) mod 2^m (src/impl/spline.h:89-117)
template <typename T>
__device__ void spline_shift(T *outc, const T *coef, T shift, int d, int m)
{
    const T neg = mod_bits<T>((T)((T)0 - shift), m);
    for (int i = 0; i <= d; i++) outc[i] = 0;
    for (int i = d + 1; i > 0; --i)
    {
        for (int j = d; j > 0; --j) outc[j] = (T)(outc[j - 1] + (T)(outc[j] * neg));
        outc[0] = (T)((T)(outc[0] * neg) + coef[i - 1]);
    }
    for (int i = 0; i <= d; i++) outc[i] = mod_bits<T>(outc[i], m);
}

template <typename T, int OB>
__global__ __launch_bounds__(group_keygen_block(sizeof(T), OB)) void spline_keygen_kernel(uint8_t *key, T *eOut, T *betaOut, const T *alpha,
                                                               const uint8_t *seed0, const uint8_t *seed1,
                                                               const T *coef, int d, const T *left, const T *right,
                                                               size_t I, int n, int m, size_t N)
{
    FSS_KERNEL_PROLOGUE
    const size_t C = (size_t)d + 1, g = I * C;
    T           *tails = (T *)(key + N * 16 * (size_t)n);
    const T      maxv  = mod_bits<T>((T)~(T)0, n);
    FSS_GRID_LOOP(e, N)
    {
        const T al   = alpha[e];
        T      *beta = betaOut + e * g;
        for (size_t q = 0; q < I; q++) spline_shift<T>(beta + q * C, coef + q * C, al, d, m);
        uint32_t a[4], b[4];
        load16(a, seed0 + 16 * e);
        load16(b, seed1 + 16 * e);
        dcf_keygen_group<T, OB>(c, key + e * 16 * (size_t)n, tails + e * (size_t)(n * g + 2 + g), (T)(al - 1), beta,
                                a, b, n, m, (int)g);
        for (size_t q = 0; q < I; q++)
        {
            const T L = mod_bits<T>(left[q], n), R = mod_bits<T>(right[q], n);
            const T Rp = mod_bits<T>((T)(R + 1), n);
            const T aL = mod_bits<T>((T)(L + al), n), aR = mod_bits<T>((T)(R + al), n);
            const T aRp = mod_bits<T>((T)(Rp + al), n);
            const T cr  = (T)((int)(aL > aR) - (int)(aL > L) + (int)(aRp > Rp) + (int)(aR == maxv));
            for (size_t j = 0; j < C; j++) eOut[e * g + q * C + j] = mod_bits<T>((T)(cr * beta[q * C + j]), m);
        }
    }
}

// Any degree: the reference's per-element workspace inside the caller's
// cache holds the two group vectors (src/impl/spline.h:64-87).
template <typename T>
__global__ __launch_bounds__(kBlock) void spline_eval_wide_kernel(T *out, const T *x, const uint8_t *key, const T *sE,
                                                                  const T *sBeta, const uint8_t *seed, int party,
                                                                  const T *left, const T *right, size_t I, int d,
                                                                  int n, int m, size_t N, uint8_t *cache)
{
    FSS_KERNEL_PROLOGUE
    const size_t C = (size_t)d + 1, g = I * C;
    const T     *tails = (const T *)(key + N * 16 * (size_t)n);
    uint8_t     *ws    = cache + dcf_cache_size((size_t)n, g, sizeof(T), N);
    FSS_GRID_LOOP(e, N)
    {
        uint32_t s[4];
        load16(s, seed + 16 * e);
        const uint8_t *sc   = key + e * 16 * (size_t)n;
        const T       *tail = tails + e * (n * g + 2 + g);
        T             *cur  = (T *)(ws + e * sizeof(T) * (2 * g + C));
        T             *nxt  = cur + g;
        T             *coef = nxt + g;
        const T        xv   = mod_bits<T>(x[e], n);
        for (size_t j = 0; j < C; j++) coef[j] = 0;
        for (size_t q = 0; q < I; q++)
        {
            const T L  = mod_bits<T>(left[q], n), R = mod_bits<T>(right[q], n);
            const T Rp = mod_bits<T>((T)(R + 1), n);
            if (q > 0 && mod_bits<T>((T)(right[q - 1] + 1), n) == L)
                for (size_t j = 0; j < g; j++) cur[j] = nxt[j];
            else
                dcf_eval_global<T>(c, cur, sc, tail, (T)(xv - 1 - L), s, party, n, m, (int)g);
            dcf_eval_global<T>(c, nxt, sc, tail, (T)(xv - 1 - Rp), s, party, n, m, (int)g);
            const T cx = (T)((int)(xv > L) - (int)(xv > Rp));
            for (size_t j = 0; j < C; j++)
                coef[j] = (T)(coef[j] + (T)(cx * sBeta[e * g + q * C + j]) - cur[q * C + j] + nxt[q * C + j] +
                              sE[e * g + q * C + j]);
        }
        T r = 0, pw = 1;
        for (size_t j = 0; j < C; j++)
        {
            r  = (T)(r + (T)(coef[j] * pw));
            pw = (T)(pw * xv);
        }
        out[e] = mod_bits<T>(r, m);
    }
}

} // namespace fss

using namespace fss;

extern "C" FAST_FSS_API int FastFss_cuda_dcfSplineKeyGen(
    void *key, size_t keyDataSize, void *e, size_t eDataSize, void *beta, size_t betaDataSize, const void *alpha,
    size_t alphaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1, size_t seedDataSize1,
    const void *coefficients, size_t coefficientsDataSize, size_t degree, const void *leftEndpoints,
    size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize, size_t intervalNum,
    size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum, void *cudaStreamPtr)
{
    const size_t n = bitWidthIn, m = bitWidthOut, es = elementSize, N = elementNum, I = intervalNum;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    const size_t C = degree + 1, g = I * C;
    if (!(0 < n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!(0 < m && m <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (I == 0) return FAST_FSS_INVALID_BOUNDARY_DATA_SIZE_ERROR;
    if (keyDataSize != dcf_key_size(n, g, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (eDataSize != N * es * g) return FAST_FSS_INVALID_Z_DATA_SIZE_ERROR;
    if (betaDataSize != N * es * g) return FAST_FSS_INVALID_BETA_DATA_SIZE_ERROR;
    if (alphaDataSize != N * es) return FAST_FSS_INVALID_ALPHA_DATA_SIZE_ERROR;
    if (seedDataSize0 != N * 16 || seedDataSize1 != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    if (leftEndpointsDataSize != I * es) return FAST_FSS_INVALID_BOUNDARY_DATA_SIZE_ERROR;
    if (rightEndpointsDataSize != I * es) return FAST_FSS_INVALID_BOUNDARY_DATA_SIZE_ERROR;
    if (coefficientsDataSize != I * C * es) return FAST_FSS_INVALID_BETA_DATA_SIZE_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed0, 8) || !aligned(seed1, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st = resolve_stream(cudaStreamPtr);
    if (es > 8)
    {
        if (!aligned(e, 8) || !aligned(beta, 8) || !aligned(alpha, 8) || !aligned(coefficients, 8) ||
            !aligned(leftEndpoints, 8) || !aligned(rightEndpoints, 8))
            return FAST_FSS_RUNTIME_ERROR;
        return wide_spline_keygen(st, key, e, beta, alpha, seed0, seed1, coefficients, (int)degree, leftEndpoints,
                                  rightEndpoints, I, (int)n, (int)m, es, N);
    }
    const int ob = keygen_chunk_ob((int)m, es);
    FSS_DISPATCH(es, FSS_KEYGEN_OB_DISPATCH(
                         T, ob,
                         (spline_keygen_kernel<T, OBV><<<grid_for(N), group_keygen_block(sizeof(T), OBV), 0, st>>>(
                             (uint8_t *)key, (T *)e, (T *)beta, (const T *)alpha, (const uint8_t *)seed0,
                             (const uint8_t *)seed1, (const T *)coefficients, (int)degree, (const T *)leftEndpoints,
                             (const T *)rightEndpoints, I, (int)n, (int)m, N))));
    return launch_status();
}

extern "C" FAST_FSS_API int FastFss_cuda_dcfSplineEval(
    void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize, const void *key,
    size_t keyDataSize, const void *sharedE, size_t sharedEDataSize, const void *sharedBeta, size_t sharedBetaDataSize,
    const void *seed, size_t seedDataSize, int partyId, const void *leftEndpoints, size_t leftEndpointsDataSize,
    const void *rightEndpoints, size_t rightEndpointsDataSize, size_t intervalNum, size_t degree, size_t bitWidthIn,
    size_t bitWidthOut, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize, void *cudaStreamPtr)
{
    const size_t n = bitWidthIn, m = bitWidthOut, es = elementSize, N = elementNum, I = intervalNum;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    const size_t C = degree + 1, g = I * C;
    if (!(0 < n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!(0 < m && m <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (partyId != 0 && partyId != 1) return FAST_FSS_INVALID_PARTY_ID_ERROR;
    if (I == 0) return FAST_FSS_INVALID_BOUNDARY_DATA_SIZE_ERROR;
    if (keyDataSize != dcf_key_size(n, g, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (sharedOutDataSize != N * es) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (maskedXDataSize != N * es) return FAST_FSS_INVALID_MASKED_X_DATA_SIZE_ERROR;
    if (sharedEDataSize != N * es * g) return FAST_FSS_INVALID_Z_DATA_SIZE_ERROR;
    if (sharedBetaDataSize != N * es * g) return FAST_FSS_INVALID_BETA_DATA_SIZE_ERROR;
    if (seedDataSize != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    if (leftEndpointsDataSize != I * es) return FAST_FSS_INVALID_BOUNDARY_DATA_SIZE_ERROR;
    if (rightEndpointsDataSize != I * es) return FAST_FSS_INVALID_BOUNDARY_DATA_SIZE_ERROR;
    if (cache == nullptr || cacheDataSize != spline_cache_size(n, I, degree, es, N))
        return FAST_FSS_INVALID_CACHE_DATA_SIZE_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed, 8)) return FAST_FSS_RUNTIME_ERROR;
    hipStream_t st = resolve_stream(cudaStreamPtr);
    if (es > 8)
    {
        if (!aligned(sharedOut, 8) || !aligned(maskedX, 8) || !aligned(sharedE, 8) || !aligned(sharedBeta, 8) ||
            !aligned(leftEndpoints, 8) || !aligned(rightEndpoints, 8) || !aligned(cache, 8))
            return FAST_FSS_RUNTIME_ERROR;
        return wide_spline_eval(st, sharedOut, maskedX, key, sharedE, sharedBeta, seed, partyId, leftEndpoints,
                                rightEndpoints, I, (int)degree, (int)n, (int)m, es, N, cache);
    }
    if (degree <= 7)
    {
        uint32_t *counter = queue_counter_acquire(st);
        if (!counter) return FAST_FSS_RUNTIME_ERROR;
        const LaunchShape ls = shape_for(N * spline_blocks_per_elem(N, I));
        const int rc = launch_spline_eval_rt(es, ls.grid, st, sharedOut, maskedX, key, sharedE, sharedBeta, seed,
                                             partyId, leftEndpoints, rightEndpoints, I, (int)degree, (int)n, (int)m,
                                             N, counter, ls.block);
        queue_counter_release(st, counter);
        return rc;
    }
    else
    {
        FSS_DISPATCH(es, spline_eval_wide_kernel<T><<<FSS_SHAPE(N), 0, st>>>(
                             (T *)sharedOut, (const T *)maskedX, (const uint8_t *)key, (const T *)sharedE,
                             (const T *)sharedBeta, (const uint8_t *)seed, partyId, (const T *)leftEndpoints,
                             (const T *)rightEndpoints, I, (int)degree, (int)n, (int)m, N, (uint8_t *)cache));
    }
    return launch_status();
}
