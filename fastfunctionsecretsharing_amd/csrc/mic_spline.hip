// DCF-based multiple-interval containment (MIC) and piecewise-polynomial
// spline on MI355X (reference src/impl/mic.h, src/impl/spline.h,
// src/kernel/mic.h, src/kernel/spline.h).
//
// Both evaluate the element's single DCF key at several points.  The spline
// needs, at each point, only the coefficient slots of the interval(s) that
// point bounds, so each point is walked once and only those group elements
// (and the AES conversion blocks that hold them) are produced; the values are
// identical to the reference's full-group evaluation.
#include "FastFss/cuda/mic.h"
#include "FastFss/cuda/spline.h"
#include "dcf_device.h"
#include "fss_host.h"
#include "fss_launch.h"

namespace fss {

#define FSS_KERNEL_PROLOGUE                 \
    __shared__ uint32_t lds[kLdsTableWords]; \
    aes_fill_lds(lds);                      \
    __syncthreads();                        \
    const AesCtx c = aes_ctx(lds);

// the same for kernels launched with workgroups below 1024 lanes (aes_fill_lds_staged)
#define FSS_KERNEL_PROLOGUE_STAGED                         \
    __shared__ __align__(16) uint32_t lds[kLdsTableWords]; \
    aes_fill_lds_staged(lds);                              \
    __syncthreads();                                       \
    const AesCtx c = aes_ctx(lds);

#define FSS_GRID_LOOP(i, count)                                                             \
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (count); \
         i += (size_t)gridDim.x * blockDim.x)

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

// Fast path: degree <= 7.  Each evaluation point x-1-L_0, x-1-(R_q+1) is walked
// once; the walk at x-1-(R_q+1) also yields interval q+1's "current" slots when
// intervals q and q+1 are adjacent (the reference copies the whole vector,
// src/impl/spline.h:214-223).  OB > 0 selects the slot conversion specialised
// for ceil(bitWidthOut/8) = OB bytes (needs 2*(degree+1)*OB <= 32); OB = 0 the
// generic one.  An element's intervals may be split into bpe blocks (a power
// of two dividing 64) on consecutive lanes: each lane sums its intervals'
// terms of the polynomial, the first interval of a block walks its own start
// point instead of taking it from its predecessor's walk (the same value), and
// the bpe partial sums are added across the lanes, whose results are equal to
// the one-lane walk's mod 2^m.
template <typename T, int OB>
__device__ __forceinline__ void spline_walk(const AesCtx &c, T *acc, const uint8_t *sc, const T *tail, T pt,
                                            const uint32_t s[4], int party, int n, int m, int g, int j0, int cnt)
{
    if constexpr (OB == 0)
        dcf_eval_range<T, 16>(c, acc, sc, tail, pt, s, party, n, m, g, j0, cnt);
    else
        dcf_eval_slots<T, OB, (32 / OB < 16 ? 32 / OB : 16)>(c, acc, sc, tail, pt, s, party, n, g, j0, cnt);
}

template <typename T, int OB, bool SPLIT>
__device__ __forceinline__ void spline_eval_body(const AesCtx &c, T *out, const T *x, const uint8_t *key, const T *sE,
                                                 const T *sBeta, const uint8_t *seed, int party, const T *left,
                                                 const T *right, size_t I, int d, int n, int m, size_t N, uint32_t bpe,
                                                 uint32_t *counter)
{
    constexpr int RS = (OB == 0) ? 16 : (32 / OB < 16 ? 32 / OB : 16); // slots per walk
    if constexpr (!SPLIT) bpe = 1;
    const int    C     = d + 1;
    const size_t g     = I * (size_t)C;
    const T     *tails = (const T *)(key + N * 16 * (size_t)n);
    const size_t span  = (I + bpe - 1) / bpe;
    for_each_item(N * bpe, counter, [&](size_t item) {
        const size_t e  = item / bpe;
        const size_t q0 = (item % bpe) * span;
        const size_t q1 = (q0 + span < I) ? q0 + span : I;
        uint32_t s[4];
        load16(s, seed + 16 * e);
        const uint8_t *sc   = key + e * 16 * (size_t)n;
        const T       *tail = tails + e * (n * g + 2 + g);
        const T        xv   = mod_bits<T>(x[e], n);
        const T       *be = sBeta + e * g, *ee = sE + e * g;
        // w[j] = sum over intervals of (cx*beta + e + nxt - cur)[j]; each walk's
        // slots are added with their sign as soon as the walk ends
        T w[8];
#pragma unroll
        for (int j = 0; j < 8; j++) w[j] = 0;
        for (size_t q = q0; q < q1; q++)
        {
            const T L  = mod_bits<T>(left[q], n), R = mod_bits<T>(right[q], n);
            const T Rp = mod_bits<T>((T)(R + 1), n);
            T       acc[RS];
            if (!(q > q0 && mod_bits<T>((T)(right[q - 1] + 1), n) == L))
            {
                spline_walk<T, OB>(c, acc, sc, tail, (T)(xv - 1 - L), s, party, n, m, (int)g, (int)(q * C), C);
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (j < C) w[j] = (T)(w[j] - acc[j < RS ? j : 0]);
            }
            const bool link = (q + 1 < q1) && mod_bits<T>((T)(right[q] + 1), n) == mod_bits<T>(left[q + 1], n);
            spline_walk<T, OB>(c, acc, sc, tail, (T)(xv - 1 - Rp), s, party, n, m, (int)g, (int)(q * C),
                               link ? 2 * C : C);
            const T cx = (T)((int)(xv > L) - (int)(xv > Rp));
#pragma unroll
            for (int j = 0; j < 8; j++)
            {
                if (j < C)
                    w[j] = (T)(w[j] + (T)(cx * be[q * C + j]) + acc[j < RS ? j : 0] + ee[q * C + j]);
            }
            // slots C..2C-1 hold interval q+1's values at its start point
            if (link)
            {
#pragma unroll
                for (int j = 0; j < 8; j++)
                {
                    if (j >= C) break;
                    T v = 0;
#pragma unroll
                    for (int r = 0; r < RS; r++) v = (r == j + C) ? acc[r] : v;
                    w[j] = (T)(w[j] - v);
                }
            }
        }
        T r = 0, pw = 1;
#pragma unroll
        for (int j = 0; j < 8; j++)
        {
            if (j < C)
            {
                r  = (T)(r + (T)(w[j] * pw));
                pw = (T)(pw * xv);
            }
        }
        // the element's bpe lanes are consecutive within one wave and all active
        for (uint32_t o = 1; o < bpe; o <<= 1)
        {
            const uint64_t rv = (uint64_t)r;
            const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)rv, (int)o);
            const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(rv >> 32), (int)o);
            r                 = (T)(r + (T)((uint64_t)lo | ((uint64_t)hi << 32)));
        }
        if (item % bpe == 0) out[e] = mod_bits<T>(r, m);
    });
}

template <typename T, int OB>
__global__ __launch_bounds__(kBlock) void spline_eval_kernel(T *out, const T *x, const uint8_t *key, const T *sE,
                                                             const T *sBeta, const uint8_t *seed, int party,
                                                             const T *left, const T *right, size_t I, int d, int n,
                                                             int m, size_t N, uint32_t *counter)
{
    FSS_KERNEL_PROLOGUE
    spline_eval_body<T, OB, false>(c, out, x, key, sE, sBeta, seed, party, left, right, I, d, n, m, N, 1, counter);
}

// The same with each element's intervals split into bpe > 1 blocks (calls
// too small to give every CU a full workgroup).
template <typename T, int OB>
__global__ __launch_bounds__(kBlock) void spline_eval_split_kernel(T *out, const T *x, const uint8_t *key,
                                                                   const T *sE, const T *sBeta, const uint8_t *seed,
                                                                   int party, const T *left, const T *right, size_t I,
                                                                   int d, int n, int m, size_t N, uint32_t bpe,
                                                                   uint32_t *counter)
{
    FSS_KERNEL_PROLOGUE_STAGED
    spline_eval_body<T, OB, true>(c, out, x, key, sE, sBeta, seed, party, left, right, I, d, n, m, N, bpe, counter);
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

static size_t mic_blocks_per_elem(size_t N, size_t I)
{
    // split each element's intervals until the launch has a lane per CU slot
    const size_t lanes = (size_t)device_cu_count() * kBlock;
    size_t       bpe   = (lanes + N - 1) / N;
    if (bpe < 1) bpe = 1;
    if (bpe > I) bpe = I;
    return bpe;
}

// Interval blocks per element for the spline walk: a power of two dividing
// 64 (so an element's lanes share one wave), at most I, and enough lanes to
// give every CU a full workgroup.
uint32_t spline_blocks_per_elem(size_t N, size_t I)
{
    const size_t lanes = (size_t)device_cu_count() * kBlock;
    const size_t want  = (lanes + N - 1) / N;
    uint32_t     b     = 1;
    while (b * 2 <= 64 && b * 2 <= want && b * 2 <= I) b *= 2;
    return b;
}

int launch_mic_eval_rt(size_t es, int grid, hipStream_t st, void *out, const void *x, const void *key, const void *z,
                       const void *seed, int party, const void *left, const void *right, size_t I, int n, int m,
                       size_t N, uint32_t *counter, int block)
{
    const size_t bpe = mic_blocks_per_elem(N, I);
    if (block < kBlock)
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

template <typename T, int OB>
static void spline_launch_ob(int grid, int block, hipStream_t st, void *out, const void *x, const void *key, const void *e,
                             const void *beta, const void *seed, int party, const void *left, const void *right,
                             size_t I, int degree, int n, int m, size_t N, uint32_t *counter)
{
    // split only as far as the launch has lanes for (the serving runtime sizes
    // its grid by elements; a split it has no lanes for only adds walks)
    uint32_t bpe = spline_blocks_per_elem(N, I);
    while (bpe > 1 && N * bpe > (size_t)grid * (size_t)block) bpe >>= 1;
    if (bpe > 1)
        spline_eval_split_kernel<T, OB><<<grid, block, 0, st>>>(
            (T *)out, (const T *)x, (const uint8_t *)key, (const T *)e, (const T *)beta, (const uint8_t *)seed, party,
            (const T *)left, (const T *)right, I, degree, n, m, N, bpe, counter);
    else
        spline_eval_kernel<T, OB><<<grid, block, 0, st>>>((T *)out, (const T *)x, (const uint8_t *)key, (const T *)e,
                                                           (const T *)beta, (const uint8_t *)seed, party,
                                                           (const T *)left, (const T *)right, I, degree, n, m, N,
                                                           counter);
}

// degree <= 7: the slot-conversion kernel for OB = ceil(m/8) when a linked
// walk's 2(degree+1) slots fit 32 bytes, else the generic conversion.
template <typename T>
static void spline_launch(int grid, int block, hipStream_t st, void *out, const void *x, const void *key, const void *e,
                          const void *beta, const void *seed, int party, const void *left, const void *right, size_t I,
                          int degree, int n, int m, size_t N, uint32_t *counter)
{
    const int ob = (m + 7) / 8;
    // (64-bit elements with 2-byte outputs keep the generic kernel: its register footprint is smaller)
    const int fast = (2 * (degree + 1) * ob <= 32 && !(sizeof(T) == 8 && ob == 2)) ? ob : 0;
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
    const size_t bpe     = mic_blocks_per_elem(N, I);
    const LaunchShape ls = shape_for(N * bpe);
    // measured on MI355X (16 intervals, n = 24): the static split is faster
    // up to 2^20 elements (2^14: 282 against 336 us), the work queue from
    // 2^22 (32.3 against 36.7 ms)
    if (N < ((size_t)1 << 21))
        return launch_mic_eval_rt(es, ls.grid, st, sharedOut, maskedX, key, sharedZ, seed, partyId, leftEndpoints,
                                  rightEndpoints, I, (int)n, (int)m, N, nullptr, ls.block);
    uint32_t *counter = queue_counter_acquire(st);
    if (!counter) return FAST_FSS_RUNTIME_ERROR;
    const int rc = launch_mic_eval_rt(es, ls.grid, st, sharedOut, maskedX, key, sharedZ, seed, partyId, leftEndpoints,
                                      rightEndpoints, I, (int)n, (int)m, N, counter, ls.block);
    queue_counter_release(st, counter);
    return rc;
}

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
