// FastFss_cuda_dpf* and FastFss_cuda_grotto* on MI355X.
//
// Validation follows the reference Task::check order (src/kernel/dpf.h,
// src/kernel/grotto.h).  Every kernel uses the persistent 1024-thread
// workgroup-per-CU shape of the AES tables and a grid-stride loop over work
// items (elements, or element x point pairs for the multi-point forms).
#include "FastFss/cuda/dpf.h"
#include "FastFss/cuda/grotto.h"
#include "FastFss/hip/grotto_full.h"
#include "dpf_device.h"
#include "fss_host.h"
#include "fss_launch.h"

#include <atomic>

namespace fss {

// ------------------------------------------------------------------ DPF ----

template <typename T>
__global__ __launch_bounds__(kBlock) void dpf_keygen_kernel(uint8_t *key, const T *alpha, const T *beta,
                                                            const uint8_t *seed0, const uint8_t *seed1, int n, int m,
                                                            int g, size_t N, uint32_t *counter)
{
    FSS_KERNEL_PROLOGUE
    T *tails = (T *)(key + N * 16 * (size_t)n);
    for_each_item(N, counter, [&](size_t e) {
        uint32_t a[4], b[4];
        load16(a, seed0 + 16 * e);
        load16(b, seed1 + 16 * e);
        dpf_keygen_one<T>(c, key + e * 16 * (size_t)n, tails + e * (size_t)(2 + g), alpha[e],
                          beta ? beta + e * (size_t)g : nullptr, a, b, n, m, g);
    });
}

template <typename T>
__global__ __launch_bounds__(kBlock) void dpf_eval_kernel(T *out, const T *x, const uint8_t *key, const uint8_t *seed,
                                                          int party, int n, int m, int g, size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    const T *tails = (const T *)(key + N * 16 * (size_t)n);
    FSS_GRID_LOOP(e, N)
    {
        uint32_t s[4];
        load16(s, seed + 16 * e);
        dpf_eval_one<T>(c, out + e * (size_t)g, key + e * 16 * (size_t)n, tails + e * (size_t)(2 + g), x[e], s,
                        party, n, m, g);
    }
}

template <typename T, int CH>
__global__ __launch_bounds__(kBlock) void dpf_eval_burst_kernel(T *out, const T *x, const uint8_t *key,
                                                                const uint8_t *seed, int party, int n, int m, size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    const T *tails = (const T *)(key + N * 16 * (size_t)n);
    FSS_GRID_LOOP(e, N)
    {
        uint32_t s[4];
        load16(s, seed + 16 * e);
        out[e] = dpf_eval_g1_burst<T, CH>(c, key + e * 16 * (size_t)n, tails + e * 3, x[e], s, party, n, m);
    }
}

// Four lanes per element (dpf_eval_g1_quad); the grid-stride step is a
// multiple of 4, so a quad always holds one element.
template <typename T>
__global__ __launch_bounds__(kBlock) void dpf_eval_quad_kernel(T *out, const T *x, const uint8_t *key,
                                                               const uint8_t *seed, int party, int n, int m, size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    const uint32_t q = threadIdx.x & 3u;
    QuadAes        qa;
    qa.init(c, q);
    const T *tails = (const T *)(key + N * 16 * (size_t)n);
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; (g >> 2) < N; g += (size_t)gridDim.x * blockDim.x)
    {
        const size_t   e = g >> 2;
        const uint32_t s = *(const uint32_t *)(seed + 16 * e + 4 * q);
        const T        v = dpf_eval_g1_quad<T>(c, qa, key + e * 16 * (size_t)n, tails + e * 3, x[e], s, party, n, m, q);
        if (q == 0) out[e] = v;
    }
}

// Work-queue form (see dcf_eval_queue_kernel): 64-element tiles per wave.
template <typename T, int CH>
__device__ __forceinline__ void dpf_eval_queue_body(const AesCtx &c, T *out, const T *x, const uint8_t *key,
                                                    const uint8_t *seed, int party, int n, int m, size_t N,
                                                    uint32_t *counter)
{
    const T       *tails = (const T *)(key + N * 16 * (size_t)n);
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
            uint32_t s[4];
            load16(s, seed + 16 * e);
            out[e] = dpf_eval_g1_burst<T, CH>(c, key + e * 16 * (size_t)n, tails + e * 3, x[e], s, party, n, m);
        }
    }
}

template <typename T, int CH>
__global__ __launch_bounds__(kBlock) void dpf_eval_queue_kernel(T *out, const T *x, const uint8_t *key,
                                                                const uint8_t *seed, int party, int n, int m, size_t N,
                                                                uint32_t *counter)
{
    FSS_KERNEL_PROLOGUE
    dpf_eval_queue_body<T, CH>(c, out, x, key, seed, party, n, m, N, counter);
}


// One work item per (element, point): out[e][j] = Eval(x_e - point_j)
// (src/kernel/dpf.h:481-534).
template <typename T>
__global__ __launch_bounds__(kBlock) void dpf_eval_multi_kernel(T *out, const T *x, const uint8_t *key,
                                                                const uint8_t *seed, int party, const T *points,
                                                                size_t P, int n, int m, int g, size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
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

__device__ __forceinline__ GrottoKeyPtrs grotto_key(const uint8_t *key, int n, size_t N)
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
    FSS_KERNEL_PROLOGUE_STAGED
    const GrottoKeyPtrs k   = grotto_key(key, n, N);
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

// Four lanes per element (grotto_eval_quad); the grid-stride step is a
// multiple of 4, so a quad always holds one element.
template <typename T>
__global__ __launch_bounds__(kBlock) void grotto_eval_quad_kernel(T *out, const T *x, const uint8_t *key,
                                                                  const uint8_t *seed, int party, int n, int mode,
                                                                  size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    const uint32_t q = threadIdx.x & 3u;
    QuadAes        qa;
    qa.init(c, q);
    const GrottoKeyPtrs k   = grotto_key(key, n, N);
    const T            *tlr = (const T *)k.tlr;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; (g >> 2) < N; g += (size_t)gridDim.x * blockDim.x)
    {
        const size_t   e  = g >> 2;
        const uint32_t s  = *(const uint32_t *)(seed + 16 * e + 4 * q);
        const uint8_t *sc = k.sCW + e * 16 * (size_t)(n - 6);
        const T v = grotto_eval_quad<T>(c, qa, sc, k.lastCW[e], tlr[2 * e], tlr[2 * e + 1], x[e], s, party, n, mode, q);
        if (q == 0) out[e] = v;
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
    FSS_KERNEL_PROLOGUE_STAGED
    const GrottoKeyPtrs k   = grotto_key(key, n, N);
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
    FSS_KERNEL_PROLOGUE_STAGED
    const GrottoKeyPtrs k   = grotto_key(key, n, N);
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

// ------------------------------------------------------ full-domain Grotto ----
//
// Lane (e, r) owns the subtree r (0 <= r < 2^L) of key e.  The workgroup
// first computes the top nodes its lanes share, breadth first through LDS:
// depth d (1 <= d < L) has one node per 2^(L-d) lanes, each made by one thread
// from its parent (one key schedule + one block, as the per-point eval's
// step), so a node is computed once rather than by each of its lanes.  Each
// lane then takes the step to its own depth-L node, and expands the remaining
// D = n-6-L levels depth-first, each
// inner node giving both children from one key schedule and two blocks.  The
// pending right children sit in a register stack; every lane of a wave visits
// its leaves in the same order, so the stack depth is wave-uniform and the
// stack slots are selected with static indices.
//
// A leaf covers 64 points.  With u the corrected leaf word and P the parity of
// the path (src/impl/grotto.h:338-389), the shares of point p are
//   eq: bit p of u
//   lt: P ^ party ^ (bits 0..p of u)        (inclusive prefix parity)
//   le: P ^ party ^ (bits 0..p-1 of u)      (exclusive prefix parity)
// The per-point eval's early exit for le (remaining input bits all zero) only
// skips left turns, which leave P unchanged, so it yields the same value.

constexpr int kFullMaxDepth = 8;

__device__ __forceinline__ uint64_t prefix_parity64(uint64_t u)
{
    u ^= u << 1;
    u ^= u << 2;
    u ^= u << 4;
    u ^= u << 8;
    u ^= u << 16;
    u ^= u << 32;
    return u;
}

// The 16-byte store q of a leaf's 64 group elements (0/1 values from `bits`).
template <typename T>
__device__ __forceinline__ void full_leaf_chunk(uint32_t w[4], uint64_t bits, int q)
{
    constexpr int per = 16 / (int)sizeof(T); // points per 16-byte store
    const uint64_t b  = bits >> (q * per);
    if constexpr (sizeof(T) == 1)
    {
#pragma unroll
        for (int j = 0; j < 4; j++)
        {
            const uint32_t nib = (uint32_t)(b >> (4 * j)) & 0xFu;
            w[j] = (nib & 1u) | ((nib & 2u) << 7) | ((nib & 4u) << 14) | ((nib & 8u) << 21);
        }
    }
    else if constexpr (sizeof(T) == 2)
    {
#pragma unroll
        for (int j = 0; j < 4; j++) w[j] = ((uint32_t)(b >> (2 * j)) & 1u) | (((uint32_t)(b >> (2 * j + 1)) & 1u) << 16);
    }
    else if constexpr (sizeof(T) == 4)
    {
#pragma unroll
        for (int j = 0; j < 4; j++) w[j] = (uint32_t)(b >> j) & 1u;
    }
    else
    {
        w[0] = (uint32_t)b & 1u;
        w[1] = 0;
        w[2] = (uint32_t)(b >> 1) & 1u;
        w[3] = 0;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void grotto_full_kernel(void *out, const uint8_t *key, const uint8_t *seed,
                                                             int party, int mode, int packed, int n, int L, int D,
                                                             size_t N)
{
    // the staged table fill: at 4096 keys the whole launch takes about 0.11 ms,
    // and the packed layout ran 2.7% faster with it (same-process A/B)
    FSS_KERNEL_PROLOGUE_STAGED
    const GrottoKeyPtrs k     = grotto_key(key, n, N);
    const T            *tlr   = (const T *)k.tlr;
    const size_t        lanes = N << L;
    const int           lv    = n - 6; // tree levels above the leaf words
    const uint32_t      lane  = threadIdx.x & 63u;
    // nodes of two consecutive top depths (at most blockDim/2 + 1 a depth):
    // seed words and t | parity << 1
    __shared__ uint32_t topS[2][4][kBlock / 2 + 1];
    __shared__ uint32_t topTP[2][kBlock / 2 + 1];
    // the workgroup iterates together (its lanes share top nodes); whole waves
    // work together (the element layout's stores are shared across the
    // wave); a lane past the end redoes its wave's first subtree and writes
    // nothing of its own
    for (size_t base = (size_t)blockIdx.x * blockDim.x; base < lanes; base += (size_t)gridDim.x * blockDim.x)
    {
        const size_t end = base + blockDim.x < lanes ? base + blockDim.x : lanes;
        for (int d = 1; d < L; d++)
        {
            // node u of depth d is the ancestor of lanes u << (L-d) ..
            const size_t u0 = base >> (L - d), cnt = ((end - 1) >> (L - d)) - u0 + 1;
            if (threadIdx.x < cnt)
            {
                const size_t u = u0 + threadIdx.x, e = u >> d;
                uint32_t     s[4], t, P;
                if (d == 1)
                {
                    load16(s, seed + 16 * e);
                    s[0] &= ~1u;
                    t = (uint32_t)party;
                    P = 0;
                }
                else
                {
                    const uint32_t j = (uint32_t)((u >> 1) - (base >> (L - d + 1)));
#pragma unroll
                    for (int q = 0; q < 4; q++) s[q] = topS[(d - 1) & 1][q][j];
                    t = topTP[(d - 1) & 1][j] & 1u;
                    P = topTP[(d - 1) & 1][j] >> 1;
                }
                const uint32_t bit = (uint32_t)u & 1u, prev = t;
                tree_eval_step<T>(c, s, t, bit, k.sCW + e * 16 * (size_t)lv + 16 * (d - 1), tlr[2 * e],
                                  tlr[2 * e + 1], d - 1);
                P ^= bit & (prev ^ t);
#pragma unroll
                for (int q = 0; q < 4; q++) topS[d & 1][q][threadIdx.x] = s[q];
                topTP[d & 1][threadIdx.x] = t | (P << 1);
            }
            __syncthreads();
        }
        const size_t   wb    = base + (threadIdx.x & ~63u);
        const bool     valid = base + threadIdx.x < lanes;
        const size_t   w     = valid ? base + threadIdx.x : (wb < lanes ? wb : base);
        const size_t   e     = w >> L;
        const uint32_t r     = (uint32_t)(w & ((1u << L) - 1u));
        const uint8_t *sc    = k.sCW + e * 16 * (size_t)lv;
        uint32_t       s[4], t, P;
        if (L < 2)
        {
            load16(s, seed + 16 * e);
            s[0] &= ~1u;
            t = (uint32_t)party;
            P = 0;
        }
        else
        {
            const uint32_t j = (uint32_t)((w >> 1) - (base >> 1));
#pragma unroll
            for (int q = 0; q < 4; q++) s[q] = topS[(L - 1) & 1][q][j];
            t = topTP[(L - 1) & 1][j] & 1u;
            P = topTP[(L - 1) & 1][j] >> 1;
        }
        // the next iteration's top nodes overwrite these buffers
        if (L >= 2) __syncthreads();
        if (wb >= lanes) continue;
        const T        tl = tlr[2 * e], tr = tlr[2 * e + 1];
        const uint64_t lastCW = k.lastCW[e];
        if (L >= 1)
        {
            const uint32_t bit = r & 1u, prev = t;
            tree_eval_step<T>(c, s, t, bit, sc + 16 * (L - 1), tl, tr, L - 1);
            P ^= bit & (prev ^ t);
        }

        // pending right children of depths 0 .. D-2; the last level's right
        // child is a leaf, kept in (pr2, pr3, prtp) and emitted right after
        // its sibling, so it never goes through the stack
        uint32_t stS[kFullMaxDepth - 1][4];
        uint32_t stTP[kFullMaxDepth - 1]; // t | parity << 1 of the pending right child
        uint32_t pr2 = 0, pr3 = 0, prtp = 0;
        const int pair = D > 0 ? 2 : 1; // leaves emitted per descent
        const uint32_t leaves = 1u << D;
        int            depth  = 0;
        for (uint32_t leaf = 0;;)
        {
            while (depth < D)
            {
                const int level = L + depth;
                uint32_t  blk[2][4];
                uint32_t  cw[4];
                load16(cw, sc + 16 * level);
                aes_encrypt_prg2(c, s, blk);
                const uint32_t mk = 0u - t;
                const uint32_t tL = (blk[0][0] & 1u) ^ ((uint32_t)(tl >> level) & 1u & t);
                const uint32_t tR = (blk[1][0] & 1u) ^ ((uint32_t)(tr >> level) & 1u & t);
                const uint32_t rs0 = (blk[1][0] & ~1u) ^ (cw[0] & mk), rs1 = blk[1][1] ^ (cw[1] & mk);
                const uint32_t rs2 = blk[1][2] ^ (cw[2] & mk), rs3 = blk[1][3] ^ (cw[3] & mk);
                const uint32_t rtp = tR | ((P ^ t ^ tR) << 1);
                if (depth == D - 1)
                {
                    pr2  = rs2;
                    pr3  = rs3;
                    prtp = rtp;
                }
                else
                {
#pragma unroll
                    for (int d = 0; d < kFullMaxDepth - 1; d++)
                        if (d == depth)
                        {
                            stS[d][0] = rs0;
                            stS[d][1] = rs1;
                            stS[d][2] = rs2;
                            stS[d][3] = rs3;
                            stTP[d]   = rtp;
                        }
                }
                s[0] = (blk[0][0] & ~1u) ^ (cw[0] & mk);
                s[1] = blk[0][1] ^ (cw[1] & mk);
                s[2] = blk[0][2] ^ (cw[2] & mk);
                s[3] = blk[0][3] ^ (cw[3] & mk);
                t    = tL;
                depth++;
            }

#pragma unroll 1
            for (int k = 0; k < pair; k++, leaf++)
            {
            if (k == 1)
            {
                s[2] = pr2;
                s[3] = pr3;
                t    = prtp & 1u;
                P    = prtp >> 1;
            }
            const uint64_t u = ((uint64_t)s[2] | ((uint64_t)s[3] << 32)) ^ (t ? lastCW : 0ull);
            uint64_t       bits;
            if (mode == 2)
                bits = u;
            else
            {
                bits = prefix_parity64(mode == 1 ? (u << 1) : u);
                if ((P ^ (uint32_t)party) & 1u) bits = ~bits;
            }
            if (packed)
            {
                if (valid) ((uint64_t *)out)[(e << lv) + (((size_t)r << D) | leaf)] = bits;
            }
            else
            {
                // The wave writes its 64 leaves' 64*sizeof(T)-byte runs together:
                // store q covers chunks 64q .. 64q+63 of the wave's leaves in
                // order, so each store fills whole lines of consecutive leaves'
                // runs instead of one 16-byte piece of 64 scattered runs.
                constexpr int CPL = 4 * (int)sizeof(T); // 16-byte chunks per leaf
#pragma unroll 1
                for (int q = 0; q < CPL; q++)
                {
                    const uint32_t gci   = (uint32_t)q * 64u + lane;
                    const uint32_t src   = gci / CPL, chunk = gci % CPL;
                    const uint64_t sbits = __shfl(bits, (int)src);
                    const size_t   sw    = wb + src;
                    if (sw < lanes)
                    {
                        const size_t sLeaf = ((sw >> L) << lv) + (((sw & ((1u << L) - 1u)) << D) | leaf);
                        uint32_t     v[4];
                        full_leaf_chunk<T>(v, sbits, (int)chunk);
                        store16_aligned((uint8_t *)out + ((sLeaf << 6) * sizeof(T)) + 16 * chunk, v);
                    }
                }
            }

            }
            if (leaf == leaves) break;
            // the next pair's path leaves the current one at depth
            // D-1-ctz(leaf) <= D-2: resume from the right child pushed there
            const int d0 = D - 1 - (int)__builtin_ctz(leaf);
#pragma unroll
            for (int d = 0; d < kFullMaxDepth - 1; d++)
                if (d == d0)
                {
                    FSS_UNIFORM_CASE;
                    s[0] = stS[d][0];
                    s[1] = stS[d][1];
                    s[2] = stS[d][2];
                    s[3] = stS[d][3];
                    t    = stTP[d] & 1u;
                    P    = stTP[d] >> 1;
                }
            depth = d0 + 1;
        }
    }
}

// DPF evalAll (src/kernel/dpf.h:182-277): element e's share at point j goes
// to out[e][(x_e - j) mod 2^n].  Lane (e, r) walks L levels to subtree r and
// expands its D = n - L levels depth-first like grotto_full_kernel (one key
// schedule and two blocks per inner node, pending right children in a
// register stack), converting each leaf seed as dpf_eval_one does, so every
// output equals dpf_eval_one at that point.
template <typename T>
__global__ __launch_bounds__(kBlock) void dpf_eval_all_dfs_kernel(T *out, const T *x, const uint8_t *key,
                                                                  const uint8_t *seed, int party, int n, int m, int g,
                                                                  int L, int D, size_t N)
{
    FSS_KERNEL_PROLOGUE
    const T     *tails = (const T *)(key + N * 16 * (size_t)n);
    const size_t lanes = N << L;
    const size_t size  = (size_t)1 << n;
    FSS_GRID_LOOP(w, lanes)
    {
        const size_t   e  = w >> L;
        const uint64_t r  = (uint64_t)w & ((1ull << L) - 1ull);
        const uint8_t *sc = key + e * 16 * (size_t)n;
        const T       *tl = tails + e * (size_t)(2 + g);
        const T        tL = tl[0], tR = tl[1];
        const size_t   xe = (size_t)mod_bits<T>(x[e], n);
        T             *oe = out + e * size * (size_t)g;
        uint32_t       s[4];
        load16(s, seed + 16 * e);
        s[0] &= ~1u;
        uint32_t t = (uint32_t)party;
        for (int i = 0; i < L; i++)
            tree_eval_step<T>(c, s, t, (uint32_t)(r >> (L - 1 - i)) & 1u, sc + 16 * i, tL, tR, i);

        // pending right children of depths 0 .. D-2; the last level's right
        // child is a leaf, kept in (pr, prt) and converted right after its
        // sibling, so it never goes through the stack
        uint32_t       stS[kFullMaxDepth - 1][4];
        uint32_t       stT[kFullMaxDepth - 1];
        uint32_t       pr[4] = {0u, 0u, 0u, 0u}, prt = 0;
        const int      pair = D > 0 ? 2 : 1; // leaves converted per descent
        const uint32_t leaves = 1u << D;
        int            depth  = 0;
        for (uint32_t leaf = 0;;)
        {
            while (depth < D)
            {
                const int level     = L + depth;
                uint32_t  blk[2][4];
                uint32_t  cw[4];
                load16(cw, sc + 16 * level);
                aes_encrypt_prg2(c, s, blk);
                const uint32_t mk = 0u - t;
                const uint32_t tl_ = (blk[0][0] & 1u) ^ ((uint32_t)(tL >> level) & 1u & t);
                const uint32_t tr_ = (blk[1][0] & 1u) ^ ((uint32_t)(tR >> level) & 1u & t);
                const uint32_t r0 = (blk[1][0] & ~1u) ^ (cw[0] & mk), r1 = blk[1][1] ^ (cw[1] & mk);
                const uint32_t r2 = blk[1][2] ^ (cw[2] & mk), r3 = blk[1][3] ^ (cw[3] & mk);
                if (depth == D - 1)
                {
                    pr[0] = r0;
                    pr[1] = r1;
                    pr[2] = r2;
                    pr[3] = r3;
                    prt   = tr_;
                }
                else
                {
#pragma unroll
                    for (int d = 0; d < kFullMaxDepth - 1; d++)
                        if (d == depth)
                        {
                            stS[d][0] = r0;
                            stS[d][1] = r1;
                            stS[d][2] = r2;
                            stS[d][3] = r3;
                            stT[d]    = tr_;
                        }
                }
                s[0] = (blk[0][0] & ~1u) ^ (cw[0] & mk);
                s[1] = blk[0][1] ^ (cw[1] & mk);
                s[2] = blk[0][2] ^ (cw[2] & mk);
                s[3] = blk[0][3] ^ (cw[3] & mk);
                t    = tl_;
                depth++;
            }

#pragma unroll 1
            for (int k = 0; k < pair; k++, leaf++)
            {
            if (k == 1)
            {
                s[0] = pr[0];
                s[1] = pr[1];
                s[2] = pr[2];
                s[3] = pr[3];
                t    = prt;
            }
            const size_t j   = (size_t)((r << D) | leaf);
            T           *dst = oe + ((xe - j) & (size - 1)) * (size_t)g;
            if (g == 1)
            {
                uint32_t conv;
                if (m <= 16)
                    conv = s[0] & 0xffu;
                else
                {
                    uint32_t b[1][4] = {{0u, 0u, 0u, 0u}};
                    aes_encrypt<1>(c, s, b);
                    conv = b[0][0] & 0xffu;
                }
                const T term = (T)((T)conv + (t ? tl[2] : (T)0));
                dst[0]       = party ? (T)(0 - term) : term;
            }
            else
            {
                ConvCtx cv;
                cv.init(s, m, g);
                for (int jj = 0; jj < g; jj++)
                {
                    const T term = (T)((T)cv.at(c, jj) + (t ? tl[2 + jj] : (T)0));
                    dst[jj]      = party ? (T)(0 - term) : term;
                }
            }

            }
            if (leaf == leaves) break;
            // resume from the right child pushed at depth D-1-ctz(leaf) <= D-2
            const int d0 = D - 1 - (int)__builtin_ctz(leaf);
#pragma unroll
            for (int d = 0; d < kFullMaxDepth - 1; d++)
                if (d == d0)
                {
                    FSS_UNIFORM_CASE;
                    s[0] = stS[d][0];
                    s[1] = stS[d][1];
                    s[2] = stS[d][2];
                    s[3] = stS[d][3];
                    t    = stT[d];
                }
            depth = d0 + 1;
        }
    }
}

// Subtree split of the full-domain evaluations (Grotto full domain, DPF
// evalAll): enough lanes to fill every CU once (or the configured target),
// and at most kFullMaxDepth levels expanded per lane.
static std::atomic<size_t> g_full_lane_target{0};

inline void full_domain_split(int lv, size_t N, int &L, int &D)
{
    const size_t conf   = g_full_lane_target.load(std::memory_order_relaxed);
    const size_t target = conf ? conf : (size_t)device_cu_count() * kBlock;
    L                   = 0;
    while (L < lv && (N << L) < target) L++;
    if (lv - L > kFullMaxDepth) L = lv - kFullMaxDepth;
    D = lv - L;
}

// Four lanes per element while they leave at most two waves per SIMD.
static bool dpf_quad_lanes(size_t N)
{
    return 4 * N <= (size_t)device_cu_count() * 512;
}

int launch_dpf_eval_g1_rt(size_t es, int grid, hipStream_t st, void *out, const void *x, const void *key,
                          const void *seed, int party, int n, int m, size_t N, int block)
{
    // the per-call shape of a small call (block < kBlock) takes the per-call
    // entry point's four-lane kernel; one workgroup per 1024 elements keeps
    // the given launch
    if (block < kBlock && dpf_quad_lanes(N))
    {
        FSS_DISPATCH(es, dpf_eval_quad_kernel<T><<<FSS_SHAPE(4 * N), 0, st>>>(
                             (T *)out, (const T *)x, (const uint8_t *)key, (const uint8_t *)seed, party, n, m, N));
        return launch_status();
    }
    FSS_DISPATCH(es, dpf_eval_burst_kernel<T, 8><<<grid, block, 0, st>>>((T *)out, (const T *)x,
                                                                           (const uint8_t *)key,
                                                                           (const uint8_t *)seed, party, n, m, N));
    return launch_status();
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
    if (es > 8)
    {
        if (!aligned(alpha, 8) || (b && !aligned(b, 8))) return FAST_FSS_RUNTIME_ERROR;
        return wide_dpf_keygen(st, key, alpha, b, seed0, seed1, (int)n, (int)m, g, es, N);
    }
    uint32_t *counter = queue_counter_acquire(st);
    if (!counter) return FAST_FSS_RUNTIME_ERROR;
    FSS_DISPATCH(es, dpf_keygen_kernel<T><<<FSS_SHAPE(N), 0, st>>>((uint8_t *)key, (const T *)alpha,
                                                                           (const T *)b, (const uint8_t *)seed0,
                                                                           (const uint8_t *)seed1, (int)n, (int)m,
                                                                           (int)g, N, counter));
    queue_counter_release(st, counter);
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
    if (es > 8)
    {
        if (!aligned(sharedOut, 8) || !aligned(maskedX, 8)) return FAST_FSS_RUNTIME_ERROR;
        return wide_dpf_eval(st, 0, sharedOut, maskedX, key, seed, partyId, (int)n, (int)m, g, es, N, nullptr, 0);
    }
    // measured on MI355X (n = m = 32): the static split beats the work queue up
    // to 2^20 elements (2^16: 55 against 82 us, 2^20: 446 against 500 us); the
    // queue evens out the waves of large launches
    if (g == 1 && N >= ((size_t)1 << 21))
    {
        uint32_t *counter = queue_counter_acquire(st);
        if (!counter) return FAST_FSS_RUNTIME_ERROR;
        FSS_DISPATCH(es, dpf_eval_queue_kernel<T, 8><<<FSS_SHAPE(N), 0, st>>>(
                             (T *)sharedOut, (const T *)maskedX, (const uint8_t *)key, (const uint8_t *)seed,
                             partyId, (int)n, (int)m, N, counter));
        queue_counter_release(st, counter);
    }
    else if (g == 1 && dpf_quad_lanes(N))
    {
        FSS_DISPATCH(es, dpf_eval_quad_kernel<T><<<FSS_SHAPE(4 * N), 0, st>>>(
                             (T *)sharedOut, (const T *)maskedX, (const uint8_t *)key, (const uint8_t *)seed,
                             partyId, (int)n, (int)m, N));
    }
    else if (g == 1)
    {
        FSS_DISPATCH(es, dpf_eval_burst_kernel<T, 8><<<FSS_SHAPE(N), 0, st>>>(
                             (T *)sharedOut, (const T *)maskedX, (const uint8_t *)key, (const uint8_t *)seed,
                             partyId, (int)n, (int)m, N));
    }
    else
    {
        FSS_DISPATCH(es, dpf_eval_kernel<T><<<FSS_SHAPE(N), 0, st>>>((T *)sharedOut, (const T *)maskedX,
                                                                             (const uint8_t *)key,
                                                                             (const uint8_t *)seed, partyId, (int)n,
                                                                             (int)m, (int)g, N));
    }
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
    hipStream_t st = resolve_stream(cudaStreamPtr);
    if (es > 8)
    {
        if (!aligned(sharedOut, 8) || !aligned(maskedX, 8)) return FAST_FSS_RUNTIME_ERROR;
        return wide_dpf_eval(st, 1, sharedOut, maskedX, key, seed, partyId, (int)n, (int)m, g, es, N, nullptr, 0);
    }
    int L, D;
    full_domain_split((int)n, N, L, D);
    FSS_DISPATCH(es, dpf_eval_all_dfs_kernel<T><<<FSS_SHAPE(N << L), 0, st>>>(
                         (T *)sharedOut, (const T *)maskedX, (const uint8_t *)key, (const uint8_t *)seed, partyId,
                         (int)n, (int)m, (int)g, L, D, N));
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
    if (es > 8)
    {
        if (!aligned(sharedOut, 8) || !aligned(maskedX, 8) || !aligned(point, 8)) return FAST_FSS_RUNTIME_ERROR;
        return wide_dpf_eval(st, 2, sharedOut, maskedX, key, seed, partyId, (int)n, (int)m, g, es, N, point, P);
    }
    FSS_DISPATCH(es, dpf_eval_multi_kernel<T><<<FSS_SHAPE(N * P), 0, st>>>(
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
    if (es > 8)
    {
        if (!aligned(alpha, 8)) return FAST_FSS_RUNTIME_ERROR;
        return wide_grotto_keygen(st, key, alpha, seed0, seed1, (int)n, es, N);
    }
    FSS_DISPATCH(es, grotto_keygen_kernel<T><<<FSS_SHAPE(N), 0, st>>>(
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
    if (es > 8)
    {
        if (!aligned(out, 8) || !aligned(x, 8)) return FAST_FSS_RUNTIME_ERROR;
        return wide_grotto_eval(st, 0, mode, out, nullptr, nullptr, x, key, seed, party, (int)n, es, N, nullptr,
                                nullptr, 0, nullptr, 0);
    }
    if (dpf_quad_lanes(N))
    {
        FSS_DISPATCH(es, grotto_eval_quad_kernel<T><<<FSS_SHAPE(4 * N), 0, st>>>(
                             (T *)out, (const T *)x, (const uint8_t *)key, (const uint8_t *)seed, party, (int)n, mode,
                             N));
    }
    else
    {
        FSS_DISPATCH(es, grotto_eval_kernel<T><<<FSS_SHAPE(N), 0, st>>>(
                             (T *)out, (const T *)x, (const uint8_t *)key, (const uint8_t *)seed, party, (int)n, mode,
                             N));
    }
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
    if (es > 8)
    {
        if (!aligned(sharedBooleanOut, 8) || !aligned(maskedX, 8) || !aligned(leftEndpoints, 8) ||
            !aligned(rightEndpoints, 8))
            return FAST_FSS_RUNTIME_ERROR;
        void *bits = nullptr;
        if (hipMallocAsync(&bits, N * I * sizeof(int), st) != hipSuccess) return FAST_FSS_RUNTIME_ERROR;
        const int rc = wide_grotto_eval(st, 1, 1, sharedBooleanOut, nullptr, bits, maskedX, key, seed, partyId,
                                        (int)n, es, N, leftEndpoints, rightEndpoints, I, nullptr, 0);
        if (hipFreeAsync(bits, st) != hipSuccess) return FAST_FSS_RUNTIME_ERROR;
        return rc;
    }
    FSS_DISPATCH(es, grotto_mic_kernel<T><<<FSS_SHAPE(N), 0, st>>>(
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
    if (es > 8)
    {
        if (!aligned(sharedOutE, 8) || !aligned(sharedOutT, 8) || !aligned(maskedX, 8) || !aligned(leftEndpoints, 8) ||
            !aligned(rightEndpoints, 8) || !aligned(lookUpTable, 8))
            return FAST_FSS_RUNTIME_ERROR;
        if (hipMallocAsync(&bits, N * I * sizeof(int), st) != hipSuccess) return FAST_FSS_RUNTIME_ERROR;
        const int rc = wide_grotto_eval(st, 2, 1, sharedOutE, sharedOutT, bits, maskedX, key, seed, partyId, (int)n,
                                        es, N, leftEndpoints, rightEndpoints, I, lookUpTable, lutNum);
        if (hipFreeAsync(bits, st) != hipSuccess) return FAST_FSS_RUNTIME_ERROR;
        return rc;
    }
    if (hipMallocAsync(&bits, N * I * es, st) != hipSuccess) return FAST_FSS_RUNTIME_ERROR;
    FSS_DISPATCH(es, grotto_lut_kernel<T><<<FSS_SHAPE(N), 0, st>>>(
                         (T *)sharedOutE, (T *)sharedOutT, (T *)bits, (const T *)maskedX, (const uint8_t *)key,
                         (const uint8_t *)seed, partyId, (const T *)leftEndpoints, (const T *)rightEndpoints,
                         (const T *)lookUpTable, lutNum, I, (int)n, N));
    int rc = launch_status();
    (void)hipFreeAsync(bits, st);
    return rc;
}

// ------------------------------------------------- C ABI: full-domain Grotto ----

extern "C" FAST_FSS_API int FastFss_hip_grottoEvalFullDomain(void *sharedOut, size_t sharedOutDataSize,
                                                             const void *key, size_t keyDataSize, const void *seed,
                                                             size_t seedDataSize, int partyId, int mode, int packed,
                                                             size_t bitWidthIn, size_t elementSize,
                                                             size_t elementNum, void *hipStreamPtr)
{
    const size_t n = bitWidthIn, es = elementSize, N = elementNum;
    if (!elem_size_narrow(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(6 <= n && n <= es * 8 && n <= 40)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (mode < FAST_FSS_GROTTO_FULL_LT || mode > FAST_FSS_GROTTO_FULL_EQ || (packed != 0 && packed != 1))
        return FAST_FSS_RUNTIME_ERROR;
    if (partyId != 0 && partyId != 1) return FAST_FSS_INVALID_PARTY_ID_ERROR;
    if (keyDataSize != grotto_key_size(n, es, N)) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (seedDataSize != N * 16) return FAST_FSS_INVALID_SEED_DATA_SIZE_ERROR;
    const size_t perKey = packed ? ((size_t)8 << (n - 6)) : (es << n);
    if (N != 0 && perKey > SIZE_MAX / N) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (sharedOutDataSize != N * perKey) return FAST_FSS_INVALID_SHARED_OUT_DATA_SIZE_ERROR;
    if (N == 0) return FAST_FSS_SUCCESS;
    if (!aligned(key, 8) || !aligned(seed, 8) || !aligned(sharedOut, 16)) return FAST_FSS_RUNTIME_ERROR;
    int L, D;
    full_domain_split((int)n - 6, N, L, D);
    hipStream_t st = resolve_stream(hipStreamPtr);
    FSS_DISPATCH(es, grotto_full_kernel<T><<<FSS_SHAPE(N << L), 0, st>>>(
                         sharedOut, (const uint8_t *)key, (const uint8_t *)seed, partyId, mode, packed, (int)n, L, D,
                         N));
    return launch_status();
}

extern "C" FAST_FSS_API int FastFss_hip_grottoFullSetLaneTarget(size_t lanes)
{
    g_full_lane_target.store(lanes, std::memory_order_relaxed);
    return FAST_FSS_SUCCESS;
}
