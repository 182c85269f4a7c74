// One-time truth tables (reference src/impl/ottt.h, src/kernel/ottt.h) and
// the AES-128-CTR PRNG (reference src/cuda/prng.cu) on MI355X.
#include <string.h>

#include <algorithm>

#include "FastFss/cuda/ottt.h"
#include "FastFss/cuda/prng.h"
#include "fss_device.h"
#include "fss_host.h"
#include "fss_launch.h"

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
// party 0 negates both (src/impl/ottt.h:26-192).
//
// For n >= 6 a wave evaluates one element at a time, one 64-bit key word per
// step: lane L takes bit L of word k (key bit b = 64 k + L), whose table index
// (x - b) mod 2^n is consecutive across the lanes, so the LDS reads of a table
// never bank-conflict.  Each lane adds its bits' table values into registers;
// at the element's end a DPP wave sum gives the total and lane 0 writes it, so
// each output is written once.  E is the popcount of the wave's ballot of set
// bits.  Key words come in with one coalesced load per 64 words (lane L loads
// word L of the block, which covers 64 / W whole elements when an element has
// W < 64 words) and are broadcast to the wave with readlane.
//
// The tables of one pass (up to kOtttTables) sit in LDS.  When they do not fit
// in kOtttLdsBytes they are staged in chunks of C entries: the key bits whose
// index falls in chunk [c0, c0 + C) form one contiguous run b in
// [x - c0 - C + 1, x - c0] (mod 2^n), C / 64 + 1 key words, and a chunked
// launch evaluates kOtttEW elements per wave per batch so that each staged
// chunk serves all of the workgroup's elements.
//
// For n < 6 a group of P = 2^n lanes evaluates one element (64 / P per wave).
constexpr int    kOtttTables   = 4;
constexpr int    kOtttEW       = 4;
// LDS budget of one workgroup's tables: two 1024-lane chunked workgroups fit a
// CU (160 KB).  80 KB rather than 64 halves the chunks at n = 14..16 (two
// barriers and one staging pass each): 2-3% faster in a same-process A/B
// (n = 16, 2^16 elements: 906 -> 881 us).  The resident tables' sizes (2^k or
// 3 * 2^k bytes) fall on the same side of 64 and 80 KB.
constexpr size_t kOtttLdsBytes = 80 * 1024;

template <typename T> struct OtttAcc { using type = uint32_t; };
template <> struct OtttAcc<uint64_t> { using type = uint64_t; };

__device__ inline uint32_t ottt_readlane(uint32_t v, uint32_t lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

// Sum over the wave's 64 lanes (DPP row shifts within rows of 16, then row
// broadcasts 15 and 31), the total read from lane 63.
__device__ inline uint32_t ottt_wave_sum(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false); // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false); // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false); // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false); // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false); // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false); // row_bcast:31
    return ottt_readlane(v, 63);
}

__device__ inline uint64_t ottt_wave_sum(uint64_t v)
{
    // the same steps on 64-bit values: both halves move, then a 64-bit add
#define FSS_OTTT_STEP(ctrl, rm)                                                                             \
    {                                                                                                      \
        const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, ctrl, rm, 0xf, false); \
        const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), ctrl, rm, 0xf, \
                                                                  false);                                  \
        v += (uint64_t)lo | ((uint64_t)hi << 32);                                                          \
    }
    FSS_OTTT_STEP(0x111, 0xf)
    FSS_OTTT_STEP(0x112, 0xf)
    FSS_OTTT_STEP(0x114, 0xf)
    FSS_OTTT_STEP(0x118, 0xf)
    FSS_OTTT_STEP(0x142, 0xa)
    FSS_OTTT_STEP(0x143, 0xc)
#undef FSS_OTTT_STEP
    return (uint64_t)ottt_readlane((uint32_t)v, 63) | ((uint64_t)ottt_readlane((uint32_t)(v >> 32), 63) << 32);
}

// 8 key bytes of an element at word k (two 4-byte loads: keys are 4-byte aligned)
__device__ inline uint64_t ottt_load_word(const uint8_t *kb, uint32_t k)
{
    const uint32_t *w = (const uint32_t *)(kb + 8 * (size_t)k);
    return (uint64_t)w[0] | ((uint64_t)w[1] << 32);
}

template <typename T, typename A>
__device__ inline void ottt_write(T *outE, T *outT, size_t e, int lutNum, int t0, int tb, int party, int writeE,
                                  const A (&acc)[kOtttTables], uint32_t E)
{
#pragma unroll
    for (int t = 0; t < kOtttTables; t++)
        if (t < tb) outT[e * (size_t)lutNum + t0 + t] = (T)(party == 0 ? (A)0 - acc[t] : acc[t]);
    if (writeE) outE[e] = (T)(party == 0 ? (A)0 - (A)E : (A)E);
}

template <typename T>
__device__ inline void ottt_stage(T *tab, const T *lut, int t0, int tb, uint32_t total, uint32_t c0, int logC)
{
    const uint32_t C = 1u << logC;
    for (uint32_t i = threadIdx.x; i < (uint32_t)tb * C; i += blockDim.x)
        tab[i] = lut[(size_t)(t0 + (int)(i >> logC)) * total + c0 + (i & (C - 1))];
}

// v : 0 per lane by the lane's bit of a wave-uniform 64-bit mask in SGPRs
// (one VALU op)
__device__ inline uint32_t ottt_select(uint64_t m, uint32_t v)
{
    uint32_t r;
    asm("v_cndmask_b32_e64 %0, 0, %1, %2" : "=v"(r) : "v"(v), "s"(m));
    return r;
}

template <typename A>
__device__ inline A ottt_select_acc(uint64_t m, A v)
{
    if constexpr (sizeof(A) == 8)
        return (A)ottt_select(m, (uint32_t)v) | ((A)ottt_select(m, (uint32_t)(v >> 32)) << 32);
    else
        return (A)ottt_select(m, (uint32_t)v);
}

// Steps over nwords (a multiple of U) consecutive key words of one element,
// lane (k0 + s) & kmask of kv holding word s (broadcast with readlane, so each
// word is wave-uniform and serves as its select's lane mask).  Lane L's table
// entry for bit L of word s sits at j0 - 64 s of the staged table; the U reads
// of a group are issued, at immediate offsets from one address, before any of
// their adds.  FULL: every set bit counts; otherwise the first and last words
// are masked to the staged run (firstMaskShift, lastMaskBits).  COUNT_E: add
// the words' set bits to E (two scalar instructions a word); callers holding
// the words in vector registers count them there instead.
template <int U, int TB, bool FULL, bool COUNT_E, typename T, typename A>
__device__ inline void ottt_words(A (&acc)[kOtttTables], uint32_t &E, const T *tab, uint32_t stride, uint64_t kv,
                                  uint32_t k0, uint32_t kmask, uint32_t nwords, uint32_t j0, uint32_t firstMaskShift,
                                  uint32_t lastMaskBits)
{
    for (uint32_t s0 = 0; s0 < nwords; s0 += U)
    {
        uint64_t  word[U];
        A         v[U][TB > 0 ? TB : 1];
        const T  *p = tab + (j0 - 64u * (s0 + U - 1));
#pragma unroll
        for (int u = 0; u < U; u++)
        {
            const uint32_t k = (k0 + s0 + u) & kmask;
            word[u] = (uint64_t)ottt_readlane((uint32_t)kv, k) | ((uint64_t)ottt_readlane((uint32_t)(kv >> 32), k) << 32);
            if (!FULL)
            {
                if (s0 + u == 0) word[u] &= ~0ull << firstMaskShift;
                if (s0 + u == nwords - 1) word[u] &= lastMaskBits == 64 ? ~0ull : ((1ull << lastMaskBits) - 1);
            }
#pragma unroll
            for (int t = 0; t < TB; t++) v[u][t] = (A)p[(size_t)t * stride + 64u * (U - 1 - u)];
        }
#pragma unroll
        for (int u = 0; u < U; u++)
        {
            if (COUNT_E) E += (uint32_t)__popcll(word[u]);
            // Either way the word is the select's lane mask.  FULL: the word
            // straight from readlane, selected by the asm form (at n = 12 the
            // builtin below ran 6% slower, at n = 8 the same); masked words:
            // the builtin, which leaves the compiler free to place the masked
            // word (the per-lane bit extraction used before cost 27% at
            // n = 16, same-process A/B with equal outputs)
            const bool take = __builtin_amdgcn_inverse_ballot_w64(word[u]);
#pragma unroll
            for (int t = 0; t < TB; t++)
                acc[t] += FULL ? ottt_select_acc<A>(word[u], v[u][t]) : (take ? v[u][t] : (A)0);
        }
    }
}

// n >= 6, the pass's tables resident in LDS, each stored twice over (entry i
// at i and i + 2^n) so that a lane's entries for all W words of an element
// lie at j0 - 64 s, j0 = (x - L) mod 2^n + 2^n, without wrapping.
//
// A wave walks a stream of 64-word key blocks with the next block's load in
// flight: a unit is one element of W >= 64 words (W / 64 blocks) or 64 / W
// whole elements of W < 64 words (one block, lane L holding word L % W of
// element L / W); lane L < 64 / W also holds x of the unit's element L.
// Each word is broadcast with readlane, so the select takes it as its mask.
template <typename T, int TB>
__global__ __launch_bounds__(256) void ottt_eval_tile_kernel(T *__restrict__ outE, T *__restrict__ outT,
                                                              const T *__restrict__ x,
                                                              const uint8_t *__restrict__ key, int party,
                                                              const T *__restrict__ lut, int lutNum, int t0, int n,
                                                              size_t N, int writeE)
{
    using A = typename OtttAcc<T>::type;
    extern __shared__ __align__(16) unsigned char ottt_lds[];
    T *tab = (T *)ottt_lds;
    const uint32_t total = 1u << n, mask = total - 1, W = total >> 6;
    for (uint32_t i = threadIdx.x; i < (uint32_t)TB * 2 * total; i += blockDim.x)
        tab[i] = lut[(size_t)(t0 + (int)(i / (2 * total))) * total + (i & mask)];
    __syncthreads();
    const uint32_t lane   = threadIdx.x & 63;
    const size_t   perKey = total / 8;
    const uint32_t ET     = W >= 64 ? 1 : 64 / W; // elements per unit
    const uint32_t B      = W >= 64 ? W / 64 : 1; // blocks per unit
    const size_t   units  = (N + ET - 1) / ET;
    const size_t   nwaves = (size_t)gridDim.x * (blockDim.x >> 6);
    size_t         u      = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (u >= units) return;

    auto load_block = [&](size_t uu, uint32_t bb) -> uint64_t {
        const size_t   e  = W >= 64 ? uu : uu * ET + lane / W;
        const uint32_t wk = W >= 64 ? bb * 64 + lane : lane % W;
        return e < N ? ottt_load_word(key + e * perKey, wk) : 0;
    };
    auto load_x = [&](size_t uu) -> uint32_t {
        const size_t e = uu * ET + lane;
        return (lane < ET && e < N) ? (uint32_t)(uint64_t)x[e] : 0u;
    };

    uint32_t blk = 0;
    uint64_t kv  = load_block(u, 0);
    uint32_t xv  = load_x(u);
    A        acc[kOtttTables];
    uint32_t E  = 0;
    uint32_t Ev = 0; // W >= 64: set bits of this lane's words of the element
#pragma unroll
    for (int t = 0; t < kOtttTables; t++) acc[t] = 0;
    while (true)
    {
        // the next block of the stream, loaded before this one is processed
        size_t   un = u;
        uint32_t bn = blk + 1;
        if (bn == B) bn = 0, un = u + nwaves;
        const bool more = un < units;
        uint64_t   kvn  = 0;
        uint32_t   xvn  = xv;
        if (more)
        {
            kvn = load_block(un, bn);
            if (bn == 0) xvn = load_x(un);
        }
        if (W >= 64)
        {
            const uint32_t j0 = ((ottt_readlane(xv, 0) - lane) & mask) + total;
            ottt_words<16, TB, true, false, T, A>(acc, E, tab, 2 * total, kv, 0, 63, 64,
                                                       j0 - 64u * 64u * blk, 0, 64);
            Ev += (uint32_t)__popcll(kv);
            if (blk == B - 1)
            {
#pragma unroll
                for (int t = 0; t < kOtttTables; t++)
                    if (t < TB) acc[t] = ottt_wave_sum(acc[t]);
                E = ottt_wave_sum(Ev);
                if (lane == 0) ottt_write<T, A>(outE, outT, u, lutNum, t0, TB, party, writeE, acc, E);
                E  = 0;
                Ev = 0;
#pragma unroll
                for (int t = 0; t < kOtttTables; t++) acc[t] = 0;
            }
        }
        else
        {
            for (uint32_t le = 0; le < ET; le++)
            {
                const size_t e = u * ET + le;
                if (e >= N) break;
                const uint32_t j0 = ((ottt_readlane(xv, le) - lane) & mask) + total;
                if (W >= 8)
                    ottt_words<8, TB, true, true, T, A>(acc, E, tab, 2 * total, kv, le * W, 63, W, j0, 0, 64);
                else if (W == 4)
                    ottt_words<4, TB, true, true, T, A>(acc, E, tab, 2 * total, kv, le * W, 63, W, j0, 0, 64);
                else if (W == 2)
                    ottt_words<2, TB, true, true, T, A>(acc, E, tab, 2 * total, kv, le * W, 63, W, j0, 0, 64);
                else
                    ottt_words<1, TB, true, true, T, A>(acc, E, tab, 2 * total, kv, le * W, 63, W, j0, 0, 64);
#pragma unroll
                for (int t = 0; t < kOtttTables; t++)
                    if (t < TB) acc[t] = ottt_wave_sum(acc[t]);
                if (lane == 0) ottt_write<T, A>(outE, outT, e, lutNum, t0, TB, party, writeE, acc, E);
                E = 0;
#pragma unroll
                for (int t = 0; t < kOtttTables; t++) acc[t] = 0;
            }
        }
        if (!more) break;
        u = un, blk = bn, kv = kvn, xv = xvn;
    }
}

// n >= 6, the pass's tables staged in chunks of C = 2^logC < 2^n entries, with
// 64 entries of padding on each side.  The key bits whose index falls in chunk
// [c0, c0 + C) are the run b in [blo, blo + C), blo = x - c0 - C + 1 (mod 2^n):
// C/64 + 1 words from word blo / 64, the first and last masked to the run.
// Lane L's entry for bit L of the run's s-th word is at C - 1 + (blo mod 64)
// - L - 64 s (+64 padding); the masked-off lanes read the padding or entries
// they do not add.  Each wave evaluates kOtttEW elements per batch.
template <typename T, int TB>
__global__ __launch_bounds__(1024) void ottt_eval_chunked_kernel(T *__restrict__ outE, T *__restrict__ outT,
                                                                  const T *__restrict__ x,
                                                                  const uint8_t *__restrict__ key, int party,
                                                                  const T *__restrict__ lut, int lutNum, int t0, int n,
                                                                  int logC, size_t N, int writeE)
{
    using A = typename OtttAcc<T>::type;
    extern __shared__ __align__(16) unsigned char ottt_lds[];
    T *tab = (T *)ottt_lds;
    const uint32_t total = 1u << n, mask = total - 1, W = total >> 6, C = 1u << logC;
    const uint32_t stride = C + 128, nchunks = total >> logC, nw = C / 64 + 1;
    const uint32_t lane   = threadIdx.x & 63;
    const size_t   perKey = total / 8;
    const size_t   batch  = (size_t)(blockDim.x >> 6) * kOtttEW;
    const uint32_t first  = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) * kOtttEW);
    for (size_t base = (size_t)blockIdx.x * batch; base < N; base += (size_t)gridDim.x * batch)
    {
        A        acc[kOtttEW][kOtttTables];
        uint32_t E[kOtttEW], Ev[kOtttEW]; // Ev: per lane, E by wave sum at the end
        uint32_t xv = 0; // lane w < kOtttEW: x of the wave's element w
        if (lane < kOtttEW && base + first + lane < N) xv = (uint32_t)(uint64_t)x[base + first + lane];
#pragma unroll
        for (int w = 0; w < kOtttEW; w++)
        {
            E[w]  = 0;
            Ev[w] = 0;
#pragma unroll
            for (int t = 0; t < kOtttTables; t++) acc[w][t] = 0;
        }
        for (uint32_t c = 0; c < nchunks; c++)
        {
            const uint32_t c0 = c << logC;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < (uint32_t)TB * stride; i += blockDim.x)
            {
                const uint32_t t = i / stride, q = (i % stride - 64) & mask; // pads wrap around the table
                tab[i]           = lut[(size_t)(t0 + (int)t) * total + ((c0 + q) & mask)];
            }
            __syncthreads();
#pragma unroll
            for (int w = 0; w < kOtttEW; w++)
            {
                const size_t e = base + first + w;
                if (e >= N) continue;
                const uint8_t *kb  = key + e * perKey;
                const uint32_t xe  = ottt_readlane(xv, w) & mask;
                const uint32_t blo = (xe - c0 - C + 1) & mask; // first key bit of the run
                const uint32_t off = blo & 63, k0 = blo >> 6;
                const uint32_t j0  = C - 1 + off - lane + 64; // +64: padding
                // the run's first C/64 words in blocks of up to 64 (one coalesced load each; a
                // multiple of 8 words), the first masked to the run, then its last word alone
                for (uint32_t blk = 0; blk * 64 < nw - 1; blk++)
                {
                    const uint32_t cnt = min(64u, nw - 1 - blk * 64);
                    const uint64_t kv  = lane < cnt ? ottt_load_word(kb, (k0 + blk * 64 + lane) & (W - 1)) : 0;
                    ottt_words<8, TB, false, false, T, A>(acc[w], E[w], tab, stride, kv, 0, 63, cnt,
                                                                j0 - 64u * 64u * blk, blk == 0 ? off : 0, 64);
                    // the run's set bits, counted in the lanes that hold its words (the
                    // first word masked as above)
                    Ev[w] += (uint32_t)__popcll(blk == 0 && lane == 0 ? kv & (~0ull << off) : kv);
                }
                const uint64_t kl = lane == 0 ? ottt_load_word(kb, (k0 + nw - 1) & (W - 1)) : 0;
                ottt_words<1, TB, false, false, T, A>(acc[w], E[w], tab, stride, kl, 0, 63, 1,
                                                            j0 - 64u * (nw - 1), 0, off);
                Ev[w] += (uint32_t)__popcll(kl & ((1ull << off) - 1)); // off < 64
            }
        }
#pragma unroll
        for (int w = 0; w < kOtttEW; w++)
        {
#pragma unroll
            for (int t = 0; t < kOtttTables; t++)
                if (t < TB) acc[w][t] = ottt_wave_sum(acc[w][t]);
            E[w] = ottt_wave_sum(Ev[w]);
            const size_t e = base + first + w;
            if (lane == 0 && e < N) ottt_write<T, A>(outE, outT, e, lutNum, t0, TB, party, writeE, acc[w], E[w]);
        }
    }
}

// n < 6: a group of P = 2^n lanes per element, 64 / P elements per round; lane
// L first loads element L of the wave's 64-element tile (its 1..4 key bytes and
// x), and each round's groups take theirs by cross-lane moves.
template <typename T, int TB>
__global__ __launch_bounds__(256) void ottt_eval_small_kernel(T *outE, T *outT, const T *x, const uint8_t *key,
                                                               int party, const T *lut, int lutNum, int t0, int n,
                                                               size_t N, int writeE)
{
    constexpr int tb = TB;
    using A = typename OtttAcc<T>::type;
    extern __shared__ __align__(16) unsigned char ottt_lds[];
    T *tab = (T *)ottt_lds;
    const uint32_t total = 1u << n, mask = total - 1;
    ottt_stage(tab, lut, t0, tb, total, 0, n);
    __syncthreads();
    const int    P = 1 << n, G = 64 / P;
    const int    lane = threadIdx.x & 63, j = lane & (P - 1), g = lane / P;
    const size_t nwaves = (size_t)gridDim.x * (blockDim.x >> 6);
    for (size_t e0 = ((size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64; e0 < N; e0 += nwaves * 64)
    {
        uint32_t kv = 0, xv = 0;
        if (e0 + lane < N)
        {
            const uint8_t *kb = key + (e0 + lane) * (size_t)(total / 8);
            kv = n == 5 ? *(const uint32_t *)kb : n == 4 ? ((uint32_t)kb[0] | ((uint32_t)kb[1] << 8)) : kb[0];
            xv = (uint32_t)(uint64_t)x[e0 + lane];
        }
        for (int r = 0; r < P; r++)
        {
            const int      src  = r * G + g;
            const uint32_t bits = (uint32_t)__shfl((int)kv, src, 64);
            const uint32_t xe   = (uint32_t)__shfl((int)xv, src, 64) & mask;
            const bool     take = (bits >> j) & 1;
            const uint32_t idx  = (xe - (uint32_t)j) & mask;
            A              acc[kOtttTables];
#pragma unroll
            for (int t = 0; t < kOtttTables; t++)
            {
                acc[t] = 0;
                if (t < tb) acc[t] = take ? (A)tab[(uint32_t)t * total + idx] : (A)0;
            }
#pragma unroll
            for (int t = 0; t < kOtttTables; t++)
                if (t < tb)
                    for (int o = P / 2; o >= 1; o >>= 1)
                    {
                        if constexpr (sizeof(A) == 8)
                        {
                            const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)acc[t], o, 64);
                            const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(acc[t] >> 32), o, 64);
                            acc[t] += (A)lo | ((A)hi << 32);
                        }
                        else
                            acc[t] += (A)(uint32_t)__shfl_xor((int)acc[t], o, 64);
                    }
            const size_t e = e0 + (size_t)src;
            if (j == 0 && e < N)
                ottt_write<T, A>(outE, outT, e, lutNum, t0, tb, party, writeE, acc, (uint32_t)__popc(bits)); // all 2^n key bits
        }
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
    if (es > 8)
    {
        if (!aligned(alpha, 8)) return FAST_FSS_RUNTIME_ERROR;
        return wide_ottt_keygen(st, key, alpha, (int)n, es, N);
    }
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
    if (n > 30) return FAST_FSS_RUNTIME_ERROR; // keys of 2^31 bits and more per element are not served
    hipStream_t st = resolve_stream(cudaStreamPtr);
    if (es > 8)
    {
        if (!aligned(sharedOutE, 8) || !aligned(sharedOutT, 8) || !aligned(maskedX, 8) || !aligned(lookUpTable, 8))
            return FAST_FSS_RUNTIME_ERROR;
        return wide_ottt_eval(st, sharedOutE, sharedOutT, maskedX, key, partyId, lookUpTable, lutNum, (int)n, es, N);
    }
    // passes of up to kOtttTables tables; each pass's tables sit in LDS, in chunks of C entries when
    // they do not fit in kOtttLdsBytes
    for (size_t t0 = 0; t0 < lutNum || t0 == 0; t0 += kOtttTables)
    {
        const int tb   = (int)std::min<size_t>(kOtttTables, lutNum - t0);
        const int tbs = std::max(tb, 1);
        // n < 6: one copy per table; n >= 6: two copies (resident) when they fit, else chunks of
        // C >= 1024 entries plus 128 of padding
        int    logC = (int)n;
        size_t lds  = ((size_t)tbs << n) * es;
        if (n >= 6)
        {
            lds = ((size_t)tbs << (n + 1)) * es;
            if (lds > kOtttLdsBytes)
            {
                logC = (int)n - 1;
                while (((size_t)tbs * (((size_t)1 << logC) + 128)) * es > kOtttLdsBytes) logC--;
                lds = (size_t)tbs * (((size_t)1 << logC) + 128) * es;
            }
        }
        const int    we  = t0 == 0;
        const size_t cus = (size_t)device_cu_count();
#define FSS_OTTT_TB(KERNEL, GRID, BLOCK, ...)                                                               \
    switch (tb)                                                                                            \
    {                                                                                                      \
        case 0: FSS_DISPATCH(es, (KERNEL<T, 0><<<GRID, BLOCK, lds, st>>>(__VA_ARGS__))); break;             \
        case 1: FSS_DISPATCH(es, (KERNEL<T, 1><<<GRID, BLOCK, lds, st>>>(__VA_ARGS__))); break;             \
        case 2: FSS_DISPATCH(es, (KERNEL<T, 2><<<GRID, BLOCK, lds, st>>>(__VA_ARGS__))); break;             \
        case 3: FSS_DISPATCH(es, (KERNEL<T, 3><<<GRID, BLOCK, lds, st>>>(__VA_ARGS__))); break;             \
        default: FSS_DISPATCH(es, (KERNEL<T, 4><<<GRID, BLOCK, lds, st>>>(__VA_ARGS__))); break;            \
    }
#define FSS_OTTT_ARGS                                                                                      \
    (T *)sharedOutE, (T *)sharedOutT, (const T *)maskedX, (const uint8_t *)key, partyId, (const T *)lookUpTable, \
        (int)lutNum, (int)t0
        if (n < 6)
        {
            const size_t grid = std::min<size_t>((N + 255) / 256, cus * 8);
            FSS_OTTT_TB(ottt_eval_small_kernel, grid, 256, FSS_OTTT_ARGS, (int)n, N, we)
        }
        else if (logC == (int)n)
        {
            // one element per wave at a time, four waves per workgroup
            const size_t perCu = std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / lds));
            const size_t grid  = std::min<size_t>((N + 3) / 4, cus * perCu);
            FSS_OTTT_TB(ottt_eval_tile_kernel, grid, 256, FSS_OTTT_ARGS, (int)n, N, we)
        }
        else
        {
            const size_t batch = (size_t)(1024 / 64) * kOtttEW;
            const size_t grid  = std::min<size_t>((N + batch - 1) / batch, cus * 2);
            FSS_OTTT_TB(ottt_eval_chunked_kernel, grid, 1024, FSS_OTTT_ARGS, (int)n, logC, N, we)
        }
#undef FSS_OTTT_ARGS
#undef FSS_OTTT_TB
        if (lutNum == 0) break;
    }
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
