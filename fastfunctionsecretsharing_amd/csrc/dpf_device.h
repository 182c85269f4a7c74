// DPF and Grotto tree walks on the device (reference src/impl/dpf.h:172-380,
// src/impl/grotto.h:111-391).
//
// Both use the 2-block PRG G(s) = (AES_s(0,1), AES_s(2,3)) (little-endian u64
// counters).  Evaluation needs one child per level: the reference's dpfEval
// computes both and keeps one (src/impl/dpf.h:331,352-361), grottoEval already
// computes one (src/impl/grotto.h:338).
#pragma once

#include "fss_device.h"

namespace fss {

// One level of key generation shared by DPF (src/impl/dpf.h:201-272) and
// Grotto (src/impl/grotto.h:137-201): expand both parties' seeds, emit the
// correction word of the lost child and move both parties to the kept child.
__device__ __forceinline__ void tree_keygen_level(const AesCtx &c, uint32_t s0[4], uint32_t s1[4], uint32_t &t0,
                                                  uint32_t &t1, int keep, uint32_t cw[4], uint32_t &tLCW,
                                                  uint32_t &tRCW)
{
    uint32_t b0[2][4], b1[2][4];
    aes_encrypt_prg2(c, s0, b0);
    aes_encrypt_prg2(c, s1, b1);
    const uint32_t tL0 = b0[0][0] & 1u, tR0 = b0[1][0] & 1u, tL1 = b1[0][0] & 1u, tR1 = b1[1][0] & 1u;
    b0[0][0] &= ~1u, b0[1][0] &= ~1u, b1[0][0] &= ~1u, b1[1][0] &= ~1u;
    // children chosen by masks: a selected array index would place b0/b1 in
    // scratch memory
    const uint32_t km = 0u - (uint32_t)keep;
#pragma unroll
    for (int k = 0; k < 4; k++) cw[k] = ((b0[0][k] ^ b1[0][k]) & km) | ((b0[1][k] ^ b1[1][k]) & ~km);
    tLCW              = tL0 ^ tL1 ^ (uint32_t)keep ^ 1u;
    tRCW              = tR0 ^ tR1 ^ (uint32_t)keep;
    const uint32_t m0 = 0u - t0, m1 = 0u - t1;
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
        s0[k] = ((b0[1][k] & km) | (b0[0][k] & ~km)) ^ (cw[k] & m0);
        s1[k] = ((b1[1][k] & km) | (b1[0][k] & ~km)) ^ (cw[k] & m1);
    }
    const uint32_t tk0 = keep ? tR0 : tL0, tk1 = keep ? tR1 : tL1, tkc = keep ? tRCW : tLCW;
    t0 = tk0 ^ (t0 & tkc);
    t1 = tk1 ^ (t1 & tkc);
}

// One evaluation step on the path: child `bit` of (s, t) with corrections.
template <typename T>
__device__ __forceinline__ void tree_eval_step(const AesCtx &c, uint32_t s[4], uint32_t &t, uint32_t bit,
                                               const uint8_t *sCW, T tl, T tr, int i)
{
    uint32_t blk[1][4] = {{2u * bit, 0u, 2u * bit + 1u, 0u}};
    uint32_t cw[4];
    load16(cw, sCW);
    aes_encrypt<1>(c, s, blk);
    const uint32_t mk = 0u - t;
    uint32_t       tn = blk[0][0] & 1u;
    s[0]              = (blk[0][0] & ~1u) ^ (cw[0] & mk);
    s[1]              = blk[0][1] ^ (cw[1] & mk);
    s[2]              = blk[0][2] ^ (cw[2] & mk);
    s[3]              = blk[0][3] ^ (cw[3] & mk);
    tn ^= (uint32_t)(((bit ? tr : tl) >> i) & 1) & t;
    t = tn;
}

// ---- DPF ----
// key tail per element: tLCW | tRCW | lastCW[g] (src/impl/dpf.h:130-156)

template <typename T>
__device__ void dpf_keygen_one(const AesCtx &c, uint8_t *sCW, T *tail, T alpha, const T *beta,
                               const uint32_t seed0[4], const uint32_t seed1[4], int n, int m, int g)
{
    uint32_t s0[4] = {seed0[0] & ~1u, seed0[1], seed0[2], seed0[3]};
    uint32_t s1[4] = {seed1[0] & ~1u, seed1[1], seed1[2], seed1[3]};
    uint32_t t0 = 0, t1 = 1;
    T        tl = 0, tr = 0;
    // the correction words of 8 levels (one 128-B line) are stored together,
    // so each line is written whole rather than 16 B at a time
    for (int i0 = 0; i0 < n; i0 += 8)
    {
        uint32_t cw[8][4];
#pragma unroll
        for (int k = 0; k < 8; k++)
        {
            const int i = i0 + k;
            if (i < n)
            {
                uint32_t tLCW, tRCW;
                tree_keygen_level(c, s0, s1, t0, t1, (int)((alpha >> (n - 1 - i)) & 1), cw[k], tLCW, tRCW);
                tl |= (T)((T)tLCW << i);
                tr |= (T)((T)tRCW << i);
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (i0 + k < n) store16(sCW + 16 * (i0 + k), cw[k]);
    }
    tail[0] = tl;
    tail[1] = tr;
    // the reference passes (bitWidthOut, groupSize) into the (groupSize,
    // bitWidth) slots of its converter (src/impl/dpf.h:74,278,283)
    ConvCtx c0, c1;
    c0.init(s0, m, g);
    c1.init(s1, m, g);
    const T sign1 = t1 ? (T)-1 : (T)1;
    for (int j = 0; j < g; j++)
    {
        T b         = beta ? beta[j] : (T)1;
        T a0        = (T)c0.at(c, j);
        T a1        = (T)c1.at(c, j);
        tail[2 + j] = (T)(sign1 * (T)(b - a0 + a1));
    }
}

// Walk to the leaf of x; returns the final (s, t).
template <typename T>
__device__ __forceinline__ void dpf_walk(const AesCtx &c, uint32_t s[4], uint32_t &t, const uint8_t *sCW, T tl, T tr,
                                         T x, int n)
{
    for (int i = 0; i < n; i++)
        tree_eval_step<T>(c, s, t, (uint32_t)(x >> (n - 1 - i)) & 1u, sCW + 16 * i, tl, tr, i);
}

template <typename T>
__device__ __forceinline__ void dpf_eval_one(const AesCtx &c, T *out, const uint8_t *sCW, const T *tail, T x,
                                             const uint32_t seed[4], int party, int n, int m, int g)
{
    uint32_t s[4] = {seed[0] & ~1u, seed[1], seed[2], seed[3]};
    uint32_t t    = (uint32_t)party;
    dpf_walk<T>(c, s, t, sCW, tail[0], tail[1], mod_bits<T>(x, n), n);
    ConvCtx cv;
    cv.init(s, m, g);
    for (int j = 0; j < g; j++)
    {
        T term = (T)((T)cv.at(c, j) + (t ? tail[2 + j] : (T)0));
        out[j] = party ? (T)(0 - term) : term;
    }
}

// groupSize-1 DPF eval with the correction words of CH levels loaded in one
// burst before they are used (one 16-byte load per level per lane otherwise
// interleaves with the AES and lets L2 evict the lane's key line between
// levels).  The value conversion for groupSize 1 takes one byte (the
// reference's swapped ConvertCtx arguments, src/impl/dpf.h:74,278,283): byte 0
// of the final seed, or of AES_seed(0^128) when bitWidthOut > 16.  Results
// equal dpf_eval_one.
template <typename T, int CH>
__device__ __forceinline__ T dpf_eval_g1_burst(const AesCtx &c, const uint8_t *sCW, const T *tail, T xin,
                                               uint32_t s[4], int party, int n, int m)
{
    s[0] &= ~1u;
    uint32_t t  = (uint32_t)party;
    const T  x  = mod_bits<T>(xin, n);
    const T  tl = tail[0], tr = tail[1];
    for (int i0 = 0; i0 < n; i0 += CH)
    {
        uint32_t cw[CH][4];
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (i0 + k < n) load16(cw[k], sCW + 16 * (i0 + k));
#pragma unroll
        for (int k = 0; k < CH; k++)
        {
            const int i = i0 + k;
            if (i < n)
            {
                const uint32_t bit       = (uint32_t)(x >> (n - 1 - i)) & 1u;
                uint32_t       blk[1][4] = {{2u * bit, 0u, 2u * bit + 1u, 0u}};
                aes_encrypt<1>(c, s, blk);
                const uint32_t mk = 0u - t;
                uint32_t       tn = blk[0][0] & 1u;
                s[0]              = (blk[0][0] & ~1u) ^ (cw[k][0] & mk);
                s[1]              = blk[0][1] ^ (cw[k][1] & mk);
                s[2]              = blk[0][2] ^ (cw[k][2] & mk);
                s[3]              = blk[0][3] ^ (cw[k][3] & mk);
                tn ^= (uint32_t)(((bit ? tr : tl) >> i) & 1) & t;
                t = tn;
            }
        }
    }
    uint32_t conv;
    if (m <= 16)
        conv = s[0] & 0xffu;
    else
    {
        uint32_t blk[1][4] = {{0u, 0u, 0u, 0u}};
        aes_encrypt<1>(c, s, blk);
        conv = blk[0][0] & 0xffu;
    }
    const T term = (T)((T)conv + (t ? tail[2] : (T)0));
    return party ? (T)(0 - term) : term;
}

// dpf_eval_g1_burst on a quad of lanes per element, one column of the seed
// block per lane (QuadAes), for calls too small to fill the GPU.  Lane 0 of
// the quad returns the share.  Same results as dpf_eval_g1_burst.
template <typename T>
__device__ __forceinline__ T dpf_eval_g1_quad(const AesCtx &c, const QuadAes &qa, const uint8_t *sCW, const T *tail,
                                              T xin, uint32_t s, int party, int n, int m, uint32_t q)
{
    if (q == 0) s &= ~1u;
    uint32_t t  = (uint32_t)party;
    const T  x  = mod_bits<T>(xin, n);
    const T  tl = tail[0], tr = tail[1];
    // this lane's column of the level's correction word, one level ahead
    uint32_t cwNext = *(const uint32_t *)(sCW + 4 * q);
    for (int i = 0; i < n; i++)
    {
        const uint32_t cw = cwNext;
        if (i + 1 < n) cwNext = *(const uint32_t *)(sCW + 16 * (i + 1) + 4 * q);
        const uint32_t bit = (uint32_t)(x >> (n - 1 - i)) & 1u;
        // plaintext columns (2 bit, 0, 2 bit + 1, 0)
        const uint32_t pt  = (q & 1u) ? 0u : 2u * bit + (q >> 1);
        const uint32_t blk = qa.encrypt(c, s, pt);
        const uint32_t c0  = quad_mov<FSS_QP(0, 0, 0, 0)>(blk);
        const uint32_t mk  = 0u - t;
        uint32_t       tn  = c0 & 1u;
        s                  = ((q == 0) ? (blk & ~1u) : blk) ^ (cw & mk);
        tn ^= (uint32_t)(((bit ? tr : tl) >> i) & 1) & t;
        t = tn;
    }
    // byte 0 of the final seed, or of AES_seed(0^128) when bitWidthOut > 16
    const uint32_t conv = (m <= 16 ? s : qa.encrypt(c, s, 0u)) & 0xffu;
    const T        term = (T)((T)conv + (t ? tail[2] : (T)0));
    return party ? (T)(0 - term) : term;
}

// ---- Grotto ----
// key: [sCW N x (n-6) x 16][lastCW N x 8][tLCW, tRCW N x 2 GE] (src/impl/grotto.h:53-79)

template <typename T>
__device__ void grotto_keygen_one(const AesCtx &c, uint8_t *sCW, uint64_t *lastCW, T *tlr, T alpha,
                                  const uint32_t seed0[4], const uint32_t seed1[4], int n)
{
    uint32_t s0[4] = {seed0[0] & ~1u, seed0[1], seed0[2], seed0[3]};
    uint32_t s1[4] = {seed1[0] & ~1u, seed1[1], seed1[2], seed1[3]};
    uint32_t t0 = 0, t1 = 1;
    T        tl = 0, tr = 0;
    for (int i = 0; i < n - 6; i++)
    {
        uint32_t cw[4], tLCW, tRCW;
        tree_keygen_level(c, s0, s1, t0, t1, (int)((alpha >> (n - 1 - i)) & 1), cw, tLCW, tRCW);
        store16(sCW + 16 * i, cw);
        tl |= (T)((T)tLCW << i);
        tr |= (T)((T)tRCW << i);
    }
    tlr[0] = tl;
    tlr[1] = tr;
    const uint64_t h0 = (uint64_t)s0[2] | ((uint64_t)s0[3] << 32);
    const uint64_t h1 = (uint64_t)s1[2] | ((uint64_t)s1[3] << 32);
    *lastCW           = h0 ^ h1 ^ (1ull << ((uint64_t)alpha & 63));
}

__device__ __forceinline__ uint32_t parity64(uint64_t x)
{
    return (uint32_t)(__popcll(x) & 1);
}

// src/impl/grotto.h:207-282
template <typename T>
__device__ __forceinline__ T grotto_eq_eval_one(const AesCtx &c, const uint8_t *sCW, uint64_t lastCW, T tl, T tr, T x,
                                                const uint32_t seed[4], int party, int n)
{
    uint32_t s[4] = {seed[0] & ~1u, seed[1], seed[2], seed[3]};
    uint32_t t    = (uint32_t)party;
    x             = mod_bits<T>(x, n);
    dpf_walk<T>(c, s, t, sCW, tl, tr, (T)(x >> 6), n - 6);
    uint64_t u = ((uint64_t)s[2] | ((uint64_t)s[3] << 32)) ^ (t ? lastCW : 0ull);
    return (T)((u >> ((uint64_t)x & 63)) & 1ull);
}

// src/impl/grotto.h:284-391
template <typename T>
__device__ __forceinline__ T grotto_eval_one(const AesCtx &c, const uint8_t *sCW, uint64_t lastCW, T tl, T tr, T x,
                                             const uint32_t seed[4], int party, int n, bool equalBound)
{
    x             = mod_bits<T>(x, n);
    uint32_t s[4] = {seed[0] & ~1u, seed[1], seed[2], seed[3]};
    uint32_t t = (uint32_t)party, parity = 0;
    for (int i = 0; i < n - 6; i++)
    {
        // the remaining low n-i bits all zero: x is on the left spine below
        if (equalBound && mod_bits<T>(x, n - i) == 0) return (T)(parity ^ (uint32_t)party);
        const uint32_t bit  = (uint32_t)(x >> (n - 1 - i)) & 1u;
        const uint32_t prev = t;
        tree_eval_step<T>(c, s, t, bit, sCW + 16 * i, tl, tr, i);
        parity ^= bit & (prev ^ t);
    }
    const uint64_t u = ((uint64_t)s[2] | ((uint64_t)s[3] << 32)) ^ (t ? lastCW : 0ull);
    const uint32_t p = (uint32_t)((uint64_t)x & 63);
    if (equalBound)
    {
        if (p != 0) parity ^= parity64(u & (~0ull >> (64 - p)));
    }
    else
        parity ^= parity64(u & (~0ull >> (63 - p)));
    return (T)(parity ^ (uint32_t)party);
}

// grotto_eval_one (mode 0 lt, 1 le) and grotto_eq_eval_one (mode 2) on a quad
// of lanes per element, one column of the seed block per lane (QuadAes), for
// calls too small to fill the GPU.  Lane 0 of the quad returns the share;
// every lane of the quad takes the same branches (one element).
template <typename T>
__device__ __forceinline__ T grotto_eval_quad(const AesCtx &c, const QuadAes &qa, const uint8_t *sCW, uint64_t lastCW,
                                              T tl, T tr, T x, uint32_t s, int party, int n, int mode, uint32_t q)
{
    x = mod_bits<T>(x, n);
    if (q == 0) s &= ~1u;
    uint32_t   t = (uint32_t)party, parity = 0;
    const bool equalBound = mode == 1;
    const T    path       = mode == 2 ? (T)(x >> 6) : x; // eq walks the leaf index, lt/le x's top bits
    const int  top        = mode == 2 ? n - 6 : n;
    for (int i = 0; i < n - 6; i++)
    {
        // the remaining low n-i bits all zero: x is on the left spine below
        if (equalBound && mod_bits<T>(x, n - i) == 0) return (T)(parity ^ (uint32_t)party);
        const uint32_t bit  = (uint32_t)(path >> (top - 1 - i)) & 1u;
        const uint32_t prev = t;
        const uint32_t cw   = *(const uint32_t *)(sCW + 16 * i + 4 * q);
        // plaintext columns (2 bit, 0, 2 bit + 1, 0)
        const uint32_t pt  = (q & 1u) ? 0u : 2u * bit + (q >> 1);
        const uint32_t blk = qa.encrypt(c, s, pt);
        const uint32_t c0  = quad_mov<FSS_QP(0, 0, 0, 0)>(blk);
        const uint32_t mk  = 0u - t;
        uint32_t       tn  = c0 & 1u;
        s                  = ((q == 0) ? (blk & ~1u) : blk) ^ (cw & mk);
        tn ^= (uint32_t)(((bit ? tr : tl) >> i) & 1) & t;
        t = tn;
        parity ^= bit & (prev ^ t);
    }
    // the leaf word: seed columns 2 and 3
    const uint32_t s2 = quad_mov<FSS_QP(2, 2, 2, 2)>(s), s3 = quad_mov<FSS_QP(3, 3, 3, 3)>(s);
    const uint64_t u  = ((uint64_t)s2 | ((uint64_t)s3 << 32)) ^ (t ? lastCW : 0ull);
    const uint32_t p  = (uint32_t)((uint64_t)x & 63);
    if (mode == 2) return (T)((u >> p) & 1ull);
    if (equalBound)
    {
        if (p != 0) parity ^= parity64(u & (~0ull >> (64 - p)));
    }
    else
        parity ^= parity64(u & (~0ull >> (63 - p)));
    return (T)(parity ^ (uint32_t)party);
}

} // namespace fss
