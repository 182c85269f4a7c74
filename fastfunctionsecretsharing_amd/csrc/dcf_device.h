// DCF tree walks on the device (reference src/impl/dcf.h:187-491).
//
// Key layout (reference src/impl/dcf.h:139-167), N elements, n = bitWidthIn:
//   [sCW: N x n x 16 bytes][per element: vCW[n*g] | tLCW | tRCW | lastCW[g]]
// The sCW region is one 16n-byte run per element; the tail is an AoS record of
// (n*g + 2 + g) group elements.
//
// Evaluation walks only the taken child at each level: the reference computes
// both children (4 AES blocks) but reads only {sL, vL} or {sR, vR}
// (src/impl/dcf.h:424,453-464), so two blocks with plaintexts (4x, 4x+1) and
// (4x+2, 4x+3) give identical shares at half the AES work.
#pragma once

#include "fss_device.h"

namespace fss {

template <typename T>
__device__ __forceinline__ T neg_if(T v, int party)
{
    return party ? (T)(0 - v) : v;
}

// One party's share for groupSize 1 (the hot path).  `s` is the raw 16-byte
// seed; `x` the masked input; outBytes = ceil(bitWidthOut / 8).
template <typename T>
__device__ __forceinline__ T dcf_eval_g1(const AesCtx &c, const uint8_t *sCW, const T *tail, T x, uint32_t s[4],
                                         int party, int n, int outBytes)
{
    s[0] &= ~1u;
    uint32_t t   = (uint32_t)party;
    x            = mod_bits<T>(x, n);
    T        acc = 0;
    const T  tl = tail[n], tr = tail[n + 1];
    uint32_t cw[4];
    load16(cw, sCW);
    T vcw = tail[0];
    for (int i = 0; i < n; i++)
    {
        const uint32_t xi = (uint32_t)(x >> (n - 1 - i)) & 1u;
        uint32_t       cwn[4];
        T              vcwn = 0;
        if (i + 1 < n)
        {
            load16(cwn, sCW + 16 * (i + 1));
            vcwn = tail[i + 1];
        }
        uint32_t blk[2][4];
        aes_encrypt_dcf_pair(c, s, xi, blk);
        const uint32_t m  = 0u - t;
        uint32_t       tn = blk[0][0] & 1u;
        s[0]              = (blk[0][0] & ~1u) ^ (cw[0] & m);
        s[1]              = blk[0][1] ^ (cw[1] & m);
        s[2]              = blk[0][2] ^ (cw[2] & m);
        s[3]              = blk[0][3] ^ (cw[3] & m);
        tn ^= (uint32_t)(((xi ? tr : tl) >> i) & 1) & t;
        T term = (T)((T)low_bytes(blk[1], outBytes) + (t ? vcw : (T)0));
        acc    = (T)(acc + neg_if<T>(term, party));
        t      = tn;
#pragma unroll
        for (int k = 0; k < 4; k++) cw[k] = cwn[k];
        vcw = vcwn;
    }
    T term = (T)((T)low_bytes(s, outBytes) + (t ? tail[n + 2] : (T)0));
    return (T)(acc + neg_if<T>(term, party));
}

// dcf_eval_g1 on a lane pair, for calls too small to fill the GPU, where one
// wave's instruction stream sets the time.  Both lanes run the key schedule;
// the even lane encrypts the taken child's seed block (4x, 4x+1) and the odd
// lane its value block (4x+2, 4x+3), so each lane issues one block per level
// instead of two.  The odd lane takes the seed block from its partner with a
// DPP quad permutation (a VALU operand modifier, no LDS), so both lanes carry
// the same seed and t bit down the tree; only the odd lane's sum is the share.
template <typename T>
__device__ __forceinline__ T dcf_eval_g1_pair(const AesCtx &c, const uint8_t *sCW, const T *tail, T x, uint32_t s[4],
                                              int party, int n, int outBytes, uint32_t odd)
{
    s[0] &= ~1u;
    uint32_t t   = (uint32_t)party;
    x            = mod_bits<T>(x, n);
    T        acc = 0;
    const T  tl = tail[n], tr = tail[n + 1];
    for (int i = 0; i < n; i++)
    {
        const uint32_t xi = (uint32_t)(x >> (n - 1 - i)) & 1u;
        uint32_t       cw[4];
        load16(cw, sCW + 16 * i);
        const T        vcw   = tail[i];
        const uint32_t p0[1] = {4u * xi + 2u * odd}, p2[1] = {4u * xi + 1u + 2u * odd};
        uint32_t       blk[1][4];
        aes_encrypt_lowpt<1>(c, s, p0, p2, blk);
        // quad_perm [0, 0, 2, 2]: every lane reads its pair's even lane
        uint32_t sb[4];
#pragma unroll
        for (int k = 0; k < 4; k++) sb[k] = (uint32_t)__builtin_amdgcn_mov_dpp((int)blk[0][k], 0xA0, 0xF, 0xF, false);
        const uint32_t m  = 0u - t;
        uint32_t       tn = sb[0] & 1u;
        s[0]              = (sb[0] & ~1u) ^ (cw[0] & m);
        s[1]              = sb[1] ^ (cw[1] & m);
        s[2]              = sb[2] ^ (cw[2] & m);
        s[3]              = sb[3] ^ (cw[3] & m);
        tn ^= (uint32_t)(((xi ? tr : tl) >> i) & 1) & t;
        T term = (T)((T)low_bytes(blk[0], outBytes) + (t ? vcw : (T)0));
        acc    = (T)(acc + neg_if<T>(term, party));
        t      = tn;
    }
    T term = (T)((T)low_bytes(s, outBytes) + (t ? tail[n + 2] : (T)0));
    return (T)(acc + neg_if<T>(term, party));
}

// dcf_eval_g1 on eight lanes per element, for calls too small to fill the
// GPU, where one wave's instruction issue sets the time.  Lanes 0-3 of the
// group (h = 0) encrypt the taken child's seed block (4x, 4x+1) and lanes 4-7
// (h = 1) its value block (4x+2, 4x+3), one column per lane (QuadAes); the
// seed block reaches the value lanes by a DPP row shift by 4.  Every lane
// carries its column of the seed and the t bit down the tree; lane 4 (h = 1,
// column 0) returns the share.  Same results as dcf_eval_g1.
template <typename T>
__device__ __forceinline__ T dcf_eval_g1_oct(const AesCtx &c, const QuadAes &qa, const uint8_t *sCW, const T *tail,
                                             T x, uint32_t s, int party, int n, int outBytes, uint32_t h, uint32_t q)
{
    if (q == 0) s &= ~1u;
    uint32_t t   = (uint32_t)party;
    x            = mod_bits<T>(x, n);
    T        acc = 0;
    const T  tl = tail[n], tr = tail[n + 1];
    // correction words one level ahead: this lane's column of level i's sCW
    uint32_t cwNext = *(const uint32_t *)(sCW + 4 * q);
    T        vNext  = tail[0];
    for (int i = 0; i < n; i++)
    {
        const uint32_t cw  = cwNext;
        const T        vcw = vNext;
        if (i + 1 < n)
        {
            cwNext = *(const uint32_t *)(sCW + 16 * (i + 1) + 4 * q);
            vNext  = tail[i + 1];
        }
        const uint32_t xi = (uint32_t)(x >> (n - 1 - i)) & 1u;
        // plaintext columns (4x + 2h, 0, 4x + 2h + 1, 0)
        const uint32_t pt  = (q & 1u) ? 0u : 4u * xi + 2u * h + (q >> 1);
        const uint32_t blk = qa.encrypt(c, s, pt);
        // the seed block's column q on both halves: lanes 4-7 read lanes 0-3
        const uint32_t fromSeed = (uint32_t)__builtin_amdgcn_update_dpp((int)blk, (int)blk, 0x114, 0xF, 0xF, false);
        const uint32_t sc       = h ? fromSeed : blk;
        const uint32_t c0       = quad_mov<FSS_QP(0, 0, 0, 0)>(sc);
        const uint32_t m        = 0u - t;
        uint32_t       tn       = c0 & 1u;
        s                       = ((q == 0) ? (sc & ~1u) : sc) ^ (cw & m);
        tn ^= (uint32_t)(((xi ? tr : tl) >> i) & 1) & t;
        // value block columns 0 and 1 (lane 4 and its right neighbour)
        const uint32_t v1   = quad_mov<FSS_QP(1, 1, 1, 1)>(blk);
        const uint32_t vb[4] = {blk, v1, 0u, 0u};
        T term = (T)((T)low_bytes(vb, outBytes) + (t ? vcw : (T)0));
        acc    = (T)(acc + neg_if<T>(term, party));
        t      = tn;
    }
    const uint32_t s1    = quad_mov<FSS_QP(1, 1, 1, 1)>(s);
    const uint32_t sb[4] = {s, s1, 0u, 0u};
    T term = (T)((T)low_bytes(sb, outBytes) + (t ? tail[n + 2] : (T)0));
    return (T)(acc + neg_if<T>(term, party));
}

// Groupsize-1 shares of E elements per lane with the correction words read in
// bursts of CH levels.  A lane's sCW run is 16n contiguous bytes and its vCW
// run n contiguous group elements; fetching CH levels back to back consumes a
// whole 16*CH-byte stretch of a cache line while it is resident, instead of
// touching each line once per level.  E > 1 interleaves independent elements
// for instruction-level parallelism.  Results equal dcf_eval_g1 per element.
template <typename T, int CH, int E>
__device__ __forceinline__ void dcf_eval_g1_burst(const AesCtx &c, T out[E], const uint8_t *const sCW[E],
                                                  const T *const tail[E], const T xin[E], uint32_t s[E][4], int party,
                                                  int n, int outBytes)
{
    uint32_t t[E];
    T        acc[E], x[E], tl[E], tr[E];
#pragma unroll
    for (int e = 0; e < E; e++)
    {
        s[e][0] &= ~1u;
        t[e]   = (uint32_t)party;
        x[e]   = mod_bits<T>(xin[e], n);
        acc[e] = 0;
        tl[e]  = tail[e][n];
        tr[e]  = tail[e][n + 1];
    }
    for (int i0 = 0; i0 < n; i0 += CH)
    {
        uint32_t cw[E][CH][4];
        T        vcw[E][CH];
#pragma unroll
        for (int e = 0; e < E; e++)
        {
#pragma unroll
            for (int k = 0; k < CH; k++)
            {
                if (i0 + k < n)
                {
                    load16(cw[e][k], sCW[e] + 16 * (i0 + k));
                    vcw[e][k] = tail[e][i0 + k];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < CH; k++)
        {
            const int i = i0 + k;
            if (i < n)
            {
                uint32_t blk[E][2][4];
                uint32_t xi[E];
#pragma unroll
                for (int e = 0; e < E; e++)
                {
                    xi[e]        = (uint32_t)(x[e] >> (n - 1 - i)) & 1u;
                    blk[e][0][0] = 4u * xi[e], blk[e][0][1] = 0u, blk[e][0][2] = 4u * xi[e] + 1u, blk[e][0][3] = 0u;
                    blk[e][1][0] = 4u * xi[e] + 2u, blk[e][1][1] = 0u, blk[e][1][2] = 4u * xi[e] + 3u,
                    blk[e][1][3] = 0u;
                }
                if constexpr (E == 1)
                    aes_encrypt_dcf_pair(c, s[0], xi[0], blk[0]);
                else
                    aes_encrypt_multi<E, 2>(c, s, blk);
#pragma unroll
                for (int e = 0; e < E; e++)
                {
                    const uint32_t m  = 0u - t[e];
                    uint32_t       tn = blk[e][0][0] & 1u;
                    s[e][0]           = (blk[e][0][0] & ~1u) ^ (cw[e][k][0] & m);
                    s[e][1]           = blk[e][0][1] ^ (cw[e][k][1] & m);
                    s[e][2]           = blk[e][0][2] ^ (cw[e][k][2] & m);
                    s[e][3]           = blk[e][0][3] ^ (cw[e][k][3] & m);
                    tn ^= (uint32_t)(((xi[e] ? tr[e] : tl[e]) >> i) & 1) & t[e];
                    T term = (T)((T)low_bytes(blk[e][1], outBytes) + (t[e] ? vcw[e][k] : (T)0));
                    acc[e] = (T)(acc[e] + neg_if<T>(term, party));
                    t[e]   = tn;
                }
            }
        }
    }
#pragma unroll
    for (int e = 0; e < E; e++)
    {
        T term = (T)((T)low_bytes(s[e], outBytes) + (t[e] ? tail[e][n + 2] : (T)0));
        out[e] = (T)(acc[e] + neg_if<T>(term, party));
    }
}

// dcf_eval_g1_burst for one element whose value corrections (n of them, plus
// tLCW, tRCW, lastCW) are all loaded when the walk starts, for elements of at
// most 32 bits.  A wave's tails are then read once, as whole lines, instead
// of once per burst: the bursts' 4-byte-per-level reads of neighbouring
// 4(n+3)-byte tails touch the same lines again after they have left the
// cache.  The seed corrections still come in bursts of CH levels (one 128-B
// line per burst at CH = 8).  The level loop is not unrolled, so the window
// of corrections shifts down by CH after each burst to keep register indices
// static.  Results equal dcf_eval_g1.
template <typename T, int CH>
__device__ __forceinline__ T dcf_eval_g1_pre(const AesCtx &c, const uint8_t *sCW, const T *tail, T xin,
                                             uint32_t s[4], int party, int n, int outBytes)
{
    static_assert(sizeof(T) <= 4, "preloaded corrections: elements of at most 32 bits");
    constexpr int NMAX = 8 * sizeof(T);
    T             vall[NMAX];
#pragma unroll
    for (int j = 0; j < NMAX; j++) vall[j] = (j < n) ? tail[j] : (T)0;
    const T tl = tail[n], tr = tail[n + 1], lastCW = tail[n + 2];
    s[0] &= ~1u;
    uint32_t t   = (uint32_t)party;
    const T  x   = mod_bits<T>(xin, n);
    T        acc = 0;
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
                const uint32_t xi = (uint32_t)(x >> (n - 1 - i)) & 1u;
                uint32_t       blk[2][4];
                aes_encrypt_dcf_pair(c, s, xi, blk);
                const uint32_t m  = 0u - t;
                uint32_t       tn = blk[0][0] & 1u;
                s[0]              = (blk[0][0] & ~1u) ^ (cw[k][0] & m);
                s[1]              = blk[0][1] ^ (cw[k][1] & m);
                s[2]              = blk[0][2] ^ (cw[k][2] & m);
                s[3]              = blk[0][3] ^ (cw[k][3] & m);
                tn ^= (uint32_t)(((xi ? tr : tl) >> i) & 1) & t;
                T term = (T)((T)low_bytes(blk[1], outBytes) + (t ? vall[k] : (T)0));
                acc    = (T)(acc + neg_if<T>(term, party));
                t      = tn;
            }
        }
#pragma unroll
        for (int j = 0; j < NMAX; j++) vall[j] = (j + CH < NMAX) ? vall[j + CH] : (T)0;
    }
    T term = (T)((T)low_bytes(s, outBytes) + (t ? lastCW : (T)0));
    return (T)(acc + neg_if<T>(term, party));
}

// Share for group elements [j0, j0 + cnt) of a groupSize-g DCF; cnt <= R.
// Used for g > 1 (R covers the group) and by the spline evaluator, which
// needs only the coefficient ranges of one or two intervals.
template <typename T, int R>
__device__ __forceinline__ void dcf_eval_range(const AesCtx &c, T acc[R], const uint8_t *sCW, const T *tail, T x,
                                               const uint32_t seed[4], int party, int n, int m, int g, int j0,
                                               int cnt)
{
    uint32_t s[4] = {seed[0] & ~1u, seed[1], seed[2], seed[3]};
    uint32_t t    = (uint32_t)party;
    x             = mod_bits<T>(x, n);
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0;
    const T tl = tail[(size_t)n * g], tr = tail[(size_t)n * g + 1];
    for (int i = 0; i < n; i++)
    {
        const uint32_t xi = (uint32_t)(x >> (n - 1 - i)) & 1u;
        uint32_t       cw[4];
        load16(cw, sCW + 16 * i);
        uint32_t blk[2][4];
        aes_encrypt_dcf_pair(c, s, xi, blk);
        const uint32_t mk = 0u - t;
        uint32_t       tn = blk[0][0] & 1u;
        s[0]              = (blk[0][0] & ~1u) ^ (cw[0] & mk);
        s[1]              = blk[0][1] ^ (cw[1] & mk);
        s[2]              = blk[0][2] ^ (cw[2] & mk);
        s[3]              = blk[0][3] ^ (cw[3] & mk);
        tn ^= (uint32_t)(((xi ? tr : tl) >> i) & 1) & t;
        ConvCtx cv;
        cv.init(blk[1], g, m);
        const T *vcw = tail + (size_t)i * g + j0;
#pragma unroll
        for (int r = 0; r < R; r++)
        {
            if (r < cnt)
            {
                T term = (T)((T)cv.at(c, j0 + r) + (t ? vcw[r] : (T)0));
                acc[r] = (T)(acc[r] + neg_if<T>(term, party));
            }
        }
        t = tn;
    }
    ConvCtx cv;
    cv.init(s, g, m);
    const T *last = tail + (size_t)n * g + 2 + j0;
#pragma unroll
    for (int r = 0; r < R; r++)
    {
        if (r < cnt)
        {
            T term = (T)((T)cv.at(c, j0 + r) + (t ? last[r] : (T)0));
            acc[r] = (T)(acc[r] + neg_if<T>(term, party));
        }
    }
}

// dcf_eval_range for a compile-time output width OB = ceil(bitWidthOut/8):
// the slots' conversion shares one key schedule per level across its AES
// windows and extracts bytes at static positions (SlotConv).  Requires
// cnt <= R and cnt * OB <= 16 * (NW - 1) (a misaligned run of slots touches
// one window more).  Same results as dcf_eval_range.
template <typename T, int OB, int R, int NW = 3>
__device__ __forceinline__ void dcf_eval_slots(const AesCtx &c, T acc[R], const uint8_t *sCW, const T *tail, T x,
                                               const uint32_t seed[4], int party, int n, int g, int j0, int cnt)
{
    constexpr int NWORDS = (R * OB + 3) / 4;
    SlotConv<OB, NW> cv;
    cv.setup(g, j0, cnt);
    uint32_t s[4] = {seed[0] & ~1u, seed[1], seed[2], seed[3]};
    uint32_t t    = (uint32_t)party;
    x             = mod_bits<T>(x, n);
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0;
    const T tl = tail[(size_t)n * g], tr = tail[(size_t)n * g + 1];
    for (int i = 0; i < n; i++)
    {
        const uint32_t xi = (uint32_t)(x >> (n - 1 - i)) & 1u;
        uint32_t       cw[4];
        load16(cw, sCW + 16 * i);
        uint32_t blk[2][4];
        aes_encrypt_dcf_pair(c, s, xi, blk);
        const uint32_t mk = 0u - t;
        uint32_t       tn = blk[0][0] & 1u;
        s[0]              = (blk[0][0] & ~1u) ^ (cw[0] & mk);
        s[1]              = blk[0][1] ^ (cw[1] & mk);
        s[2]              = blk[0][2] ^ (cw[2] & mk);
        s[3]              = blk[0][3] ^ (cw[3] & mk);
        tn ^= (uint32_t)(((xi ? tr : tl) >> i) & 1) & t;
        uint32_t words[NWORDS];
        cv.template fill<NWORDS>(c, blk[1], words);
        const T *vcw = tail + (size_t)i * g + j0;
#pragma unroll
        for (int r = 0; r < R; r++)
        {
            if (r < cnt)
            {
                T term = (T)((T)SlotConv<OB, NW>::template at<NWORDS>(words, r) + (t ? vcw[r] : (T)0));
                acc[r] = (T)(acc[r] + neg_if<T>(term, party));
            }
        }
        t = tn;
    }
    uint32_t words[NWORDS];
    cv.template fill<NWORDS>(c, s, words);
    const T *last = tail + (size_t)n * g + 2 + j0;
#pragma unroll
    for (int r = 0; r < R; r++)
    {
        if (r < cnt)
        {
            T term = (T)((T)SlotConv<OB, NW>::template at<NWORDS>(words, r) + (t ? last[r] : (T)0));
            acc[r] = (T)(acc[r] + neg_if<T>(term, party));
        }
    }
}

// Any groupSize: accumulate straight into the element's output slice.
template <typename T>
__device__ void dcf_eval_global(const AesCtx &c, T *out, const uint8_t *sCW, const T *tail, T x,
                                const uint32_t seed[4], int party, int n, int m, int g)
{
    uint32_t s[4] = {seed[0] & ~1u, seed[1], seed[2], seed[3]};
    uint32_t t    = (uint32_t)party;
    x             = mod_bits<T>(x, n);
    for (int j = 0; j < g; j++) out[j] = 0;
    const T tl = tail[(size_t)n * g], tr = tail[(size_t)n * g + 1];
    for (int i = 0; i < n; i++)
    {
        const uint32_t xi = (uint32_t)(x >> (n - 1 - i)) & 1u;
        uint32_t       cw[4];
        load16(cw, sCW + 16 * i);
        uint32_t blk[2][4];
        aes_encrypt_dcf_pair(c, s, xi, blk);
        const uint32_t mk = 0u - t;
        uint32_t       tn = blk[0][0] & 1u;
        s[0]              = (blk[0][0] & ~1u) ^ (cw[0] & mk);
        s[1]              = blk[0][1] ^ (cw[1] & mk);
        s[2]              = blk[0][2] ^ (cw[2] & mk);
        s[3]              = blk[0][3] ^ (cw[3] & mk);
        tn ^= (uint32_t)(((xi ? tr : tl) >> i) & 1) & t;
        ConvCtx cv;
        cv.init(blk[1], g, m);
        const T *vcw = tail + (size_t)i * g;
        for (int j = 0; j < g; j++)
        {
            T term = (T)((T)cv.at(c, j) + (t ? vcw[j] : (T)0));
            out[j] = (T)(out[j] + neg_if<T>(term, party));
        }
        t = tn;
    }
    ConvCtx cv;
    cv.init(s, g, m);
    const T *last = tail + (size_t)n * g + 2;
    for (int j = 0; j < g; j++)
    {
        T term = (T)((T)cv.at(c, j) + (t ? last[j] : (T)0));
        out[j] = (T)(out[j] + neg_if<T>(term, party));
    }
}

// Key generation for one element (reference src/impl/dcf.h:187-368).
// `beta` == nullptr means beta = (1, 0, ...).  lastCW doubles as the running
// vAlpha accumulator, as in the reference.
// groupSize-1 keyGen (bitWidthOut <= 64, so the value conversion is the low
// ceil(bitWidthOut/8) bytes of a block): the two parties' blocks are produced
// one party at a time and only the values each level needs are kept, so the
// walk stays in registers; the running lastCW correction is a register, and
// the correction words of CH levels are stored in one burst per lane.
// Same outputs as dcf_keygen_one with g = 1 (src/impl/dcf.h:187-368).
template <typename T, int CH>
__device__ void dcf_keygen_g1(const AesCtx &c, uint8_t *sCW, T *tail, T alpha, T beta, const uint32_t seed0[4],
                              const uint32_t seed1[4], int n, int m)
{
    const int ob    = (m + 7) / 8;
    uint32_t  s0[4] = {seed0[0] & ~1u, seed0[1], seed0[2], seed0[3]};
    uint32_t  s1[4] = {seed1[0] & ~1u, seed1[1], seed1[2], seed1[3]};
    uint32_t  t0 = 0, t1 = 1;
    T         tlBits = 0, trBits = 0, last = 0;
    for (int i0 = 0; i0 < n; i0 += CH)
    {
        uint32_t cwb[CH][4];
        T        vcw[CH];
#pragma unroll
        for (int k = 0; k < CH; k++)
        {
            const int i = i0 + k;
            if (i >= n) break;
            const uint32_t keep = (uint32_t)(alpha >> (n - 1 - i)) & 1u;
            uint32_t       b[4][4];
            uint32_t       sk0[4], sl0[4];
            aes_encrypt_prg4(c, s0, b);
            const uint32_t tL0 = b[0][0] & 1u, tR0 = b[2][0] & 1u;
            b[0][0] &= ~1u, b[2][0] &= ~1u;
#pragma unroll
            for (int w = 0; w < 4; w++) sk0[w] = keep ? b[2][w] : b[0][w], sl0[w] = keep ? b[0][w] : b[2][w];
            const T vl0 = (T)low_bytes(keep ? b[1] : b[3], ob), vk0 = (T)low_bytes(keep ? b[3] : b[1], ob);
            aes_encrypt_prg4(c, s1, b);
            const uint32_t tL1 = b[0][0] & 1u, tR1 = b[2][0] & 1u;
            b[0][0] &= ~1u, b[2][0] &= ~1u;
            const T vl1 = (T)low_bytes(keep ? b[1] : b[3], ob), vk1 = (T)low_bytes(keep ? b[3] : b[1], ob);
            uint32_t cw[4], sk1[4];
#pragma unroll
            for (int w = 0; w < 4; w++)
            {
                sk1[w] = keep ? b[2][w] : b[0][w];
                cw[w]  = sl0[w] ^ (keep ? b[0][w] : b[2][w]);
                cwb[k][w] = cw[w];
            }
            const T sign1 = t1 ? (T)-1 : (T)1;
            T       v     = (T)((T)(vl1 - vl0 - last) * sign1);
            if (keep) v = (T)(v + (T)(sign1 * beta)); // the lost child is the left one
            vcw[k] = v;
            last   = (T)(last - vk1 + vk0 + (T)(sign1 * v));
            const uint32_t tLCW = tL0 ^ tL1 ^ keep ^ 1u;
            const uint32_t tRCW = tR0 ^ tR1 ^ keep;
            tlBits |= (T)((T)tLCW << i);
            trBits |= (T)((T)tRCW << i);
            const uint32_t m0 = 0u - t0, m1 = 0u - t1;
#pragma unroll
            for (int w = 0; w < 4; w++)
            {
                s0[w] = sk0[w] ^ (cw[w] & m0);
                s1[w] = sk1[w] ^ (cw[w] & m1);
            }
            const uint32_t tk0 = keep ? tR0 : tL0, tk1 = keep ? tR1 : tL1, tkc = keep ? tRCW : tLCW;
            t0 = tk0 ^ (t0 & tkc);
            t1 = tk1 ^ (t1 & tkc);
        }
#pragma unroll
        for (int k = 0; k < CH; k++)
        {
            if (i0 + k >= n) break;
            store16(sCW + 16 * (i0 + k), cwb[k]);
            tail[i0 + k] = vcw[k];
        }
    }
    tail[n]       = tlBits;
    tail[n + 1]   = trBits;
    const T sign1 = t1 ? (T)-1 : (T)1;
    tail[n + 2]   = (T)((T)((T)(0 - last) + (T)low_bytes(s1, ob) - (T)low_bytes(s0, ob)) * sign1);
}

template <typename T>
__device__ void dcf_keygen_one(const AesCtx &c, uint8_t *sCW, T *tail, T alpha, const T *beta,
                               const uint32_t seed0[4], const uint32_t seed1[4], int n, int m, int g)
{
    T *vcw  = tail;
    T *last = tail + (size_t)n * g + 2;
    for (int j = 0; j < g; j++) last[j] = 0;
    uint32_t s0[4] = {seed0[0] & ~1u, seed0[1], seed0[2], seed0[3]};
    uint32_t s1[4] = {seed1[0] & ~1u, seed1[1], seed1[2], seed1[3]};
    uint32_t t0 = 0, t1 = 1;
    T        tlBits = 0, trBits = 0;
    for (int i = 0; i < n; i++)
    {
        uint32_t b0[4][4], b1[4][4];
        aes_encrypt_prg4(c, s0, b0);
        aes_encrypt_prg4(c, s1, b1);
        const uint32_t tL0 = b0[0][0] & 1u, tR0 = b0[2][0] & 1u, tL1 = b1[0][0] & 1u, tR1 = b1[2][0] & 1u;
        b0[0][0] &= ~1u, b0[2][0] &= ~1u, b1[0][0] &= ~1u, b1[2][0] &= ~1u;
        const int keep = (int)((alpha >> (n - 1 - i)) & 1);
        const int lose = 1 - keep;
        uint32_t  cw[4];
#pragma unroll
        for (int k = 0; k < 4; k++) cw[k] = (lose ? b0[2][k] ^ b1[2][k] : b0[0][k] ^ b1[0][k]);
        store16(sCW + 16 * i, cw);

        const T sign1 = t1 ? (T)-1 : (T)1;
        ConvCtx cl1, cl0, ck1, ck0;
        uint32_t vl0[4], vl1[4], vk0[4], vk1[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
        {
            vl0[k] = lose ? b0[3][k] : b0[1][k];
            vl1[k] = lose ? b1[3][k] : b1[1][k];
            vk0[k] = keep ? b0[3][k] : b0[1][k];
            vk1[k] = keep ? b1[3][k] : b1[1][k];
        }
        cl1.init(vl1, g, m), cl0.init(vl0, g, m), ck1.init(vk1, g, m), ck0.init(vk0, g, m);
        for (int j = 0; j < g; j++)
        {
            T v = (T)((T)cl1.at(c, j) - (T)cl0.at(c, j) - last[j]);
            v   = (T)(v * sign1);
            if (lose == 0)
            {
                T b = beta ? beta[j] : (T)(j == 0 ? 1 : 0);
                v   = (T)(v + (T)(sign1 * b));
            }
            vcw[(size_t)i * g + j] = v;
            last[j]                = (T)(last[j] - (T)ck1.at(c, j) + (T)ck0.at(c, j) + (T)(sign1 * v));
        }
        const uint32_t tLCW = tL0 ^ tL1 ^ (uint32_t)keep ^ 1u;
        const uint32_t tRCW = tR0 ^ tR1 ^ (uint32_t)keep;
        tlBits |= (T)((T)tLCW << i);
        trBits |= (T)((T)tRCW << i);
        const uint32_t m0 = 0u - t0, m1 = 0u - t1;
#pragma unroll
        for (int k = 0; k < 4; k++)
        {
            s0[k] = (keep ? b0[2][k] : b0[0][k]) ^ (cw[k] & m0);
            s1[k] = (keep ? b1[2][k] : b1[0][k]) ^ (cw[k] & m1);
        }
        const uint32_t tk0 = keep ? tR0 : tL0, tk1 = keep ? tR1 : tL1, tkc = keep ? tRCW : tLCW;
        t0 = tk0 ^ (t0 & tkc);
        t1 = tk1 ^ (t1 & tkc);
    }
    tail[(size_t)n * g]     = tlBits;
    tail[(size_t)n * g + 1] = trBits;
    const T sign1 = t1 ? (T)-1 : (T)1;
    ConvCtx c1, c0;
    c1.init(s1, g, m), c0.init(s0, g, m);
    for (int j = 0; j < g; j++)
    {
        T v     = (T)((T)(0 - last[j]) + (T)c1.at(c, j) - (T)c0.at(c, j));
        last[j] = (T)(v * sign1);
    }
}

// Slot r of a converted chunk: bytes [r*OB, (r+1)*OB) of the windows w[]
// (r static, OB in {1, 2, 4, 8}).
template <typename T, int OB, int NWC>
__device__ __forceinline__ T chunk_slot(const uint32_t w[NWC][4], int r)
{
    const int o = r * OB, q = o >> 4, wd = (o & 15) >> 2;
    if constexpr (OB == 8)
        return (T)((uint64_t)w[q][wd] | ((uint64_t)w[q][wd + 1] << 32));
    else if constexpr (OB == 4)
        return (T)w[q][wd];
    else
        return (T)((w[q][wd] >> (8 * (o & 3))) & ((1u << (8 * OB)) - 1u));
}

// Converted chunk of slots [j0, j0 + SPC) under the keys ka, kb together: the
// value blocks themselves when the conversion is direct (OB*g <= 16, j0 = 0),
// else the AES windows w = j0*OB/16 .. + NWC of each key, plaintext
// (0, 16w/OB) (reference DcfConvertCtx, src/impl/dcf.h:69-137).
template <int OB, int NWC>
__device__ __forceinline__ void convert_chunk_pair(const AesCtx &c, uint32_t out[2][NWC][4], const uint32_t ka[4],
                                                   const uint32_t kb[4], bool direct, int j0)
{
    if (direct)
    {
#pragma unroll
        for (int k = 0; k < NWC; k++)
#pragma unroll
            for (int i = 0; i < 4; i++) out[0][k][i] = k ? 0u : ka[i], out[1][k][i] = k ? 0u : kb[i];
        return;
    }
    const int w0 = j0 * OB / 16;
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int k = 0; k < NWC; k++)
            out[p][k][0] = 0u, out[p][k][1] = 0u, out[p][k][2] = (uint32_t)((16 * (w0 + k)) / OB), out[p][k][3] = 0u;
    const uint32_t keys[2][4] = {{ka[0], ka[1], ka[2], ka[3]}, {kb[0], kb[1], kb[2], kb[3]}};
    aes_encrypt_multi<2, NWC>(c, keys, out);
}

// dcf_keygen_one for OB = ceil(bitWidthOut/8) in {1, 2, 4, 8} (compile time),
// slot chunk by slot chunk: the tree walk is repeated for each chunk of SPC
// group slots (NWC conversion windows), so the chunk's running lastCW values
// stay in registers for the whole walk and every value correction is written
// once, instead of reading and rewriting all g lastCW values at every level.
// The first pass also writes the seed and control-bit corrections.  Each
// chunk's windows share one key schedule per key, and the two keys a
// difference needs are encrypted together.  Same key bytes as dcf_keygen_one
// (reference src/impl/dcf.h:187-368).
template <typename T, int OB>
__device__ void dcf_keygen_chunked(const AesCtx &c, uint8_t *sCW, T *tail, T alpha, const T *beta,
                                   const uint32_t seed0[4], const uint32_t seed1[4], int n, int g)
{
    constexpr int NWC    = OB < 4 ? OB : 4;
    constexpr int SPC    = 16 * NWC / OB;
    const bool    direct = OB * g <= 16;
    T            *vcw    = tail;
    T            *lastCW = tail + (size_t)n * g + 2;
    for (int j0 = 0; j0 < g; j0 += SPC)
    {
        const bool first = j0 == 0;
        T          last[SPC], bj[SPC];
#pragma unroll
        for (int r = 0; r < SPC; r++)
        {
            const int j = j0 + r;
            last[r]     = 0;
            bj[r]       = (j < g) ? (beta ? beta[j] : (T)(j == 0 ? 1 : 0)) : (T)0;
        }
        uint32_t s0[4] = {seed0[0] & ~1u, seed0[1], seed0[2], seed0[3]};
        uint32_t s1[4] = {seed1[0] & ~1u, seed1[1], seed1[2], seed1[3]};
        uint32_t t0 = 0, t1 = 1;
        T        tlBits = 0, trBits = 0;
        for (int i = 0; i < n; i++)
        {
            const int keep = (int)((alpha >> (n - 1 - i)) & 1);
            uint32_t  vl0[4], vl1[4], vk0[4], vk1[4];
            const T   sign1 = t1 ? (T)-1 : (T)1;
            {
                // per party: the lost seed, kept seed, both value blocks and
                // the t bits
                const uint32_t km = 0u - (uint32_t)keep;
                uint32_t       sl[2][4], sk[2][4], tl[2], tr[2];
#pragma unroll
                for (int p = 0; p < 2; p++)
                {
                    const uint32_t *sp = p ? s1 : s0;
                    uint32_t        b[4][4];
                    aes_encrypt_prg4(c, sp, b);
                    tl[p] = b[0][0] & 1u, tr[p] = b[2][0] & 1u;
                    b[0][0] &= ~1u, b[2][0] &= ~1u;
                    uint32_t *vl = p ? vl1 : vl0, *vk = p ? vk1 : vk0;
#pragma unroll
                    for (int k = 0; k < 4; k++)
                    {
                        // blocks: 0 = left seed, 1 = left value, 2 = right seed, 3 = right value
                        sk[p][k] = (b[2][k] & km) | (b[0][k] & ~km);
                        sl[p][k] = (b[0][k] & km) | (b[2][k] & ~km);
                        vl[k]    = (b[1][k] & km) | (b[3][k] & ~km);
                        vk[k]    = (b[3][k] & km) | (b[1][k] & ~km);
                    }
                }
                uint32_t cw[4];
#pragma unroll
                for (int k = 0; k < 4; k++) cw[k] = sl[0][k] ^ sl[1][k];
                if (first) store16(sCW + 16 * i, cw);
                const uint32_t tLCW = tl[0] ^ tl[1] ^ (uint32_t)keep ^ 1u;
                const uint32_t tRCW = tr[0] ^ tr[1] ^ (uint32_t)keep;
                tlBits |= (T)((T)tLCW << i);
                trBits |= (T)((T)tRCW << i);
                const uint32_t m0 = 0u - t0, m1 = 0u - t1;
#pragma unroll
                for (int k = 0; k < 4; k++)
                {
                    s0[k] = sk[0][k] ^ (cw[k] & m0);
                    s1[k] = sk[1][k] ^ (cw[k] & m1);
                }
                const uint32_t tk0 = keep ? tr[0] : tl[0], tk1 = keep ? tr[1] : tl[1], tkc = keep ? tRCW : tLCW;
                t0 = tk0 ^ (t0 & tkc);
                t1 = tk1 ^ (t1 & tkc);
            }
            // this chunk's value corrections at level i (sign1 is the party-1
            // control bit on entering the level)
            uint32_t w[2][NWC][4];
            T        v[SPC];
            convert_chunk_pair<OB, NWC>(c, w, vl1, vl0, direct, j0);
#pragma unroll
            for (int r = 0; r < SPC; r++)
            {
                T x = (T)((T)(chunk_slot<T, OB, NWC>(w[0], r) - chunk_slot<T, OB, NWC>(w[1], r)) - last[r]);
                x   = (T)(x * sign1);
                if (keep) x = (T)(x + (T)(sign1 * bj[r]));
                v[r] = x;
                if (j0 + r < g) vcw[(size_t)i * g + j0 + r] = x;
            }
            convert_chunk_pair<OB, NWC>(c, w, vk1, vk0, direct, j0);
#pragma unroll
            for (int r = 0; r < SPC; r++)
                last[r] = (T)(last[r] - chunk_slot<T, OB, NWC>(w[0], r) + chunk_slot<T, OB, NWC>(w[1], r) +
                              (T)(sign1 * v[r]));
        }
        if (first)
        {
            tail[(size_t)n * g]     = tlBits;
            tail[(size_t)n * g + 1] = trBits;
        }
        const T  sign1 = t1 ? (T)-1 : (T)1;
        uint32_t w[2][NWC][4];
        convert_chunk_pair<OB, NWC>(c, w, s1, s0, direct, j0);
#pragma unroll
        for (int r = 0; r < SPC; r++)
            if (j0 + r < g)
                lastCW[j0 + r] = (T)((T)((T)(0 - last[r]) + chunk_slot<T, OB, NWC>(w[0], r) -
                                         chunk_slot<T, OB, NWC>(w[1], r)) *
                                     sign1);
    }
}

// Value-conversion width class for the chunked keyGen: ceil(bitWidthOut/8)
// when it is 1, 2, 4 or 8 bytes and fits the element, else 0 (generic path).
inline int keygen_chunk_ob(int m, size_t es)
{
    const int ob = (m + 7) / 8;
    return ((ob == 1 || ob == 2 || ob == 4 || ob == 8) && (size_t)ob <= es) ? ob : 0;
}

// dcf_keygen_chunked when OB > 0, dcf_keygen_one otherwise.
template <typename T, int OB>
__device__ __forceinline__ void dcf_keygen_group(const AesCtx &c, uint8_t *sCW, T *tail, T alpha, const T *beta,
                                                 const uint32_t seed0[4], const uint32_t seed1[4], int n, int m,
                                                 int g)
{
    if constexpr (OB == 0)
        dcf_keygen_one<T>(c, sCW, tail, alpha, beta, seed0, seed1, n, m, g);
    else
        dcf_keygen_chunked<T, OB>(c, sCW, tail, alpha, beta, seed0, seed1, n, g);
}

// Calls LAUNCH(OBV) with the compile-time width class of keygen_chunk_ob.
#define FSS_KEYGEN_OB_DISPATCH(T, OBRT, LAUNCH)                          \
    switch (OBRT)                                                        \
    {                                                                    \
        case 1: {                                                        \
            constexpr int OBV = 1;                                       \
            LAUNCH;                                                      \
        }                                                                \
        break;                                                           \
        case 2: {                                                        \
            constexpr int OBV = sizeof(T) >= 2 ? 2 : 0;                  \
            LAUNCH;                                                      \
        }                                                                \
        break;                                                           \
        case 4: {                                                        \
            constexpr int OBV = sizeof(T) >= 4 ? 4 : 0;                  \
            LAUNCH;                                                      \
        }                                                                \
        break;                                                           \
        case 8: {                                                        \
            constexpr int OBV = sizeof(T) >= 8 ? 8 : 0;                  \
            LAUNCH;                                                      \
        }                                                                \
        break;                                                           \
        default: {                                                       \
            constexpr int OBV = 0;                                       \
            LAUNCH;                                                      \
        }                                                                \
    }

} // namespace fss
