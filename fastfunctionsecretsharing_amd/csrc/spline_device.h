// DCF-spline evaluation kernels for degree <= 7 (reference src/impl/spline.h,
// src/kernel/spline.h), shared by the library's spline launcher.
//
// The element's single DCF key (group size I*(degree+1)) is evaluated at the
// interval end points.  Each point needs only the coefficient slots of the
// interval(s) it bounds, so each point is walked once and only those group
// elements (and the AES conversion blocks that hold them) are produced; the
// values are identical to the reference's full-group evaluation.
#pragma once

#include "dcf_device.h"
#include "fss_host.h"

namespace fss {

// Fast path: degree <= 7.  Each evaluation point x-1-L_0, x-1-(R_q+1) is walked
// once; the walk at x-1-(R_q+1) also yields interval q+1's "current" slots when
// intervals q and q+1 are adjacent (the reference copies the whole vector,
// src/impl/spline.h:214-223).  OB > 0 selects the slot conversion specialised
// for ceil(bitWidthOut/8) = OB bytes (needs 2*(degree+1)*OB <= BY, the walk's
// byte budget: 32, or 48 for 8-byte slots); OB = 0 the generic one.  An
// element's intervals may be split into bpe blocks (a power of two dividing
// 64) on consecutive lanes of one wave, which then read the element's seed
// corrections together: each lane sums its terms of the polynomial and the
// bpe partial sums are added across the lanes.  The linked walk that ends a
// block also subtracts the next block's first "current" slots, so the split
// walks the same points as one lane (I + 1 for I adjacent intervals), and
// since every term enters the sum exactly once the result equals the one-lane
// walk's mod 2^m.
template <typename T, int OB, int BY>
__device__ __forceinline__ void spline_walk(const AesCtx &c, T *acc, const uint8_t *sc, const T *tail, T pt,
                                            const uint32_t s[4], int party, int n, int m, int g, int j0, int cnt)
{
    if constexpr (OB == 0)
        dcf_eval_range<T, 16>(c, acc, sc, tail, pt, s, party, n, m, g, j0, cnt);
    else
        dcf_eval_slots<T, OB, (BY / OB < 16 ? BY / OB : 16), BY / 16 + 1>(c, acc, sc, tail, pt, s, party, n, g, j0,
                                                                           cnt);
}

template <typename T, int OB, bool SPLIT, int BY>
__device__ __forceinline__ void spline_eval_body(const AesCtx &c, T *out, const T *x, const uint8_t *key, const T *sE,
                                                 const T *sBeta, const uint8_t *seed, int party, const T *left,
                                                 const T *right, size_t I, int d, int n, int m, size_t N, uint32_t bpe,
                                                 uint32_t *counter)
{
    constexpr int RS = (OB == 0) ? 16 : (BY / OB < 16 ? BY / OB : 16); // slots per walk
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
            // interval q's start slots come from the linked end walk of q - 1
            // (on this lane or, at a block boundary, on the previous one)
            if (!(q > 0 && mod_bits<T>((T)(right[q - 1] + 1), n) == L))
            {
                spline_walk<T, OB, BY>(c, acc, sc, tail, (T)(xv - 1 - L), s, party, n, m, (int)g, (int)(q * C), C);
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (j < C) w[j] = (T)(w[j] - acc[j < RS ? j : 0]);
            }
            const bool link = (q + 1 < I) && mod_bits<T>((T)(right[q] + 1), n) == mod_bits<T>(left[q + 1], n);
            spline_walk<T, OB, BY>(c, acc, sc, tail, (T)(xv - 1 - Rp), s, party, n, m, (int)g, (int)(q * C),
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

template <typename T, int OB, int BY = 32>
__global__ __launch_bounds__(kBlock) void spline_eval_kernel(T *out, const T *x, const uint8_t *key, const T *sE,
                                                             const T *sBeta, const uint8_t *seed, int party,
                                                             const T *left, const T *right, size_t I, int d, int n,
                                                             int m, size_t N, uint32_t *counter)
{
    FSS_KERNEL_PROLOGUE
    spline_eval_body<T, OB, false, BY>(c, out, x, key, sE, sBeta, seed, party, left, right, I, d, n, m, N, 1, counter);
}

// The same with each element's intervals split into bpe > 1 blocks: calls too
// small to give every CU a full workgroup, and calls of many intervals, whose
// 2-way split pays for its extra wave step even when every lane is busy.
template <typename T, int OB, int BY = 32>
__global__ __launch_bounds__(kBlock) void spline_eval_split_kernel(T *out, const T *x, const uint8_t *key,
                                                                   const T *sE, const T *sBeta, const uint8_t *seed,
                                                                   int party, const T *left, const T *right, size_t I,
                                                                   int d, int n, int m, size_t N, uint32_t bpe,
                                                                   uint32_t *counter)
{
    FSS_KERNEL_PROLOGUE_STAGED
    spline_eval_body<T, OB, true, BY>(c, out, x, key, sE, sBeta, seed, party, left, right, I, d, n, m, N, bpe, counter);
}

// ---- small calls: each walk on eight lanes (4-byte slots) ----
//
// One walk of a groupSize-g DCF at `pt` for slots [j0, j0 + cnt), on the
// eight lanes of a group (QuadAes): per level, lanes 0-3 (h = 0) encrypt the
// taken child's seed block and lanes 4-7 its value block; the value block,
// the conversion key, reaches lanes 0-3 by a DPP row shift, and the two quads
// encrypt conversion windows w0 and w0 + 1 (a third, when the slots touch it,
// on lanes 0-3).  With 4-byte slots a slot is one word of one window, so the
// lane holding that word accumulates the slot's share itself: slot r1 of the
// first pass, r2 of the second (-1: none).  Same sums as dcf_eval_slots.
struct OctSlots
{
    int w0, nw, b0w; // first window, windows, first slot's word in window w0
    int r1, r2;

    __device__ __forceinline__ void setup(int j0, int cnt, uint32_t h, uint32_t q)
    {
        const int start = 4 * j0, end = 4 * (j0 + cnt);
        w0              = start >> 4;
        nw              = ((end - 1) >> 4) - w0 + 1;
        b0w             = (start - 16 * w0) >> 2;
        const int a     = 4 * (int)h + (int)q - b0w;
        const int b     = 8 + (int)q - b0w;
        r1              = (a >= 0 && a < cnt) ? a : -1;
        r2              = (h == 0 && nw == 3 && b >= 0 && b < cnt) ? b : -1;
    }
};

template <typename T>
__device__ __forceinline__ void spline_walk_oct(const AesCtx &c, const QuadAes &qa, const OctSlots &sl, T &acc1,
                                                T &acc2, const uint8_t *sCW, const T *tail, T pt, uint32_t seedCol,
                                                int party, int n, int g, int j0, uint32_t h, uint32_t q)
{
    uint32_t s = (q == 0) ? (seedCol & ~1u) : seedCol;
    uint32_t t = (uint32_t)party;
    pt         = mod_bits<T>(pt, n);
    acc1 = 0, acc2 = 0;
    const T        tl = tail[(size_t)n * g], tr = tail[(size_t)n * g + 1];
    const uint32_t wpt0 = (q == 2) ? 4u * (uint32_t)(sl.w0 + (int)h) : 0u; // window plaintext (0, 0, 16w/4, 0)
    const uint32_t wpt2 = (q == 2) ? 4u * (uint32_t)(sl.w0 + 2) : 0u;
    for (int i = 0; i < n; i++)
    {
        const uint32_t xi  = (uint32_t)(pt >> (n - 1 - i)) & 1u;
        const uint32_t cw  = *(const uint32_t *)(sCW + 16 * i + 4 * q);
        const T       *vcw = tail + (size_t)i * g + j0;
        const T        v1  = sl.r1 >= 0 ? vcw[sl.r1] : (T)0;
        const T        v2  = sl.r2 >= 0 ? vcw[sl.r2] : (T)0;
        const uint32_t blk = qa.encrypt(c, s, (q & 1u) ? 0u : 4u * xi + 2u * h + (q >> 1));
        // seed block on both quads (lanes 4-7 read lanes 0-3), value block on
        // both quads (lanes 0-3 read lanes 4-7); the moves run on every lane
        // of the group before the selects (a DPP move under a partial exec
        // mask would read disabled lanes)
        const uint32_t fromA = (uint32_t)__builtin_amdgcn_update_dpp((int)blk, (int)blk, 0x114, 0xF, 0xF, false);
        const uint32_t fromB = (uint32_t)__builtin_amdgcn_update_dpp((int)blk, (int)blk, 0x104, 0xF, 0xF, false);
        const uint32_t seedB = h ? fromA : blk;
        const uint32_t valB  = h ? blk : fromB;
        const uint32_t c0    = quad_mov<FSS_QP(0, 0, 0, 0)>(seedB);
        const uint32_t mk    = 0u - t;
        uint32_t       tn    = c0 & 1u;
        tn ^= (uint32_t)(((xi ? tr : tl) >> i) & 1) & t;
        const uint32_t win = qa.encrypt(c, valB, wpt0);
        if (sl.r1 >= 0) acc1 = (T)(acc1 + neg_if<T>((T)((T)win + (t ? v1 : (T)0)), party));
        if (sl.nw == 3)
        {
            const uint32_t win2 = qa.encrypt(c, valB, wpt2);
            if (sl.r2 >= 0) acc2 = (T)(acc2 + neg_if<T>((T)((T)win2 + (t ? v2 : (T)0)), party));
        }
        s = ((q == 0) ? (seedB & ~1u) : seedB) ^ (cw & mk);
        t = tn;
    }
    // the final seed converted like a level's value block, with lastCW
    const T *last = tail + (size_t)n * g + 2 + j0;
    const T  l1   = sl.r1 >= 0 ? last[sl.r1] : (T)0;
    const T  l2   = sl.r2 >= 0 ? last[sl.r2] : (T)0;
    const uint32_t win = qa.encrypt(c, s, wpt0);
    if (sl.r1 >= 0) acc1 = (T)(acc1 + neg_if<T>((T)((T)win + (t ? l1 : (T)0)), party));
    if (sl.nw == 3)
    {
        const uint32_t win2 = qa.encrypt(c, s, wpt2);
        if (sl.r2 >= 0) acc2 = (T)(acc2 + neg_if<T>((T)((T)win2 + (t ? l2 : (T)0)), party));
    }
}

// x^k (k < 8) mod 2^bits
template <typename T>
__device__ __forceinline__ T pow_small(T x, int k)
{
    T r = 1;
#pragma unroll
    for (int j = 0; j < 8; j++)
        if (j < k) r = (T)(r * x);
    return r;
}

// The spline of degree <= 7 with 4-byte slots and AES conversion (4g > 16)
// for small calls: work item = (element, block of intervals) as in the split
// kernel, on eight lanes; the element's 8 * bpe lanes (at most 64: one wave)
// add their partial sums with cross-lane shuffles, and lane 4 of the
// element's first group writes the share.  Same results as
// spline_eval_split_kernel.
template <typename T>
__global__ __launch_bounds__(kBlock) void spline_eval_oct_kernel(T *out, const T *x, const uint8_t *key, const T *sE,
                                                                 const T *sBeta, const uint8_t *seed, int party,
                                                                 const T *left, const T *right, size_t I, int d,
                                                                 int n, int m, size_t N, uint32_t bpe)
{
    FSS_KERNEL_PROLOGUE_STAGED
    const uint32_t q = threadIdx.x & 3u, h = (threadIdx.x >> 2) & 1u;
    QuadAes        qa;
    qa.init(c, q);
    const int    C     = d + 1;
    const size_t g     = I * (size_t)C;
    const T     *tails = (const T *)(key + N * 16 * (size_t)n);
    const size_t span  = (I + bpe - 1) / bpe;
    const size_t lanes = N * bpe * 8;
    for (size_t lg = (size_t)blockIdx.x * blockDim.x + threadIdx.x; lg < lanes; lg += (size_t)gridDim.x * blockDim.x)
    {
        const size_t item = lg >> 3;
        const size_t e    = item / bpe;
        const size_t q0   = (item % bpe) * span;
        const size_t q1   = (q0 + span < I) ? q0 + span : I;
        const uint32_t seedCol = *(const uint32_t *)(seed + 16 * e + 4 * q);
        const uint8_t *sc      = key + e * 16 * (size_t)n;
        const T       *tail    = tails + e * (n * g + 2 + g);
        const T        xv      = mod_bits<T>(x[e], n);
        const T       *be = sBeta + e * g, *ee = sE + e * g;
        T              r  = 0;
        for (size_t qi = q0; qi < q1; qi++)
        {
            const T L  = mod_bits<T>(left[qi], n), R = mod_bits<T>(right[qi], n);
            const T Rp = mod_bits<T>((T)(R + 1), n);
            OctSlots sl;
            T        a1, a2;
            // interval qi's start slots come from the linked end walk of qi - 1
            if (!(qi > 0 && mod_bits<T>((T)(right[qi - 1] + 1), n) == L))
            {
                sl.setup((int)(qi * C), C, h, q);
                spline_walk_oct<T>(c, qa, sl, a1, a2, sc, tail, (T)(xv - 1 - L), seedCol, party, n, (int)g,
                                   (int)(qi * C), h, q);
                if (sl.r1 >= 0) r = (T)(r - (T)(a1 * pow_small<T>(xv, sl.r1)));
                if (sl.r2 >= 0) r = (T)(r - (T)(a2 * pow_small<T>(xv, sl.r2)));
            }
            const bool link = (qi + 1 < I) && mod_bits<T>((T)(right[qi] + 1), n) == mod_bits<T>(left[qi + 1], n);
            const int  cnt  = link ? 2 * C : C;
            sl.setup((int)(qi * C), cnt, h, q);
            spline_walk_oct<T>(c, qa, sl, a1, a2, sc, tail, (T)(xv - 1 - Rp), seedCol, party, n, (int)g,
                               (int)(qi * C), h, q);
            // slots below C: interval qi's values at its end (+); from C:
            // interval qi+1's at its start (-)
            if (sl.r1 >= 0)
                r = (sl.r1 < C) ? (T)(r + (T)(a1 * pow_small<T>(xv, sl.r1)))
                                : (T)(r - (T)(a1 * pow_small<T>(xv, sl.r1 - C)));
            if (sl.r2 >= 0)
                r = (sl.r2 < C) ? (T)(r + (T)(a2 * pow_small<T>(xv, sl.r2)))
                                : (T)(r - (T)(a2 * pow_small<T>(xv, sl.r2 - C)));
            if ((threadIdx.x & 7u) == 4u)
            {
                const T cx = (T)((int)(xv > L) - (int)(xv > Rp));
                T       pw = 1;
                for (int j = 0; j < C; j++)
                {
                    r  = (T)(r + (T)((T)((T)(cx * be[qi * C + j]) + ee[qi * C + j]) * pw));
                    pw = (T)(pw * xv);
                }
            }
        }
        // the element's 8 * bpe lanes are consecutive within one wave and all active
        for (uint32_t o = 1; o < 8 * bpe; o <<= 1)
        {
            const uint64_t rv = (uint64_t)r;
            const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)rv, (int)o);
            const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(rv >> 32), (int)o);
            r                 = (T)(r + (T)((uint64_t)lo | ((uint64_t)hi << 32)));
        }
        if ((lg % (8 * bpe)) == 4) out[e] = mod_bits<T>(r, m);
    }
}

} // namespace fss
