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

} // namespace fss
