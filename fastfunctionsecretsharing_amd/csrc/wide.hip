// Elements of 16, 24 and 32 bytes (the reference's wideint::uint<2>, <3>,
// <4>, dispatched by FAST_FSS_DISPATCH_INTEGRAL_TYPES, src/impl/def.h:37-44)
// for every primitive: DCF, DPF (eval, evalAll, evalMulti), Grotto (eval,
// eqEval, MIC, interval LUT), DCF-MIC, DCF-spline and OTTT.
//
// No benchmark configuration uses these widths, so this is the plain form:
// one lane per element on a grid-stride loop, the AES tables in LDS as
// everywhere else, and each element's per-slot vectors (value corrections,
// the running vAlpha, output accumulators, the spline's current/next vectors)
// kept in the caller's buffers rather than registers.  The group elements are
// WideInt<W> (wide_int.h), arithmetic modulo 2^(64 W), so the steps below are
// those of the per-element reference functions with GroupElement = W limbs:
//   DCF     keyGen src/impl/dcf.h:187-368, eval :370-491 (value conversion
//           DcfConvertCtx :69-137: ceil(bitWidthOut/8) bytes per element,
//           assembled little-endian, refilled with AES_v(0, element index)
//           when a group needs more than 16 bytes)
//   DPF     src/impl/dpf.h:172-380 (its conversion takes (bitWidthOut,
//           groupSize) in swapped order, :74,278,283,375), evalAll
//           src/kernel/dpf.h:182-277, evalMulti :383-535
//   Grotto  src/impl/grotto.h:111-545
//   MIC     src/impl/mic.h:9-109
//   spline  src/impl/spline.h:9-253
//   OTTT    src/impl/ottt.h:9-192
// Key layouts are the reference's (src/impl/*.h); the entry points validate
// arguments before they come here, exactly as for the narrow widths.
#include "dcf_device.h"
#include "fss_host.h"
#include "fss_launch.h"
#include "wide_int.h"

namespace fss {

namespace {

// PRG counter blocks {2k, 2k+1}, k = first .. first+NB-1, under key s
template <int NB>
__device__ __forceinline__ void w_prg(const AesCtx &c, const uint32_t s[4], int first, uint32_t out[NB][4])
{
#pragma unroll
    for (int k = 0; k < NB; k++)
        out[k][0] = 2u * (uint32_t)(first + k), out[k][1] = 0u, out[k][2] = 2u * (uint32_t)(first + k) + 1u,
        out[k][3] = 0u;
    aes_encrypt<NB>(c, s, out);
}

// bit 0 of a child block as its control bit; the block becomes the seed
__device__ __forceinline__ int w_take_t(uint32_t b[4])
{
    const int t = (int)(b[0] & 1u);
    b[0] &= ~1u;
    return t;
}

__device__ __forceinline__ void w_xor(uint32_t a[4], const uint32_t b[4])
{
    a[0] ^= b[0], a[1] ^= b[1], a[2] ^= b[2], a[3] ^= b[3];
}

__device__ __forceinline__ void w_copy(uint32_t a[4], const uint32_t b[4])
{
    a[0] = b[0], a[1] = b[1], a[2] = b[2], a[3] = b[3];
}

// the reference's DcfConvertCtx/DpfConvertCtx on a 16-byte seed: `count`
// elements of ceil(bitWidth/8) bytes each
struct WConv
{
    uint32_t key[4], buf[4];
    int      bytes, pos;
    bool     direct;

    __device__ void init(const uint32_t s[4], size_t count, size_t bitWidth)
    {
        bytes  = (int)((bitWidth + 7) / 8);
        direct = (size_t)bytes * count <= 16;
        pos    = 0;
        w_copy(key, s);
        w_copy(buf, s);
    }

    template <int W>
    __device__ WideInt<W> next(const AesCtx &c)
    {
        WideInt<W>   r(0);
        const size_t off = (size_t)pos * (size_t)bytes;
        for (int i = 0; i < bytes; i++)
        {
            const int idx = (int)((off + (size_t)i) % 16);
            if (!direct && idx == 0)
            {
                // AES_key(u64 0 || u64 element index), little-endian
                uint32_t b[1][4] = {{0u, 0u, (uint32_t)pos, 0u}};
                aes_encrypt<1>(c, key, b);
                w_copy(buf, b[0]);
            }
            if (i < 8 * W) r = r | (WideInt<W>(byte_of(buf, idx)) << (8 * i));
        }
        pos++;
        return r;
    }
};

template <int W>
__device__ __forceinline__ int w_bit(const WideInt<W> &x, int i)
{
    return (int)(uint32_t)(x >> i) & 1;
}

template <int W>
__device__ __forceinline__ WideInt<W> w_sign(int neg)
{
    return neg ? WideInt<W>(-1) : WideInt<W>(1);
}

// ---------------------------------------------------------------- DCF ----

// keyGen of one element; vCW rows, tLCW, tRCW, lastCW in `tail`, lastCW
// doubling as the running vAlpha (src/impl/dcf.h:187-368)
template <int W>
__device__ void w_dcf_keygen_one(const AesCtx &c, uint8_t *sCW, WideInt<W> *tail, WideInt<W> alpha,
                                 const WideInt<W> *beta, const uint8_t *seed0, const uint8_t *seed1, int n, int m,
                                 size_t g)
{
    using GE         = WideInt<W>;
    GE *const vcw    = tail;
    GE *const last   = tail + (size_t)n * g + 2;
    for (size_t j = 0; j < g; j++) last[j] = GE(0);
    uint32_t s0[4], s1[4];
    load16(s0, seed0);
    load16(s1, seed1);
    s0[0] &= ~1u, s1[0] &= ~1u;
    int t0 = 0, t1 = 1;
    GE  tl(0), tr(0);
    for (int i = 0; i < n; i++)
    {
        uint32_t c0[4][4], c1[4][4]; // sL vL sR vR
        w_prg<4>(c, s0, 0, c0);
        w_prg<4>(c, s1, 0, c1);
        const int tL0 = w_take_t(c0[0]), tR0 = w_take_t(c0[2]);
        const int tL1 = w_take_t(c1[0]), tR1 = w_take_t(c1[2]);
        const int keep = w_bit(alpha, n - 1 - i), lose = 1 - keep;
        const uint32_t *sLose0 = lose ? c0[2] : c0[0], *sLose1 = lose ? c1[2] : c1[0];
        const uint32_t *vLose0 = lose ? c0[3] : c0[1], *vLose1 = lose ? c1[3] : c1[1];
        const uint32_t *sKeep0 = keep ? c0[2] : c0[0], *sKeep1 = keep ? c1[2] : c1[0];
        const uint32_t *vKeep0 = keep ? c0[3] : c0[1], *vKeep1 = keep ? c1[3] : c1[1];
        uint32_t        cw[4] = {sLose0[0] ^ sLose1[0], sLose0[1] ^ sLose1[1], sLose0[2] ^ sLose1[2],
                                 sLose0[3] ^ sLose1[3]};
        const GE        sign1 = w_sign<W>(t1);
        WConv           a1, a0, k1, k0;
        a1.init(vLose1, g, m), a0.init(vLose0, g, m), k1.init(vKeep1, g, m), k0.init(vKeep0, g, m);
        for (size_t j = 0; j < g; j++)
        {
            const GE x1 = a1.next<W>(c), x0 = a0.next<W>(c), y1 = k1.next<W>(c), y0 = k0.next<W>(c);
            const GE vA = last[j];
            GE       v  = (x1 - x0 - vA) * sign1;
            if (lose == 0) v = v + sign1 * (beta ? beta[j] : GE(j == 0 ? 1 : 0));
            vcw[(size_t)i * g + j] = v;
            last[j]                = vA - y1 + y0 + sign1 * v;
        }
        const int tLCW = tL0 ^ tL1 ^ keep ^ 1, tRCW = tR0 ^ tR1 ^ keep;
        store16(sCW + 16 * (size_t)i, cw);
        tl = tl | (GE(tLCW) << i);
        tr = tr | (GE(tRCW) << i);
        uint32_t k0s[4], k1s[4];
        w_copy(k0s, sKeep0), w_copy(k1s, sKeep1);
        if (t0) w_xor(k0s, cw);
        if (t1) w_xor(k1s, cw);
        w_copy(s0, k0s), w_copy(s1, k1s);
        const int tKeepCW = keep ? tRCW : tLCW;
        const int tk0 = keep ? tR0 : tL0, tk1 = keep ? tR1 : tL1;
        t0 = tk0 ^ (t0 & tKeepCW);
        t1 = tk1 ^ (t1 & tKeepCW);
    }
    tail[(size_t)n * g]     = tl;
    tail[(size_t)n * g + 1] = tr;
    const GE sign1          = w_sign<W>(t1);
    WConv    b1, b0;
    b1.init(s1, g, m), b0.init(s0, g, m);
    for (size_t j = 0; j < g; j++)
    {
        const GE x1 = b1.next<W>(c), x0 = b0.next<W>(c);
        last[j]     = (GE(0) - last[j] + x1 - x0) * sign1;
    }
}

// eval of one element into out[0 .. g) (src/impl/dcf.h:370-491)
template <int W>
__device__ void w_dcf_eval_one(const AesCtx &c, WideInt<W> *out, const uint8_t *sCW, const WideInt<W> *tail,
                               WideInt<W> x, const uint8_t *seed, int party, int n, int m, size_t g)
{
    using GE      = WideInt<W>;
    const GE sign = w_sign<W>(party);
    uint32_t s[4];
    load16(s, seed);
    s[0] &= ~1u;
    int t = party;
    x     = wide_mod_bits(x, n);
    for (size_t j = 0; j < g; j++) out[j] = GE(0);
    const GE tl = tail[(size_t)n * g], tr = tail[(size_t)n * g + 1];
    for (int i = 0; i < n; i++)
    {
        uint32_t cb[4][4];
        w_prg<4>(c, s, 0, cb);
        int tL = w_take_t(cb[0]), tR = w_take_t(cb[2]);
        if (t)
        {
            uint32_t cw[4];
            load16(cw, sCW + 16 * (size_t)i);
            w_xor(cb[0], cw), w_xor(cb[2], cw);
            tL ^= w_bit(tl, i), tR ^= w_bit(tr, i);
        }
        const int xi = w_bit(x, n - 1 - i);
        WConv     cv;
        cv.init(xi ? cb[3] : cb[1], g, m);
        for (size_t j = 0; j < g; j++)
        {
            const GE term = cv.next<W>(c) + (t ? tail[(size_t)i * g + j] : GE(0));
            out[j]        = out[j] + sign * term;
        }
        w_copy(s, xi ? cb[2] : cb[0]);
        t = xi ? tR : tL;
    }
    WConv cv;
    cv.init(s, g, m);
    for (size_t j = 0; j < g; j++)
    {
        const GE term = cv.next<W>(c) + (t ? tail[(size_t)n * g + 2 + j] : GE(0));
        out[j]        = out[j] + sign * term;
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void w_dcf_keygen_kernel(uint8_t *key, const WideInt<W> *alpha,
                                                              const WideInt<W> *beta, const uint8_t *seed0,
                                                              const uint8_t *seed1, int n, int m, size_t g, size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    const size_t es = sizeof(WideInt<W>);
    FSS_GRID_LOOP(e, N)
    {
        WideInt<W> *tail = (WideInt<W> *)(key + N * 16 * (size_t)n + e * es * ((size_t)n * g + 2 + g));
        w_dcf_keygen_one<W>(c, key + e * 16 * (size_t)n, tail, alpha[e], beta ? beta + e * g : nullptr,
                            seed0 + 16 * e, seed1 + 16 * e, n, m, g);
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void w_dcf_eval_kernel(WideInt<W> *out, const WideInt<W> *x, const uint8_t *key,
                                                            const uint8_t *seed, int party, int n, int m, size_t g,
                                                            size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    const size_t es = sizeof(WideInt<W>);
    FSS_GRID_LOOP(e, N)
    {
        const WideInt<W> *tail =
            (const WideInt<W> *)(key + N * 16 * (size_t)n + e * es * ((size_t)n * g + 2 + g));
        w_dcf_eval_one<W>(c, out + e * g, key + e * 16 * (size_t)n, tail, x[e], seed + 16 * e, party, n, m, g);
    }
}

// ---------------------------------------------------------------- DPF ----

// one level of a two-block (DPF/Grotto) tree for both parties; writes the
// level's seed correction, returns (tLCW, tRCW) (src/impl/dpf.h:201-272)
__device__ void w_two_block_level(const AesCtx &c, uint32_t s0[4], uint32_t s1[4], int &t0, int &t1, int keep,
                                  uint32_t cw[4], int &tLCW, int &tRCW)
{
    uint32_t c0[2][4], c1[2][4];
    w_prg<2>(c, s0, 0, c0);
    w_prg<2>(c, s1, 0, c1);
    const int tL0 = w_take_t(c0[0]), tR0 = w_take_t(c0[1]);
    const int tL1 = w_take_t(c1[0]), tR1 = w_take_t(c1[1]);
    const int lose = 1 - keep;
    const uint32_t *sLose0 = lose ? c0[1] : c0[0], *sLose1 = lose ? c1[1] : c1[0];
    uint32_t        k0s[4], k1s[4];
    w_copy(k0s, keep ? c0[1] : c0[0]);
    w_copy(k1s, keep ? c1[1] : c1[0]);
    for (int q = 0; q < 4; q++) cw[q] = sLose0[q] ^ sLose1[q];
    tLCW = tL0 ^ tL1 ^ keep ^ 1;
    tRCW = tR0 ^ tR1 ^ keep;
    if (t0) w_xor(k0s, cw);
    if (t1) w_xor(k1s, cw);
    const int tKeepCW = keep ? tRCW : tLCW;
    const int tk0 = keep ? tR0 : tL0, tk1 = keep ? tR1 : tL1;
    t0 = tk0 ^ (t0 & tKeepCW);
    t1 = tk1 ^ (t1 & tKeepCW);
    w_copy(s0, k0s), w_copy(s1, k1s);
}

template <int W>
__global__ __launch_bounds__(kBlock) void w_dpf_keygen_kernel(uint8_t *key, const WideInt<W> *alpha,
                                                              const WideInt<W> *beta, const uint8_t *seed0,
                                                              const uint8_t *seed1, int n, int m, size_t g, size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    using GE        = WideInt<W>;
    const size_t es = sizeof(GE);
    FSS_GRID_LOOP(e, N)
    {
        uint8_t *sCW  = key + e * 16 * (size_t)n;
        GE      *tail = (GE *)(key + N * 16 * (size_t)n + e * es * (2 + g));
        uint32_t s0[4], s1[4];
        load16(s0, seed0 + 16 * e);
        load16(s1, seed1 + 16 * e);
        s0[0] &= ~1u, s1[0] &= ~1u;
        int      t0 = 0, t1 = 1;
        GE       tl(0), tr(0);
        const GE a = alpha[e];
        for (int i = 0; i < n; i++)
        {
            uint32_t cw[4];
            int      tLCW, tRCW;
            w_two_block_level(c, s0, s1, t0, t1, w_bit(a, n - 1 - i), cw, tLCW, tRCW);
            store16(sCW + 16 * (size_t)i, cw);
            tl = tl | (GE(tLCW) << i);
            tr = tr | (GE(tRCW) << i);
        }
        tail[0] = tl;
        tail[1] = tr;
        // the conversion's (groupSize, bitWidth) receive (bitWidthOut, groupSize), as in the reference
        const GE sign1 = w_sign<W>(t1);
        WConv    c0, c1;
        c0.init(s0, (size_t)m, g), c1.init(s1, (size_t)m, g);
        for (size_t j = 0; j < g; j++)
        {
            const GE x0 = c0.next<W>(c), x1 = c1.next<W>(c);
            const GE b  = beta ? beta[e * g + j] : GE(1);
            tail[2 + j] = sign1 * (b - x0 + x1);
        }
    }
}

// the leaf (seed, t) of one DPF walk at point x
template <int W>
__device__ void w_dpf_walk(const AesCtx &c, uint32_t s[4], int &t, const uint8_t *sCW, const WideInt<W> *tail,
                           WideInt<W> x, const uint8_t *seed, int party, int n)
{
    load16(s, seed);
    s[0] &= ~1u;
    t                     = party;
    x                     = wide_mod_bits(x, n);
    const WideInt<W> tl = tail[0], tr = tail[1];
    for (int i = 0; i < n; i++)
    {
        uint32_t cb[2][4];
        w_prg<2>(c, s, 0, cb);
        int tL = w_take_t(cb[0]), tR = w_take_t(cb[1]);
        if (t)
        {
            uint32_t cw[4];
            load16(cw, sCW + 16 * (size_t)i);
            w_xor(cb[0], cw), w_xor(cb[1], cw);
            tL ^= w_bit(tl, i), tR ^= w_bit(tr, i);
        }
        const int xi = w_bit(x, n - 1 - i);
        w_copy(s, xi ? cb[1] : cb[0]);
        t = xi ? tR : tL;
    }
}

template <int W>
__device__ void w_dpf_eval_one(const AesCtx &c, WideInt<W> *out, const uint8_t *sCW, const WideInt<W> *tail,
                               WideInt<W> x, const uint8_t *seed, int party, int n, int m, size_t g)
{
    uint32_t s[4];
    int      t;
    w_dpf_walk<W>(c, s, t, sCW, tail, x, seed, party, n);
    const WideInt<W> sign = w_sign<W>(party);
    WConv            cv;
    cv.init(s, (size_t)m, g); // the same argument swap as keyGen
    for (size_t j = 0; j < g; j++) out[j] = sign * (cv.next<W>(c) + (t ? tail[2 + j] : WideInt<W>(0)));
}

// mode 0: eval at x; 1: evalAll (out[(x - j) mod 2^n] = Eval(j)); 2: evalMulti at x - point[p]
template <int W>
__global__ __launch_bounds__(kBlock) void w_dpf_eval_kernel(WideInt<W> *out, const WideInt<W> *x, const uint8_t *key,
                                                            const uint8_t *seed, int party, int n, int m, size_t g,
                                                            size_t N, int mode, const WideInt<W> *point, size_t P)
{
    FSS_KERNEL_PROLOGUE_STAGED
    using GE        = WideInt<W>;
    const size_t es = sizeof(GE);
    FSS_GRID_LOOP(e, N)
    {
        const uint8_t *sCW  = key + e * 16 * (size_t)n;
        const GE      *tail = (const GE *)(key + N * 16 * (size_t)n + e * es * (2 + g));
        if (mode == 0)
            w_dpf_eval_one<W>(c, out + e * g, sCW, tail, x[e], seed + 16 * e, party, n, m, g);
        else if (mode == 1)
        {
            const size_t size = (size_t)1 << n;
            for (size_t j = 0; j < size; j++)
            {
                const size_t idx = (size_t)(uint64_t)(x[e] - GE(j)) & (size - 1);
                w_dpf_eval_one<W>(c, out + (e * size + idx) * g, sCW, tail, GE(j), seed + 16 * e, party, n, m, g);
            }
        }
        else
            for (size_t p = 0; p < P; p++)
                w_dpf_eval_one<W>(c, out + (e * P + p) * g, sCW, tail, x[e] - point[p], seed + 16 * e, party, n, m,
                                  g);
    }
}

// ------------------------------------------------------------- Grotto ----

template <int W>
struct WGrottoKey
{
    const uint8_t    *sCW;
    const uint8_t    *lastCW;
    const WideInt<W> *tlr;
};

template <int W>
__device__ __forceinline__ WGrottoKey<W> w_grotto_key(const uint8_t *key, int n, size_t e, size_t N)
{
    const size_t lv = (size_t)n - 6;
    return {key + e * 16 * lv, key + N * 16 * lv + 8 * e,
            (const WideInt<W> *)(key + N * 16 * lv + N * 8 + e * 2 * sizeof(WideInt<W>))};
}

template <int W>
__global__ __launch_bounds__(kBlock) void w_grotto_keygen_kernel(uint8_t *key, const WideInt<W> *alpha,
                                                                 const uint8_t *seed0, const uint8_t *seed1, int n,
                                                                 size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    using GE        = WideInt<W>;
    const size_t lv = (size_t)n - 6;
    FSS_GRID_LOOP(e, N)
    {
        uint8_t *sCW = key + e * 16 * lv;
        GE      *tlr = (GE *)(key + N * 16 * lv + N * 8 + e * 2 * sizeof(GE));
        uint32_t s0[4], s1[4];
        load16(s0, seed0 + 16 * e);
        load16(s1, seed1 + 16 * e);
        s0[0] &= ~1u, s1[0] &= ~1u;
        int      t0 = 0, t1 = 1;
        GE       tl(0), tr(0);
        const GE a = alpha[e];
        for (int i = 0; i < (int)lv; i++)
        {
            uint32_t cw[4];
            int      tLCW, tRCW;
            w_two_block_level(c, s0, s1, t0, t1, w_bit(a, n - 1 - i), cw, tLCW, tRCW);
            store16(sCW + 16 * (size_t)i, cw);
            tl = tl | (GE(tLCW) << i);
            tr = tr | (GE(tRCW) << i);
        }
        tlr[0] = tl;
        tlr[1] = tr;
        const uint64_t hi0  = (uint64_t)s0[2] | ((uint64_t)s0[3] << 32);
        const uint64_t hi1  = (uint64_t)s1[2] | ((uint64_t)s1[3] << 32);
        const uint64_t leaf = hi0 ^ hi1 ^ (1ull << ((uint32_t)a & 63u));
        *(uint64_t *)(key + N * 16 * lv + 8 * e) = leaf;
    }
}

// one path level: a single block chosen by the input bit (src/impl/grotto.h:336-362)
template <int W>
__device__ __forceinline__ void w_grotto_step(const AesCtx &c, const WGrottoKey<W> &k, int i, int bit, uint32_t s[4],
                                              int &t, const WideInt<W> &tl, const WideInt<W> &tr)
{
    uint32_t cb[1][4];
    w_prg<1>(c, s, bit, cb);
    int ti = w_take_t(cb[0]);
    if (t)
    {
        uint32_t cw[4];
        load16(cw, k.sCW + 16 * (size_t)i);
        w_xor(cb[0], cw);
        ti ^= w_bit(bit ? tr : tl, i);
    }
    w_copy(s, cb[0]);
    t = ti;
}

__device__ __forceinline__ int w_parity64(uint64_t x)
{
    return __popcll(x) & 1;
}

// mode 0: lt ([x < alpha]), 1: le (equalBound), 2: eq (src/impl/grotto.h:207-391)
template <int W>
__device__ int w_grotto_point(const AesCtx &c, const WGrottoKey<W> &k, WideInt<W> x, const uint8_t *seed, int party,
                              int n, int mode)
{
    using GE = WideInt<W>;
    x        = wide_mod_bits(x, n);
    uint32_t s[4];
    load16(s, seed);
    s[0] &= ~1u;
    int      t = party, parity = 0;
    const GE tl = k.tlr[0], tr = k.tlr[1];
    for (int i = 0; i < n - 6; i++)
    {
        // le: early exit when the remaining low n - i bits are all zero
        if (mode == 1 && !(bool)(x & wide_mod_bits(~GE(0), n - i))) return parity ^ party;
        const int bit = w_bit(x, n - 1 - i), prev = t;
        w_grotto_step<W>(c, k, i, bit, s, t, tl, tr);
        if (bit) parity ^= prev ^ t;
    }
    const uint64_t lcw = *(const uint64_t *)k.lastCW;
    const uint64_t hi  = (uint64_t)s[2] | ((uint64_t)s[3] << 32);
    const uint64_t u   = t ? (hi ^ lcw) : hi;
    const unsigned p   = (uint32_t)x & 63u;
    if (mode == 2) return (int)((u >> p) & 1);
    if (mode == 1)
    {
        if (p != 0) parity ^= w_parity64(u & (~0ull >> (64 - p)));
    }
    else
        parity ^= w_parity64(u & (~0ull >> (63 - p)));
    return parity ^ party;
}

// interval shares of one element with the reuse of the previous right end
// (src/impl/grotto.h:393-447); share q in sh[q] as 0/1
template <int W>
__device__ void w_grotto_mic_one(const AesCtx &c, int *sh, const WGrottoKey<W> &k, WideInt<W> x, const uint8_t *seed,
                                 int party, const WideInt<W> *left, const WideInt<W> *right, size_t I, int n)
{
    using GE = WideInt<W>;
    int sp = 0, sq = 0;
    for (size_t i = 0; i < I; i++)
    {
        const GE L = left[i], R = right[i];
        const GE xP = wide_mod_bits(x - (L - GE(1)), n);
        const GE xQ = wide_mod_bits(x - R, n);
        if (i > 0 && L == wide_mod_bits(right[i - 1] + GE(1), n))
            sp = sq;
        else
            sp = w_grotto_point<W>(c, k, xP, seed, party, n, 1);
        sq    = w_grotto_point<W>(c, k, xQ, seed, party, n, 1);
        sh[i] = (sp ^ sq) ^ ((xQ >= xP) ? party : 0);
    }
}

// op 0: eval/eqEval (mode as w_grotto_point); 1: MIC; 2: interval LUT
template <int W>
__global__ __launch_bounds__(kBlock) void w_grotto_eval_kernel(WideInt<W> *out, WideInt<W> *outT, int *bits,
                                                               const WideInt<W> *x, const uint8_t *key,
                                                               const uint8_t *seed, int party, int n, size_t N,
                                                               int op, int mode, const WideInt<W> *left,
                                                               const WideInt<W> *right, size_t I,
                                                               const WideInt<W> *lut, size_t lutNum)
{
    FSS_KERNEL_PROLOGUE_STAGED
    using GE = WideInt<W>;
    FSS_GRID_LOOP(e, N)
    {
        const WGrottoKey<W> k = w_grotto_key<W>(key, n, e, N);
        if (op == 0)
        {
            out[e] = GE(w_grotto_point<W>(c, k, x[e], seed + 16 * e, party, n, mode));
            continue;
        }
        int *sh = bits + e * I;
        w_grotto_mic_one<W>(c, sh, k, x[e], seed + 16 * e, party, left, right, I, n);
        if (op == 1)
        {
            for (size_t q = 0; q < I; q++) out[e * I + q] = GE(sh[q]);
            continue;
        }
        GE E(0);
        for (size_t q = 0; q < I; q++) E = E + GE(sh[q] & 1);
        for (size_t j = 0; j < lutNum; j++)
        {
            GE acc(0);
            for (size_t q = 0; q < I; q++)
                if (sh[q] & 1) acc = acc + lut[q + j * I];
            outT[e * lutNum + j] = party ? -acc : acc;
        }
        out[e] = party ? -E : E;
    }
}

// --------------------------------------------------------------- MIC ----

template <int W>
__global__ __launch_bounds__(kBlock) void w_mic_keygen_kernel(uint8_t *key, WideInt<W> *z, const WideInt<W> *alpha,
                                                              const uint8_t *seed0, const uint8_t *seed1,
                                                              const WideInt<W> *left, const WideInt<W> *right,
                                                              size_t I, int n, int m, size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    using GE        = WideInt<W>;
    const size_t es = sizeof(GE);
    const GE     MAX = wide_mod_bits(~GE(0), n);
    FSS_GRID_LOOP(e, N)
    {
        const GE a    = alpha[e];
        GE      *tail = (GE *)(key + N * 16 * (size_t)n + e * es * ((size_t)n + 3));
        const GE one(1);
        // DCF on alpha - 1 with beta = 1 (src/impl/mic.h:22)
        w_dcf_keygen_one<W>(c, key + e * 16 * (size_t)n, tail, a - GE(1), &one, seed0 + 16 * e, seed1 + 16 * e, n,
                            m, 1);
        for (size_t q = 0; q < I; q++)
        {
            const GE L = left[q], R = right[q];
            const GE qP  = wide_mod_bits(R + GE(1), n);
            const GE aP  = wide_mod_bits(L + a, n);
            const GE aQ  = wide_mod_bits(R + a, n);
            const GE aQp = wide_mod_bits(R + GE(1) + a, n);
            const GE p   = wide_mod_bits(L, n);
            const int zv = (int)(aP > aQ) - (int)(aP > p) + (int)(aQp > qP) + (int)(aQ == MAX);
            z[e * I + q] = GE(zv);
        }
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void w_mic_eval_kernel(WideInt<W> *out, const WideInt<W> *x, const uint8_t *key,
                                                            const WideInt<W> *z, const uint8_t *seed, int party,
                                                            const WideInt<W> *left, const WideInt<W> *right, size_t I,
                                                            int n, int m, size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    using GE        = WideInt<W>;
    const size_t es = sizeof(GE);
    FSS_GRID_LOOP(e, N)
    {
        const uint8_t *sCW  = key + e * 16 * (size_t)n;
        const GE      *tail = (const GE *)(key + N * 16 * (size_t)n + e * es * ((size_t)n + 3));
        const GE       xv   = wide_mod_bits(x[e], n);
        GE             sp(0), sq(0);
        for (size_t i = 0; i < I; i++)
        {
            const GE L = left[i], R = right[i];
            const GE qP  = wide_mod_bits(R + GE(1), n);
            const GE xP  = wide_mod_bits(xv - GE(1) - L, n);
            const GE xQp = wide_mod_bits(xv - GE(1) - (R + GE(1)), n);
            const GE p   = wide_mod_bits(L, n);
            if (i > 0 && L == wide_mod_bits(right[i - 1] + GE(1), n))
                sp = sq;
            else
                w_dcf_eval_one<W>(c, &sp, sCW, tail, xP, seed + 16 * e, party, n, m, 1);
            w_dcf_eval_one<W>(c, &sq, sCW, tail, xQp, seed + 16 * e, party, n, m, 1);
            GE o = sq - sp + z[e * I + i];
            if (party == 1) o = o + GE((int)(xv > p) - (int)(xv > qP));
            out[e * I + i] = o;
        }
    }
}

// ------------------------------------------------------------ spline ----

template <int W>
__global__ __launch_bounds__(kBlock) void w_spline_keygen_kernel(uint8_t *key, WideInt<W> *eOut, WideInt<W> *betaOut,
                                                                 const WideInt<W> *alpha, const uint8_t *seed0,
                                                                 const uint8_t *seed1, const WideInt<W> *coef, int d,
                                                                 const WideInt<W> *left, const WideInt<W> *right,
                                                                 size_t I, int n, int m, size_t N)
{
    FSS_KERNEL_PROLOGUE_STAGED
    using GE        = WideInt<W>;
    const size_t es = sizeof(GE);
    const size_t C = (size_t)d + 1, g = I * C;
    const GE     maxv = wide_mod_bits(~GE(0), n);
    FSS_GRID_LOOP(e, N)
    {
        const GE a    = alpha[e];
        GE      *beta = betaOut + e * g;
        // coefficients of f(x - alpha) mod 2^m (src/impl/spline.h:89-117)
        const GE neg = wide_mod_bits(GE(0) - a, m);
        for (size_t q = 0; q < I; q++)
        {
            GE *o = beta + q * C;
            for (size_t i = 0; i < C; i++) o[i] = GE(0);
            for (size_t i = C; i > 0; --i)
            {
                for (size_t j = C - 1; j > 0; --j) o[j] = o[j - 1] + o[j] * neg;
                o[0] = o[0] * neg + coef[q * C + i - 1];
            }
            for (size_t i = 0; i < C; i++) o[i] = wide_mod_bits(o[i], m);
        }
        GE *tail = (GE *)(key + N * 16 * (size_t)n + e * es * ((size_t)n * g + 2 + g));
        w_dcf_keygen_one<W>(c, key + e * 16 * (size_t)n, tail, a - GE(1), beta, seed0 + 16 * e, seed1 + 16 * e, n,
                            m, g);
        for (size_t q = 0; q < I; q++)
        {
            const GE L = wide_mod_bits(left[q], n), R = wide_mod_bits(right[q], n);
            const GE Rp = wide_mod_bits(R + GE(1), n);
            const GE aL = wide_mod_bits(L + a, n), aR = wide_mod_bits(R + a, n), aRp = wide_mod_bits(Rp + a, n);
            const GE cr = GE((int)(aL > aR) - (int)(aL > L) + (int)(aRp > Rp) + (int)(aR == maxv));
            for (size_t j = 0; j < C; j++) eOut[e * g + q * C + j] = wide_mod_bits(cr * beta[q * C + j], m);
        }
    }
}

// the reference's two group vectors and the coefficient sums in the
// caller's cache (src/impl/spline.h:186-253)
template <int W>
__global__ __launch_bounds__(kBlock) void w_spline_eval_kernel(WideInt<W> *out, const WideInt<W> *x,
                                                               const uint8_t *key, const WideInt<W> *sE,
                                                               const WideInt<W> *sBeta, const uint8_t *seed, int party,
                                                               const WideInt<W> *left, const WideInt<W> *right,
                                                               size_t I, int d, int n, int m, size_t N, uint8_t *cache)
{
    FSS_KERNEL_PROLOGUE_STAGED
    using GE        = WideInt<W>;
    const size_t es = sizeof(GE);
    const size_t C = (size_t)d + 1, g = I * C;
    uint8_t     *ws = cache + dcf_cache_size((size_t)n, g, es, N);
    FSS_GRID_LOOP(e, N)
    {
        const uint8_t *sCW  = key + e * 16 * (size_t)n;
        const GE      *tail = (const GE *)(key + N * 16 * (size_t)n + e * es * ((size_t)n * g + 2 + g));
        GE            *cur  = (GE *)(ws + e * es * (2 * g + C));
        GE            *nxt  = cur + g;
        GE            *w    = nxt + g;
        const GE       xv   = wide_mod_bits(x[e], n);
        for (size_t j = 0; j < C; j++) w[j] = GE(0);
        for (size_t q = 0; q < I; q++)
        {
            const GE L = wide_mod_bits(left[q], n), R = wide_mod_bits(right[q], n);
            const GE Rp = wide_mod_bits(R + GE(1), n);
            if (q > 0 && wide_mod_bits(right[q - 1] + GE(1), n) == L)
                for (size_t j = 0; j < g; j++) cur[j] = nxt[j];
            else
                w_dcf_eval_one<W>(c, cur, sCW, tail, xv - GE(1) - L, seed + 16 * e, party, n, m, g);
            w_dcf_eval_one<W>(c, nxt, sCW, tail, xv - GE(1) - Rp, seed + 16 * e, party, n, m, g);
            const GE cx = GE((int)(xv > L) - (int)(xv > Rp));
            for (size_t j = 0; j < C; j++)
                w[j] = w[j] + cx * sBeta[e * g + q * C + j] - cur[q * C + j] + nxt[q * C + j] + sE[e * g + q * C + j];
        }
        GE r(0), pw(1);
        for (size_t j = 0; j < C; j++)
        {
            r  = r + w[j] * pw;
            pw = pw * xv;
        }
        out[e] = wide_mod_bits(r, m);
    }
}

// -------------------------------------------------------------- OTTT ----

template <int W>
__global__ __launch_bounds__(kBlock) void w_ottt_keygen_kernel(uint8_t *key, const WideInt<W> *alpha, int n, size_t N)
{
    const size_t per = ((size_t)1 << n) / 8;
    FSS_GRID_LOOP(e, N)
    {
        const uint64_t a = (uint64_t)wide_mod_bits(alpha[e], n);
        key[e * per + a / 8] ^= (uint8_t)(1u << (a % 8));
    }
}

// key bit b selects LUT[(x - b) mod 2^n]; party 0 negates (src/impl/ottt.h:26-192)
template <int W>
__global__ __launch_bounds__(kBlock) void w_ottt_eval_kernel(WideInt<W> *outE, WideInt<W> *outT, const WideInt<W> *x,
                                                             const uint8_t *key, int party, const WideInt<W> *lut,
                                                             size_t lutNum, int n, size_t N)
{
    using GE           = WideInt<W>;
    const size_t total = (size_t)1 << n, per = total / 8;
    FSS_GRID_LOOP(e, N)
    {
        const uint8_t *kb = key + e * per;
        const uint64_t xv = (uint64_t)x[e];
        GE             E(0);
        for (size_t j = 0; j < lutNum; j++) outT[e * lutNum + j] = GE(0);
        for (size_t b = 0; b < total; b++)
        {
            if (!((kb[b / 8] >> (b % 8)) & 1)) continue;
            E                = E + GE(1);
            const size_t idx = (size_t)(xv - b) & (total - 1);
            for (size_t j = 0; j < lutNum; j++) outT[e * lutNum + j] = outT[e * lutNum + j] + lut[idx + j * total];
        }
        if (party == 0)
        {
            E = -E;
            for (size_t j = 0; j < lutNum; j++) outT[e * lutNum + j] = -outT[e * lutNum + j];
        }
        outE[e] = E;
    }
}

} // namespace

// ------------------------------------------------------- launchers ----

#define FSS_WIDE_DISPATCH(ES, ...)                 \
    switch (ES)                                    \
    {                                              \
        case 16: {                                 \
            constexpr int W = 2;                   \
            __VA_ARGS__;                           \
        }                                          \
        break;                                     \
        case 24: {                                 \
            constexpr int W = 3;                   \
            __VA_ARGS__;                           \
        }                                          \
        break;                                     \
        case 32: {                                 \
            constexpr int W = 4;                   \
            __VA_ARGS__;                           \
        }                                          \
        break;                                     \
        default: return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR; \
    }

#define FSS_WSHAPE(N) ::fss::shape_for((N) ? (N) : 1).grid, ::fss::shape_for((N) ? (N) : 1).block, 0, st

int wide_dcf_keygen(hipStream_t st, void *key, const void *alpha, const void *beta, const void *seed0,
                    const void *seed1, int n, int m, size_t g, size_t es, size_t N)
{
    FSS_WIDE_DISPATCH(es, w_dcf_keygen_kernel<W><<<FSS_WSHAPE(N)>>>(
                              (uint8_t *)key, (const WideInt<W> *)alpha, (const WideInt<W> *)beta,
                              (const uint8_t *)seed0, (const uint8_t *)seed1, n, m, g, N));
    return launch_status();
}

int wide_dcf_eval(hipStream_t st, void *out, const void *x, const void *key, const void *seed, int party, int n, int m,
                  size_t g, size_t es, size_t N)
{
    FSS_WIDE_DISPATCH(es, w_dcf_eval_kernel<W><<<FSS_WSHAPE(N)>>>((WideInt<W> *)out, (const WideInt<W> *)x,
                                                                  (const uint8_t *)key, (const uint8_t *)seed, party,
                                                                  n, m, g, N));
    return launch_status();
}

int wide_dpf_keygen(hipStream_t st, void *key, const void *alpha, const void *beta, const void *seed0,
                    const void *seed1, int n, int m, size_t g, size_t es, size_t N)
{
    FSS_WIDE_DISPATCH(es, w_dpf_keygen_kernel<W><<<FSS_WSHAPE(N)>>>(
                              (uint8_t *)key, (const WideInt<W> *)alpha, (const WideInt<W> *)beta,
                              (const uint8_t *)seed0, (const uint8_t *)seed1, n, m, g, N));
    return launch_status();
}

int wide_dpf_eval(hipStream_t st, int mode, void *out, const void *x, const void *key, const void *seed, int party,
                  int n, int m, size_t g, size_t es, size_t N, const void *point, size_t P)
{
    FSS_WIDE_DISPATCH(es, w_dpf_eval_kernel<W><<<FSS_WSHAPE(N)>>>(
                              (WideInt<W> *)out, (const WideInt<W> *)x, (const uint8_t *)key, (const uint8_t *)seed,
                              party, n, m, g, N, mode, (const WideInt<W> *)point, P));
    return launch_status();
}

int wide_grotto_keygen(hipStream_t st, void *key, const void *alpha, const void *seed0, const void *seed1, int n,
                       size_t es, size_t N)
{
    FSS_WIDE_DISPATCH(es, w_grotto_keygen_kernel<W><<<FSS_WSHAPE(N)>>>((uint8_t *)key, (const WideInt<W> *)alpha,
                                                                       (const uint8_t *)seed0, (const uint8_t *)seed1,
                                                                       n, N));
    return launch_status();
}

int wide_grotto_eval(hipStream_t st, int op, int mode, void *out, void *outT, void *bits, const void *x,
                     const void *key, const void *seed, int party, int n, size_t es, size_t N, const void *left,
                     const void *right, size_t I, const void *lut, size_t lutNum)
{
    FSS_WIDE_DISPATCH(es, w_grotto_eval_kernel<W><<<FSS_WSHAPE(N)>>>(
                              (WideInt<W> *)out, (WideInt<W> *)outT, (int *)bits, (const WideInt<W> *)x,
                              (const uint8_t *)key, (const uint8_t *)seed, party, n, N, op, mode,
                              (const WideInt<W> *)left, (const WideInt<W> *)right, I, (const WideInt<W> *)lut,
                              lutNum));
    return launch_status();
}

int wide_mic_keygen(hipStream_t st, void *key, void *z, const void *alpha, const void *seed0, const void *seed1,
                    const void *left, const void *right, size_t I, int n, int m, size_t es, size_t N)
{
    FSS_WIDE_DISPATCH(es, w_mic_keygen_kernel<W><<<FSS_WSHAPE(N)>>>(
                              (uint8_t *)key, (WideInt<W> *)z, (const WideInt<W> *)alpha, (const uint8_t *)seed0,
                              (const uint8_t *)seed1, (const WideInt<W> *)left, (const WideInt<W> *)right, I, n, m,
                              N));
    return launch_status();
}

int wide_mic_eval(hipStream_t st, void *out, const void *x, const void *key, const void *z, const void *seed,
                  int party, const void *left, const void *right, size_t I, int n, int m, size_t es, size_t N)
{
    FSS_WIDE_DISPATCH(es, w_mic_eval_kernel<W><<<FSS_WSHAPE(N)>>>(
                              (WideInt<W> *)out, (const WideInt<W> *)x, (const uint8_t *)key, (const WideInt<W> *)z,
                              (const uint8_t *)seed, party, (const WideInt<W> *)left, (const WideInt<W> *)right, I, n,
                              m, N));
    return launch_status();
}

int wide_spline_keygen(hipStream_t st, void *key, void *e, void *beta, const void *alpha, const void *seed0,
                       const void *seed1, const void *coef, int d, const void *left, const void *right, size_t I,
                       int n, int m, size_t es, size_t N)
{
    FSS_WIDE_DISPATCH(es, w_spline_keygen_kernel<W><<<FSS_WSHAPE(N)>>>(
                              (uint8_t *)key, (WideInt<W> *)e, (WideInt<W> *)beta, (const WideInt<W> *)alpha,
                              (const uint8_t *)seed0, (const uint8_t *)seed1, (const WideInt<W> *)coef, d,
                              (const WideInt<W> *)left, (const WideInt<W> *)right, I, n, m, N));
    return launch_status();
}

int wide_spline_eval(hipStream_t st, void *out, const void *x, const void *key, const void *e, const void *beta,
                     const void *seed, int party, const void *left, const void *right, size_t I, int d, int n, int m,
                     size_t es, size_t N, void *cache)
{
    FSS_WIDE_DISPATCH(es, w_spline_eval_kernel<W><<<FSS_WSHAPE(N)>>>(
                              (WideInt<W> *)out, (const WideInt<W> *)x, (const uint8_t *)key, (const WideInt<W> *)e,
                              (const WideInt<W> *)beta, (const uint8_t *)seed, party, (const WideInt<W> *)left,
                              (const WideInt<W> *)right, I, d, n, m, N, (uint8_t *)cache));
    return launch_status();
}

int wide_ottt_keygen(hipStream_t st, void *key, const void *alpha, int n, size_t es, size_t N)
{
    FSS_WIDE_DISPATCH(es, w_ottt_keygen_kernel<W><<<FSS_WSHAPE(N)>>>((uint8_t *)key, (const WideInt<W> *)alpha, n,
                                                                     N));
    return launch_status();
}

int wide_ottt_eval(hipStream_t st, void *outE, void *outT, const void *x, const void *key, int party, const void *lut,
                   size_t lutNum, int n, size_t es, size_t N)
{
    FSS_WIDE_DISPATCH(es, w_ottt_eval_kernel<W><<<FSS_WSHAPE(N)>>>(
                              (WideInt<W> *)outE, (WideInt<W> *)outT, (const WideInt<W> *)x, (const uint8_t *)key,
                              party, (const WideInt<W> *)lut, lutNum, n, N));
    return launch_status();
}

} // namespace fss
