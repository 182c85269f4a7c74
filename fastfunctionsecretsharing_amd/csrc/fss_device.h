// Device-side building blocks for the FSS kernels on gfx950 (CDNA4).
//
// AES-128 with a per-lane key (the PRG of every tree level is AES keyed by the
// node seed, reference src/impl/dcf.h:423-424 / src/impl/aes.h:466-469), written
// as a T-table cipher whose four tables live in LDS:
//
//   * State words are little-endian columns of the FIPS-197 byte array
//     (byte r of column c sits in bits 8r..8r+7 of word c), so 16-byte seeds
//     and blocks load and store with plain dword accesses.
//   * TE0[x] = 2S(x) | S(x)<<8 | S(x)<<16 | 3S(x)<<24 and TE_k = rotl(TE0, 8k),
//     so a column costs 4 lookups and 2 three-way XORs (v_bitop3_b32), the
//     round key folded into the second.
//   * Each table entry is replicated 32 times with lane-private copies: entry
//     x of table k for lane L is at byte x*256 + (k&1)*128 + (L%32)*4 +
//     (k>>1)*64K.  A wave's lookups are served 32 lanes at a time and each
//     half-wave hits 32 distinct banks whatever the data (measured: 32 and 64
//     replicas give the same lookup rate).  A single v_perm_b32 builds the LDS
//     address from the state byte and a per-lane, per-table constant.  The
//     four tables take 128 KB per workgroup.
//   * The last round and the key schedule take S(x) out of the table entries
//     (S sits in bytes 1-2 of TE0 and bytes 2-3 of TE1).
//
// The S-box is computed at compile time from the GF(2^8) inverse and the
// affine map of FIPS-197 section 5.1.1.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

// Opens a conditional block whose condition is wave-uniform and keeps it a
// branch: an empty volatile asm cannot be speculated, so the compiler does
// not flatten the block into per-lane selects of every value it assigns.
#ifndef FSS_UNIFORM_CASE
#define FSS_UNIFORM_CASE asm volatile("")
#endif

namespace fss {

// ---------------------------------------------------------------------------
// compile-time AES tables
// ---------------------------------------------------------------------------

struct SboxTable
{
    uint8_t s[256];
};

constexpr uint8_t gf_xtime(uint8_t a)
{
    return (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
}

constexpr uint8_t gf_mul3(uint8_t a)
{
    return (uint8_t)(gf_xtime(a) ^ a);
}

constexpr SboxTable make_sbox()
{
    // powers of the generator 3 give log/antilog tables; inverse = 3^(255-log x)
    uint8_t exp_t[256] = {};
    uint8_t log_t[256] = {};
    uint8_t p          = 1;
    for (int i = 0; i < 255; i++)
    {
        exp_t[i] = p;
        log_t[p] = (uint8_t)i;
        p        = gf_mul3(p);
    }
    SboxTable t = {};
    for (int x = 0; x < 256; x++)
    {
        uint8_t inv = x ? exp_t[(255 - log_t[x]) % 255] : 0;
        uint8_t s   = inv;
        uint8_t r   = inv;
        for (int k = 0; k < 4; k++)
        {
            r = (uint8_t)((r << 1) | (r >> 7));
            s ^= r;
        }
        t.s[x] = (uint8_t)(s ^ 0x63);
    }
    return t;
}

struct TeTable
{
    uint32_t v[256];
};

constexpr TeTable make_te0()
{
    SboxTable sb = make_sbox();
    TeTable   t  = {};
    for (int x = 0; x < 256; x++)
    {
        uint32_t s = sb.s[x];
        t.v[x]     = (uint32_t)gf_xtime((uint8_t)s) | (s << 8) | (s << 16) | ((uint32_t)gf_mul3((uint8_t)s) << 24);
    }
    return t;
}

__constant__ static const TeTable kTe0 = make_te0();

constexpr int kLdsTableWords = 2 * 256 * 64; // 128 KB

// ---------------------------------------------------------------------------
// bit helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// (a | b) ^ c
__device__ __forceinline__ uint32_t or_xor(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x56);
}

__device__ __forceinline__ uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel)
{
    return __builtin_amdgcn_perm(hi, lo, sel);
}

// ---------------------------------------------------------------------------
// LDS tables
// ---------------------------------------------------------------------------

// Fill both tables; every thread of the block takes part.  Call once per
// workgroup, followed by __syncthreads().
__device__ __forceinline__ void aes_fill_lds(uint32_t *lds)
{
    // word w: half (k>>1) = w>>14, entry x = (w>>6)&255, k&1 = (w>>5)&1, lane copy w&31
    for (int w = threadIdx.x; w < kLdsTableWords; w += blockDim.x)
    {
        const uint32_t k = ((w >> 14) << 1) | ((w >> 5) & 1);
        const uint32_t v = kTe0.v[(w >> 6) & 255];
        lds[w]           = k ? __builtin_amdgcn_alignbit(v, v, 32 - 8 * k) : v; // rotl(v, 8k)
    }
}

// aes_fill_lds for kernels that often run in small workgroups (the per-call
// evaluations below a full GPU's worth of elements, launched with as few as
// 64 lanes per workgroup): TE0 is staged in LDS first, so the workgroup pays
// one global-memory round trip for it instead of one per word written (512
// per lane at 64 lanes), and the lane copies are written 16 bytes (four
// copies of one entry) per store.  `lds` must be 16-byte aligned.  Kept apart
// from aes_fill_lds because the long-running persistent kernels measured
// 0.8% slower with it: the fill itself is negligible there, but it changes
// their main loop's register allocation and schedule.
__device__ __forceinline__ void aes_fill_lds_staged(uint32_t *lds)
{
    __shared__ uint32_t stage[256];
    for (int x = threadIdx.x; x < 256; x += blockDim.x) stage[x] = kTe0.v[x];
    __syncthreads();
    // chunk q = words 4q .. 4q+3: lane copies 4(q&7) .. +3 of entry x = (q>>4)&255
    // of table k = (q>>12)<<1 | (q>>3)&1
#pragma unroll 4
    for (int q = threadIdx.x; q < kLdsTableWords / 4; q += blockDim.x)
    {
        const uint32_t k  = (((uint32_t)q >> 12) << 1) | (((uint32_t)q >> 3) & 1u);
        const uint32_t v0 = stage[((uint32_t)q >> 4) & 255u];
        const uint32_t v  = __builtin_amdgcn_alignbit(v0, v0, (32u - 8u * k) & 31u); // rotl(v0, 8k)
        reinterpret_cast<uint4 *>(lds)[q] = make_uint4(v, v, v, v);
    }
}

struct AesCtx
{
    const char *base; // LDS table base
    uint32_t    l[4]; // byte offset of this lane's copy of table k (address bytes 0 and 2)
};

__device__ __forceinline__ AesCtx aes_ctx(const uint32_t *lds)
{
    AesCtx         c;
    const uint32_t lane = (threadIdx.x & 31u) * 4u;
    c.base              = (const char *)lds;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) c.l[k] = lane | ((k & 1u) << 7) | ((k >> 1) << 16);
    return c;
}

// LDS byte address of the table entry for byte k of s: [lane byte, s.byte k, table, 0]
#define FSS_SEL(k) (0x0C020000u | ((4u + (k)) << 8))

// TE_t[byte k of s]
template <int t>
__device__ __forceinline__ uint32_t te(const AesCtx &c, uint32_t s, int k)
{
    return *(const uint32_t *)(c.base + perm(s, c.l[t], FSS_SEL(k)));
}

__device__ __forceinline__ uint32_t te0(const AesCtx &c, uint32_t s, int k)
{
    return te<0>(c, s, k);
}

__device__ __forceinline__ uint32_t te1(const AesCtx &c, uint32_t s, int k)
{
    return te<1>(c, s, k);
}

// S-bytes of (a.b1, b.b1) into bytes 0,1 from two TE0 entries
#define FSS_SEL_S01 0x0C0C0105u
// S-bytes of (a.b2, b.b3) into bytes 2,3 from two TE1 entries
#define FSS_SEL_S23 0x03060C0Cu

// next AES-128 round key (little-endian words), round r in 1..10
__device__ __forceinline__ void aes_next_key(const AesCtx &c, uint32_t k[4], uint32_t rcon)
{
    uint32_t lo = perm(te0(c, k[3], 1), te0(c, k[3], 2), FSS_SEL_S01);
    uint32_t hi = perm(te1(c, k[3], 3), te1(c, k[3], 0), FSS_SEL_S23);
    k[0]        = or_xor(lo, hi, k[0]) ^ rcon;
    k[1] ^= k[0];
    k[2] ^= k[1];
    k[3] ^= k[2];
}

// one full round; h[c] = round key word c
__device__ __forceinline__ void aes_round(const AesCtx &c, uint32_t s[4], const uint32_t h[4])
{
    uint32_t n0 = xor3(te<0>(c, s[0], 0), te<1>(c, s[1], 1), xor3(te<2>(c, s[2], 2), te<3>(c, s[3], 3), h[0]));
    uint32_t n1 = xor3(te<0>(c, s[1], 0), te<1>(c, s[2], 1), xor3(te<2>(c, s[3], 2), te<3>(c, s[0], 3), h[1]));
    uint32_t n2 = xor3(te<0>(c, s[2], 0), te<1>(c, s[3], 1), xor3(te<2>(c, s[0], 2), te<3>(c, s[1], 3), h[2]));
    uint32_t n3 = xor3(te<0>(c, s[3], 0), te<1>(c, s[0], 1), xor3(te<2>(c, s[1], 2), te<3>(c, s[2], 3), h[3]));
    s[0] = n0, s[1] = n1, s[2] = n2, s[3] = n3;
}

__device__ __forceinline__ uint32_t aes_last_col(const AesCtx &c, uint32_t a, uint32_t b, uint32_t d, uint32_t e,
                                                 uint32_t k)
{
    uint32_t lo = perm(te0(c, a, 0), te0(c, b, 1), FSS_SEL_S01);
    uint32_t hi = perm(te1(c, d, 2), te1(c, e, 3), FSS_SEL_S23);
    return or_xor(lo, hi, k);
}

__device__ __forceinline__ void aes_last(const AesCtx &c, uint32_t s[4], const uint32_t k[4])
{
    uint32_t n0 = aes_last_col(c, s[0], s[1], s[2], s[3], k[0]);
    uint32_t n1 = aes_last_col(c, s[1], s[2], s[3], s[0], k[1]);
    uint32_t n2 = aes_last_col(c, s[2], s[3], s[0], s[1], k[2]);
    uint32_t n3 = aes_last_col(c, s[3], s[0], s[1], s[2], k[3]);
    s[0] = n0, s[1] = n1, s[2] = n2, s[3] = n3;
}

// ---- column-parallel AES on a quad of lanes ----
//
// Lane q (0..3) of a quad holds column q of the state and of the round key.
// A T-table round needs byte k of column q+k (ShiftRows); those columns come
// from the quad's other lanes through DPP quad permutations (VALU operand
// moves, no LDS), so each lane issues 4 table lookups per round instead of
// 16, and the key schedule's SubWord splits into one S-box lookup per lane.

// DPP quad_perm controls: lane q reads lane sel[q] of its quad
#define FSS_QP(a, b, c, d) ((a) | ((b) << 2) | ((c) << 4) | ((d) << 6))

template <int CTRL>
__device__ __forceinline__ uint32_t quad_mov(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false);
}

struct QuadAes
{
    uint32_t ksAddr; // this lane's key-schedule table offset and byte selector
    uint32_t ksSel;
    uint32_t ksMask; // 0xff << 8q
    uint32_t m1, m2; // all ones when q >= 1, q >= 2 (prefix XOR of the key words)

    __device__ __forceinline__ void init(const AesCtx &c, uint32_t q)
    {
        // byte q of SubWord(RotWord(w3)) is S[w3 byte (q+1)&3], read from the
        // rotated table whose entry holds S in byte q: TE_{(q+3)&3}
        const uint32_t t = (q + 3u) & 3u;
        ksAddr           = t == 0 ? c.l[0] : t == 1 ? c.l[1] : t == 2 ? c.l[2] : c.l[3];
        ksSel            = 0x0C020000u | ((4u + ((q + 1u) & 3u)) << 8);
        ksMask           = 0xFFu << (8u * q);
        m1               = q >= 1 ? ~0u : 0u;
        m2               = q >= 2 ? ~0u : 0u;
    }

    // this lane's column of the next round key
    __device__ __forceinline__ uint32_t next_key(const AesCtx &c, uint32_t k, uint32_t rcon) const
    {
        const uint32_t w3 = quad_mov<FSS_QP(3, 3, 3, 3)>(k);
        uint32_t       sb = *(const uint32_t *)(c.base + perm(w3, ksAddr, ksSel)) & ksMask;
        sb |= quad_mov<FSS_QP(1, 0, 3, 2)>(sb);
        sb |= quad_mov<FSS_QP(2, 3, 0, 1)>(sb);
        // k'_q = SubWord(RotWord(k_3)) ^ rcon ^ k_0 ^ ... ^ k_q
        uint32_t p = k;
        p ^= quad_mov<FSS_QP(0, 0, 1, 2)>(p) & m1;
        p ^= quad_mov<FSS_QP(0, 0, 0, 1)>(p) & m2;
        return xor3(p, sb, rcon);
    }

    __device__ __forceinline__ static uint32_t round(const AesCtx &c, uint32_t s, uint32_t k)
    {
        const uint32_t a1 = quad_mov<FSS_QP(1, 2, 3, 0)>(s);
        const uint32_t a2 = quad_mov<FSS_QP(2, 3, 0, 1)>(s);
        const uint32_t a3 = quad_mov<FSS_QP(3, 0, 1, 2)>(s);
        return xor3(te<0>(c, s, 0), te<1>(c, a1, 1), xor3(te<2>(c, a2, 2), te<3>(c, a3, 3), k));
    }

    __device__ __forceinline__ static uint32_t last(const AesCtx &c, uint32_t s, uint32_t k)
    {
        const uint32_t a1 = quad_mov<FSS_QP(1, 2, 3, 0)>(s);
        const uint32_t a2 = quad_mov<FSS_QP(2, 3, 0, 1)>(s);
        const uint32_t a3 = quad_mov<FSS_QP(3, 0, 1, 2)>(s);
        return aes_last_col(c, s, a1, a2, a3, k);
    }

    // AES-128 of this lane's plaintext column under the key column `key`
    __device__ __forceinline__ uint32_t encrypt(const AesCtx &c, uint32_t key, uint32_t pt) const
    {
        const uint32_t rcon[10] = {0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80, 0x1b, 0x36};
        uint32_t       k        = key;
        uint32_t       s        = pt ^ k;
#pragma unroll
        for (int r = 0; r < 9; r++)
        {
            k = next_key(c, k, rcon[r]);
            s = round(c, s, k);
        }
        k = next_key(c, k, rcon[9]);
        return last(c, s, k);
    }
};

// Encrypt NB blocks (in place) under the 16-byte key `key`, key schedule on the fly.
template <int NB>
__device__ __forceinline__ void aes_encrypt(const AesCtx &c, const uint32_t key[4], uint32_t st[NB][4])
{
    uint32_t k[4] = {key[0], key[1], key[2], key[3]};
#pragma unroll
    for (int b = 0; b < NB; b++)
    {
        st[b][0] ^= k[0], st[b][1] ^= k[1], st[b][2] ^= k[2], st[b][3] ^= k[3];
    }
    const uint32_t rcon[10] = {0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80, 0x1b, 0x36};
#pragma unroll
    for (int r = 0; r < 9; r++)
    {
        aes_next_key(c, k, rcon[r]);
        const uint32_t *h = k;
#pragma unroll
        for (int b = 0; b < NB; b++) aes_round(c, st[b], h);
    }
    aes_next_key(c, k, rcon[9]);
#pragma unroll
    for (int b = 0; b < NB; b++) aes_last(c, st[b], k);
}

// The first `nb` (<= NB, wave-uniform) of NB blocks under one key schedule;
// blocks nb..NB-1 are left as they are.
template <int NB>
__device__ __forceinline__ void aes_encrypt_upto(const AesCtx &c, const uint32_t key[4], uint32_t st[NB][4], int nb)
{
    uint32_t k[4] = {key[0], key[1], key[2], key[3]};
#pragma unroll
    for (int b = 0; b < NB; b++)
        if (b < nb) st[b][0] ^= k[0], st[b][1] ^= k[1], st[b][2] ^= k[2], st[b][3] ^= k[3];
    const uint32_t rcon[10] = {0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80, 0x1b, 0x36};
#pragma unroll
    for (int r = 0; r < 9; r++)
    {
        aes_next_key(c, k, rcon[r]);
        const uint32_t *h = k;
#pragma unroll
        for (int b = 0; b < NB; b++)
            if (b < nb) aes_round(c, st[b], h);
    }
    aes_next_key(c, k, rcon[9]);
#pragma unroll
    for (int b = 0; b < NB; b++)
        if (b < nb) aes_last(c, st[b], k);
}

// E independent keys, NB blocks each, rounds interleaved across keys so the
// lookups of all E*NB states are in flight together.
template <int E, int NB>
__device__ __forceinline__ void aes_encrypt_multi(const AesCtx &c, const uint32_t key[E][4], uint32_t st[E][NB][4])
{
    uint32_t k[E][4];
#pragma unroll
    for (int e = 0; e < E; e++)
    {
#pragma unroll
        for (int i = 0; i < 4; i++) k[e][i] = key[e][i];
#pragma unroll
        for (int b = 0; b < NB; b++)
            st[e][b][0] ^= k[e][0], st[e][b][1] ^= k[e][1], st[e][b][2] ^= k[e][2], st[e][b][3] ^= k[e][3];
    }
    const uint32_t rcon[10] = {0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80, 0x1b, 0x36};
#pragma unroll
    for (int r = 0; r < 9; r++)
    {
#pragma unroll
        for (int e = 0; e < E; e++)
        {
            aes_next_key(c, k[e], rcon[r]);
            const uint32_t *h = k[e];
#pragma unroll
            for (int b = 0; b < NB; b++) aes_round(c, st[e][b], h);
        }
    }
#pragma unroll
    for (int e = 0; e < E; e++)
    {
        aes_next_key(c, k[e], rcon[9]);
#pragma unroll
        for (int b = 0; b < NB; b++) aes_last(c, st[e][b], k[e]);
    }
}

// NB blocks under one key whose plaintexts are (p0[q], 0, p2[q], 0) as
// little-endian u32 words with every p0[q], p2[q] < 256: the PRG counter
// blocks of a tree level (DCF eval: (4x, 4x+1), (4x+2, 4x+3); keyGen:
// (2q, 2q+1), q = 0..3).  After the whitening the blocks differ only in byte 0
// of words 0 and 2.  Round 1 output word w reads byte 0 of word w and bytes 1-3
// of the other words, so words 1 and 3 come out equal for every block and
// words 0 and 2 differ only in their TE0 term: round 1 costs 14 shared table
// lookups plus 2 per block instead of 16 per block.  Rounds 2-10 are the
// ordinary ones.  Same output as aes_encrypt<NB> on those plaintexts.
template <int NB>
__device__ __forceinline__ void aes_encrypt_lowpt(const AesCtx &c, const uint32_t key[4], const uint32_t p0[NB],
                                                  const uint32_t p2[NB], uint32_t st[NB][4])
{
    uint32_t       k[4] = {key[0], key[1], key[2], key[3]};
    const uint32_t a0 = k[0], a1 = k[1], a2 = k[2], a3 = k[3]; // bytes 1-3 of the whitened words 0, 2
    const uint32_t rcon[10] = {0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80, 0x1b, 0x36};
    aes_next_key(c, k, rcon[0]);
    const uint32_t r0 = xor3(te<1>(c, a1, 1), te<2>(c, a2, 2), te<3>(c, a3, 3)) ^ k[0];
    const uint32_t r1 = xor3(te<0>(c, a1, 0), te<1>(c, a2, 1), xor3(te<2>(c, a3, 2), te<3>(c, a0, 3), k[1]));
    const uint32_t r2 = xor3(te<1>(c, a3, 1), te<2>(c, a0, 2), te<3>(c, a1, 3)) ^ k[2];
    const uint32_t r3 = xor3(te<0>(c, a3, 0), te<1>(c, a0, 1), xor3(te<2>(c, a1, 2), te<3>(c, a2, 3), k[3]));
#pragma unroll
    for (int q = 0; q < NB; q++)
    {
        st[q][0] = r0 ^ te<0>(c, p0[q] ^ a0, 0);
        st[q][1] = r1;
        st[q][2] = r2 ^ te<0>(c, p2[q] ^ a2, 0);
        st[q][3] = r3;
    }
#pragma unroll
    for (int r = 1; r < 9; r++)
    {
        aes_next_key(c, k, rcon[r]);
#pragma unroll
        for (int q = 0; q < NB; q++) aes_round(c, st[q], k);
    }
    aes_next_key(c, k, rcon[9]);
#pragma unroll
    for (int q = 0; q < NB; q++) aes_last(c, st[q], k);
}

// aes_encrypt_lowpt for the conversion windows: plaintexts (0, 0, p2[q], 0)
// with every p2[q] < 256, only the first `nb` (<= NB, wave-uniform) blocks.
// The blocks differ only in byte 0 of word 2, so round 1 costs 15 shared
// lookups plus 1 per block.
template <int NB>
__device__ __forceinline__ void aes_encrypt_win_upto(const AesCtx &c, const uint32_t key[4], const uint32_t p2[NB],
                                                     uint32_t st[NB][4], int nb)
{
    uint32_t       k[4] = {key[0], key[1], key[2], key[3]};
    const uint32_t a0 = k[0], a1 = k[1], a2 = k[2], a3 = k[3];
    const uint32_t rcon[10] = {0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80, 0x1b, 0x36};
    aes_next_key(c, k, rcon[0]);
    const uint32_t r0 = xor3(te<0>(c, a0, 0), te<1>(c, a1, 1), xor3(te<2>(c, a2, 2), te<3>(c, a3, 3), k[0]));
    const uint32_t r1 = xor3(te<0>(c, a1, 0), te<1>(c, a2, 1), xor3(te<2>(c, a3, 2), te<3>(c, a0, 3), k[1]));
    const uint32_t r2 = xor3(te<1>(c, a3, 1), te<2>(c, a0, 2), te<3>(c, a1, 3)) ^ k[2];
    const uint32_t r3 = xor3(te<0>(c, a3, 0), te<1>(c, a0, 1), xor3(te<2>(c, a1, 2), te<3>(c, a2, 3), k[3]));
#pragma unroll
    for (int q = 0; q < NB; q++)
        if (q < nb) st[q][0] = r0, st[q][1] = r1, st[q][2] = r2 ^ te<0>(c, p2[q] ^ a2, 0), st[q][3] = r3;
#pragma unroll
    for (int r = 1; r < 9; r++)
    {
        aes_next_key(c, k, rcon[r]);
#pragma unroll
        for (int q = 0; q < NB; q++)
            if (q < nb) aes_round(c, st[q], k);
    }
    aes_next_key(c, k, rcon[9]);
#pragma unroll
    for (int q = 0; q < NB; q++)
        if (q < nb) aes_last(c, st[q], k);
}

// The two blocks a DCF walk needs at one level (x = input bit).
__device__ __forceinline__ void aes_encrypt_dcf_pair(const AesCtx &c, const uint32_t key[4], uint32_t xi,
                                                     uint32_t st[2][4])
{
    const uint32_t p0[2] = {4u * xi, 4u * xi + 2u}, p2[2] = {4u * xi + 1u, 4u * xi + 3u};
    aes_encrypt_lowpt<2>(c, key, p0, p2, st);
}

// The four counter blocks of a keyGen level: plaintexts (2q, 2q+1), q = 0..3.
__device__ __forceinline__ void aes_encrypt_prg4(const AesCtx &c, const uint32_t key[4], uint32_t st[4][4])
{
    const uint32_t p0[4] = {0u, 2u, 4u, 6u}, p2[4] = {1u, 3u, 5u, 7u};
    aes_encrypt_lowpt<4>(c, key, p0, p2, st);
}

// The two counter blocks of a DPF/Grotto keyGen level: (0, 1) and (2, 3).
__device__ __forceinline__ void aes_encrypt_prg2(const AesCtx &c, const uint32_t key[4], uint32_t st[2][4])
{
    const uint32_t p0[2] = {0u, 2u}, p2[2] = {1u, 3u};
    aes_encrypt_lowpt<2>(c, key, p0, p2, st);
}

// ---------------------------------------------------------------------------
// PRG and value conversion
// ---------------------------------------------------------------------------

// Plaintext of counter block q: little-endian u64 pair (2q, 2q+1).
__device__ __forceinline__ void prg_pt(uint32_t p[4], uint32_t q)
{
    p[0] = 2u * q, p[1] = 0u, p[2] = 2u * q + 1u, p[3] = 0u;
}

// Low `bytes` bytes (1..8) of a 16-byte block as an integer.
__device__ __forceinline__ uint64_t low_bytes(const uint32_t v[4], int bytes)
{
    uint64_t w = (uint64_t)v[0] | ((uint64_t)v[1] << 32);
    return bytes >= 8 ? w : (w & ((1ull << (8 * bytes)) - 1));
}

__device__ __forceinline__ uint32_t sel_word(const uint32_t v[4], int i)
{
    uint32_t r = v[0];
    r          = (i == 1) ? v[1] : r;
    r          = (i == 2) ? v[2] : r;
    r          = (i == 3) ? v[3] : r;
    return r;
}

__device__ __forceinline__ uint32_t byte_of(const uint32_t v[4], int idx)
{
    return (sel_word(v, idx >> 2) >> (8 * (idx & 3))) & 0xffu;
}

// AES_key(u64 0 || u64 idx): the conversion's refill block, kept out of line
// so unrolled conversion loops stay small (the groupSize-1 paths never call it).
__device__ __noinline__ static uint4 conv_refill(AesCtx c, uint32_t k0, uint32_t k1, uint32_t k2, uint32_t k3,
                                                uint32_t idx)
{
    const uint32_t key[4] = {k0, k1, k2, k3};
    uint32_t       blk[1][4] = {{0u, 0u, idx, 0u}};
    aes_encrypt<1>(c, key, blk);
    return make_uint4(blk[0][0], blk[0][1], blk[0][2], blk[0][3]);
}

// Group-element conversion of a 16-byte value (reference DcfConvertCtx,
// src/impl/dcf.h:69-137).  Element j of a group of `count` elements of
// `bytes` bytes: if bytes*count <= 16 it is bytes [j*bytes, (j+1)*bytes) of the
// value; otherwise byte offset o = j*bytes+i comes from the block that the
// reference refilled at the start of o's 16-byte window, AES_value(0, e) with
// e = (16*floor(o/16)) / bytes the element being produced at that moment.
// The struct caches the last block so sequential elements cost one AES block
// per 16 bytes, like the reference.
struct ConvCtx
{
    uint32_t key[4];
    uint32_t buf[4];
    int      bytes;
    int      direct;
    int      bufBlock; // window index held in buf (-1: none)

    __device__ __forceinline__ void init(const uint32_t v[4], int count, int bitWidth)
    {
        bytes    = (bitWidth + 7) / 8;
        direct   = bytes * count <= 16;
        bufBlock = -1;
#pragma unroll
        for (int i = 0; i < 4; i++) key[i] = v[i], buf[i] = v[i];
    }

    __device__ __forceinline__ uint64_t at(const AesCtx &c, int j)
    {
        uint64_t r = 0;
        for (int i = 0; i < bytes; i++)
        {
            int o = j * bytes + i;
            if (!direct)
            {
                int win = o >> 4;
                if (win != bufBlock)
                {
                    uint4 b  = conv_refill(c, key[0], key[1], key[2], key[3], (uint32_t)((win * 16) / bytes));
                    buf[0]   = b.x, buf[1] = b.y, buf[2] = b.z, buf[3] = b.w;
                    bufBlock = win;
                }
            }
            if (i < 8) r |= (uint64_t)byte_of(buf, o & 15) << (8 * i);
        }
        return r;
    }
};

// Value conversion restricted to the slots [j0, j0 + cnt) of a g-element
// group with OB = ceil(bitWidthOut / 8) bytes per element (compile time); the
// same bytes as ConvCtx::at (reference DcfConvertCtx, src/impl/dcf.h:69-137).
// When OB * g > 16 the slots' bytes come from the AES windows w0 .. w0+nw-1
// (window w = AES_value(0, 16w / OB)), all encrypted under one key schedule;
// at most NW windows (the caller guarantees (j0*OB mod 16) + cnt*OB <= 16*NW).
// fill() leaves the bytes starting at slot j0 in words[], so slot j0 + r is
// bytes [r*OB, (r+1)*OB) of words[] -- static positions.
template <int OB, int NW>
struct SlotConv
{
    int direct, w0, nw, b0; // wave-uniform

    __device__ __forceinline__ void setup(int g, int j0, int cnt)
    {
        const int start = j0 * OB, end = (j0 + cnt) * OB;
        direct          = OB * g <= 16;
        w0              = direct ? 0 : (start >> 4);
        nw              = direct ? 1 : (((end - 1) >> 4) - w0 + 1);
        b0              = start - 16 * w0;
    }

    template <int NWORDS>
    __device__ __forceinline__ void fill(const AesCtx &c, const uint32_t v[4], uint32_t words[NWORDS]) const
    {
        static_assert(NWORDS + 4 <= 4 * (NW + 1), "window buffer too small");
        uint32_t w[NW + 1][4];
        if (direct)
        {
#pragma unroll
            for (int i = 0; i < 4; i++) w[0][i] = v[i];
#pragma unroll
            for (int k = 1; k <= NW; k++)
#pragma unroll
                for (int i = 0; i < 4; i++) w[k][i] = 0;
        }
        else
        {
            uint32_t blk[NW][4];
            // (only for 4- and 8-byte slots: with narrower slots the extra
            // shared terms push the spline kernel past its register budget)
            if ((OB == 4 || OB == 8) && (16 * (w0 + NW - 1)) / OB < 256)
            {
                uint32_t p2[NW];
#pragma unroll
                for (int k = 0; k < NW; k++) p2[k] = (uint32_t)((16 * (w0 + k)) / OB);
                aes_encrypt_win_upto<NW>(c, v, p2, blk, nw);
            }
            else
            {
#pragma unroll
                for (int k = 0; k < NW; k++)
                    blk[k][0] = 0u, blk[k][1] = 0u, blk[k][2] = (uint32_t)((16 * (w0 + k)) / OB), blk[k][3] = 0u;
                aes_encrypt_upto<NW>(c, v, blk, nw);
            }
#pragma unroll
            for (int k = 0; k < NW; k++)
#pragma unroll
                for (int i = 0; i < 4; i++) w[k][i] = blk[k][i];
#pragma unroll
            for (int i = 0; i < 4; i++) w[NW][i] = 0;
        }
        // shift the window bytes left by b0 (0..15): whole words by b0/4, then bytes by b0%4
        const int      ws = b0 >> 2;
        const uint32_t bs = (uint32_t)(b0 & 3);
#pragma unroll
        for (int q = 0; q < NWORDS; q++)
        {
            uint32_t lo = 0, hi = 0;
#pragma unroll
            for (int d = 0; d < 4; d++)
            {
                const int i = q + d;
                if (ws == d)
                {
                    lo = w[i >> 2][i & 3];
                    hi = w[(i + 1) >> 2][(i + 1) & 3];
                }
            }
            words[q] = __builtin_amdgcn_alignbyte(hi, lo, bs);
        }
    }

    // slot j0 + r from the filled words (r static)
    template <int NWORDS>
    __device__ __forceinline__ static uint64_t at(const uint32_t words[NWORDS], int r)
    {
        uint64_t v = 0;
#pragma unroll
        for (int i = 0; i < OB; i++)
        {
            const int o = r * OB + i;
            v |= (uint64_t)((words[o >> 2] >> (8 * (o & 3))) & 0xffu) << (8 * i);
        }
        return v;
    }
};

// ---------------------------------------------------------------------------
// work distribution
// ---------------------------------------------------------------------------

// body(i) for every work item i in [0, count), in 64-item tiles: wave w takes
// tile t and lane L item 64*t + L.  With a queue counter (zeroed before the
// launch) each wave takes the next tile from it until none are left; without
// one, wave w takes tiles w, w + W, w + 2W, ... (W waves in the grid).  One
// loop body serves both, so the kernel is compiled once.  Every lane of a
// wave reaches the loop head together, so the tile read by lane 0 is the
// wave's.
template <typename Body>
__device__ __forceinline__ void for_each_item(size_t count, uint32_t *counter, Body &&body)
{
    const uint32_t lane   = threadIdx.x & 63u;
    const size_t   tiles  = (count + 63) / 64;
    const size_t   waves  = ((size_t)gridDim.x * blockDim.x) >> 6;
    size_t         stat   = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    for (;;)
    {
        size_t tile;
        if (counter)
        {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(counter, 1u);
            tile = __builtin_amdgcn_readfirstlane(t);
        }
        else
        {
            tile = stat;
            stat += waves;
        }
        if (tile >= tiles) break;
        const size_t i = tile * 64 + lane;
        if (i < count) body(i);
    }
}

// ---------------------------------------------------------------------------
// memory helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ void load16(uint32_t d[4], const uint8_t *p)
{
    uint2 a = *(const uint2 *)p;
    uint2 b = *(const uint2 *)(p + 8);
    d[0] = a.x, d[1] = a.y, d[2] = b.x, d[3] = b.y;
}

__device__ __forceinline__ void store16(uint8_t *p, const uint32_t d[4])
{
    *(uint2 *)p       = make_uint2(d[0], d[1]);
    *(uint2 *)(p + 8) = make_uint2(d[2], d[3]);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// one 16-byte store to a 16-byte aligned address
__device__ __forceinline__ void store16_aligned(uint8_t *p, const uint32_t d[4])
{
    *(u32x4 *)p = u32x4{d[0], d[1], d[2], d[3]};
}

template <typename T>
__device__ __forceinline__ T mod_bits(T x, int bw)
{
    return bw >= (int)(8 * sizeof(T)) ? x : (T)(x & (T)(((T)1 << bw) - 1));
}

// Kernel prologue: fill this workgroup's AES tables in LDS, then bind them.
#define FSS_KERNEL_PROLOGUE                      \
    __shared__ uint32_t lds[kLdsTableWords]; \
    aes_fill_lds(lds);                           \
    __syncthreads();                             \
    const AesCtx c = aes_ctx(lds);

// the same for kernels launched with workgroups below 1024 lanes, or whose
// whole launch lasts about 0.1 ms (aes_fill_lds_staged)
#define FSS_KERNEL_PROLOGUE_STAGED                         \
    __shared__ __align__(16) uint32_t lds[kLdsTableWords]; \
    aes_fill_lds_staged(lds);                              \
    __syncthreads();                                       \
    const AesCtx c = aes_ctx(lds);

#define FSS_GRID_LOOP(i, count)                                                  \
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (count); \
         i += (size_t)gridDim.x * blockDim.x)

} // namespace fss
