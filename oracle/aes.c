/*
 * Oracle AES-128 (test infrastructure only: never linked into the product).
 *
 * Two independent encryptors of the same FIPS-197 cipher:
 *   - a portable byte-oriented one derived from the specification (S-box
 *     computed from the GF(2^8) inverse and the affine map, MixColumns by
 *     xtime), used as the cross-check;
 *   - an AES-NI one (x86 only), used for speed by the oracle's bulk paths,
 *     matching the reference CPU backend's choice of AES-NI
 *     (src/impl/aesni.h:42-79 key expansion, :109-144 block encrypt).
 *
 * Byte order: keys, plaintexts and ciphertexts are plain 16-byte FIPS-197
 * arrays, exactly as the reference feeds them (src/impl/aesni.h:118 loads the
 * raw bytes; src/impl/aes.h:10-18 loads big-endian words of the same bytes).
 */
#include "fss_oracle.h"

#include <string.h>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

static uint8_t g_sbox[256];
static int     g_sbox_ready = 0;

static uint8_t gf_mul(uint8_t a, uint8_t b)
{
    uint8_t r = 0;
    while (b)
    {
        if (b & 1) r ^= a;
        a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
        b >>= 1;
    }
    return r;
}

static void sbox_init(void)
{
    if (g_sbox_ready) return;
    for (int x = 0; x < 256; x++)
    {
        /* multiplicative inverse in GF(2^8) mod x^8+x^4+x^3+x+1 (0 -> 0) */
        uint8_t inv = 0;
        if (x)
        {
            for (int y = 1; y < 256; y++)
            {
                if (gf_mul((uint8_t)x, (uint8_t)y) == 1)
                {
                    inv = (uint8_t)y;
                    break;
                }
            }
        }
        /* affine transform (FIPS-197 eq. 5.1) */
        uint8_t s = inv;
        uint8_t r = inv;
        for (int k = 0; k < 4; k++)
        {
            r = (uint8_t)((r << 1) | (r >> 7));
            s ^= r;
        }
        g_sbox[x] = (uint8_t)(s ^ 0x63);
    }
    g_sbox_ready = 1;
}

uint8_t oracle_aes_sbox(uint8_t x)
{
    sbox_init();
    return g_sbox[x];
}

void oracle_aes128_expand_portable(const uint8_t key[16], uint8_t rk[176])
{
    sbox_init();
    memcpy(rk, key, 16);
    uint8_t rcon = 1;
    for (int i = 4; i < 44; i++)
    {
        uint8_t t[4];
        memcpy(t, rk + 4 * (i - 1), 4);
        if (i % 4 == 0)
        {
            uint8_t t0 = t[0];
            t[0]       = (uint8_t)(g_sbox[t[1]] ^ rcon);
            t[1]       = g_sbox[t[2]];
            t[2]       = g_sbox[t[3]];
            t[3]       = g_sbox[t0];
            rcon       = gf_mul(rcon, 2);
        }
        for (int b = 0; b < 4; b++) rk[4 * i + b] = (uint8_t)(rk[4 * (i - 4) + b] ^ t[b]);
    }
}

void oracle_aes128_encrypt_portable(const uint8_t rk[176], const uint8_t in[16], uint8_t out[16])
{
    sbox_init();
    uint8_t s[16];
    for (int b = 0; b < 16; b++) s[b] = (uint8_t)(in[b] ^ rk[b]);
    for (int round = 1; round <= 10; round++)
    {
        uint8_t t[16];
        /* SubBytes + ShiftRows: state byte (row r, col c) lives at s[4c+r] */
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < 4; r++) t[4 * c + r] = g_sbox[s[4 * ((c + r) & 3) + r]];
        if (round != 10)
        {
            for (int c = 0; c < 4; c++)
            {
                uint8_t a0 = t[4 * c], a1 = t[4 * c + 1], a2 = t[4 * c + 2], a3 = t[4 * c + 3];
                s[4 * c + 0] = (uint8_t)(gf_mul(a0, 2) ^ gf_mul(a1, 3) ^ a2 ^ a3);
                s[4 * c + 1] = (uint8_t)(a0 ^ gf_mul(a1, 2) ^ gf_mul(a2, 3) ^ a3);
                s[4 * c + 2] = (uint8_t)(a0 ^ a1 ^ gf_mul(a2, 2) ^ gf_mul(a3, 3));
                s[4 * c + 3] = (uint8_t)(gf_mul(a0, 3) ^ a1 ^ a2 ^ gf_mul(a3, 2));
            }
        }
        else
        {
            memcpy(s, t, 16);
        }
        for (int b = 0; b < 16; b++) s[b] ^= rk[16 * round + b];
    }
    memcpy(out, s, 16);
}

#if defined(__x86_64__)

__attribute__((target("aes,sse2"))) static inline __m128i ni_expand_step(__m128i k, __m128i assist)
{
    assist = _mm_shuffle_epi32(assist, 0xff);
    k      = _mm_xor_si128(k, _mm_slli_si128(k, 4));
    k      = _mm_xor_si128(k, _mm_slli_si128(k, 4));
    k      = _mm_xor_si128(k, _mm_slli_si128(k, 4));
    return _mm_xor_si128(k, assist);
}

#define NI_STEP(i, rc)                                               \
    k       = ni_expand_step(k, _mm_aeskeygenassist_si128(k, (rc))); \
    rks[(i)] = k;

__attribute__((target("aes,sse2"))) static void ni_expand(const uint8_t key[16], __m128i rks[11])
{
    __m128i k = _mm_loadu_si128((const __m128i *)key);
    rks[0]    = k;
    NI_STEP(1, 0x01)
    NI_STEP(2, 0x02)
    NI_STEP(3, 0x04)
    NI_STEP(4, 0x08)
    NI_STEP(5, 0x10)
    NI_STEP(6, 0x20)
    NI_STEP(7, 0x40)
    NI_STEP(8, 0x80)
    NI_STEP(9, 0x1b)
    NI_STEP(10, 0x36)
}

__attribute__((target("aes,sse2"))) static void ni_encrypt_blocks(const uint8_t key[16],
                                                                  const uint8_t *in,
                                                                  uint8_t       *out,
                                                                  size_t         nblocks)
{
    __m128i rks[11];
    ni_expand(key, rks);
    size_t b = 0;
    for (; b + 4 <= nblocks; b += 4)
    {
        __m128i s0 = _mm_xor_si128(_mm_loadu_si128((const __m128i *)(in + 16 * b)), rks[0]);
        __m128i s1 = _mm_xor_si128(_mm_loadu_si128((const __m128i *)(in + 16 * b + 16)), rks[0]);
        __m128i s2 = _mm_xor_si128(_mm_loadu_si128((const __m128i *)(in + 16 * b + 32)), rks[0]);
        __m128i s3 = _mm_xor_si128(_mm_loadu_si128((const __m128i *)(in + 16 * b + 48)), rks[0]);
        for (int r = 1; r < 10; r++)
        {
            s0 = _mm_aesenc_si128(s0, rks[r]);
            s1 = _mm_aesenc_si128(s1, rks[r]);
            s2 = _mm_aesenc_si128(s2, rks[r]);
            s3 = _mm_aesenc_si128(s3, rks[r]);
        }
        _mm_storeu_si128((__m128i *)(out + 16 * b), _mm_aesenclast_si128(s0, rks[10]));
        _mm_storeu_si128((__m128i *)(out + 16 * b + 16), _mm_aesenclast_si128(s1, rks[10]));
        _mm_storeu_si128((__m128i *)(out + 16 * b + 32), _mm_aesenclast_si128(s2, rks[10]));
        _mm_storeu_si128((__m128i *)(out + 16 * b + 48), _mm_aesenclast_si128(s3, rks[10]));
    }
    for (; b < nblocks; b++)
    {
        __m128i s = _mm_xor_si128(_mm_loadu_si128((const __m128i *)(in + 16 * b)), rks[0]);
        for (int r = 1; r < 10; r++) s = _mm_aesenc_si128(s, rks[r]);
        _mm_storeu_si128((__m128i *)(out + 16 * b), _mm_aesenclast_si128(s, rks[10]));
    }
}

static int have_aesni(void)
{
    static int cached = -1;
    if (cached < 0)
    {
        __builtin_cpu_init();
        cached = __builtin_cpu_supports("aes") ? 1 : 0;
    }
    return cached;
}
#else
static int have_aesni(void)
{
    return 0;
}
#endif

static int g_force_portable = 0;

void oracle_aes_force_portable(int on)
{
    g_force_portable = on;
}

int oracle_aes_uses_aesni(void)
{
    return have_aesni() && !g_force_portable;
}

void oracle_aes128_encrypt_blocks(const uint8_t key[16], const uint8_t *in, uint8_t *out, size_t nblocks)
{
#if defined(__x86_64__)
    if (oracle_aes_uses_aesni())
    {
        ni_encrypt_blocks(key, in, out, nblocks);
        return;
    }
#endif
    uint8_t rk[176];
    oracle_aes128_expand_portable(key, rk);
    for (size_t b = 0; b < nblocks; b++) oracle_aes128_encrypt_portable(rk, in + 16 * b, out + 16 * b);
}
