/*
 * FSS oracle: DCF / DPF / Grotto / MIC / spline / OTTT key generation and
 * evaluation, size queries and the AES-CTR PRNG, restated in plain C from the
 * reference CPU path.  TEST INFRASTRUCTURE ONLY (see fss_oracle.h).
 *
 * Group elements are carried in uint64_t and reduced to the element width
 * (8*elementSize bits) whenever the reference would store them in its
 * GroupElement type, which reproduces the reference's modular arithmetic for
 * 1/2/4/8-byte elements.
 */
#include <stdlib.h>
#include <string.h>

#include "fss_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* include/FastFss/errors.h:4-27 */
enum
{
    E_OK            = 0,
    E_RUNTIME       = -1,
    E_KEY_SIZE      = -2,
    E_ALPHA_SIZE    = -3,
    E_BETA_SIZE     = -4,
    E_SEED_SIZE     = -5,
    E_BOUNDARY_SIZE = -6,
    E_Z_SIZE        = -7,
    E_OUT_SIZE      = -8,
    E_X_SIZE        = -9,
    E_BITWIDTH      = -10,
    E_ELEM_SIZE     = -11,
    E_PARTY         = -12,
    E_CACHE_SIZE    = -13,
    E_POINT_SIZE    = -14,
    E_LUT_SIZE      = -15,
    E_GROUP_SIZE    = -16,
    E_PRNG_ARG      = -17,
    E_PRNG_BW       = -18,
};

/* ------------------------------------------------------------------------ */
/* group-element helpers                                                     */
/* ------------------------------------------------------------------------ */

static int es_ok(size_t es)
{
    return es == 1 || es == 2 || es == 4 || es == 8;
}

static uint64_t ge_mask(size_t es)
{
    return es >= 8 ? ~0ULL : ((1ULL << (8 * es)) - 1);
}

/* src/impl/number.h:30-42 (modBits) */
static uint64_t mod_bits(uint64_t x, size_t bw)
{
    return bw >= 64 ? x : (x & ((1ULL << bw) - 1));
}

static uint64_t ld(const void *p, size_t es, size_t idx)
{
    switch (es)
    {
        case 1: return ((const uint8_t *)p)[idx];
        case 2: return ((const uint16_t *)p)[idx];
        case 4: return ((const uint32_t *)p)[idx];
        default: return ((const uint64_t *)p)[idx];
    }
}

static void st(void *p, size_t es, size_t idx, uint64_t v)
{
    switch (es)
    {
        case 1: ((uint8_t *)p)[idx] = (uint8_t)v; break;
        case 2: ((uint16_t *)p)[idx] = (uint16_t)v; break;
        case 4: ((uint32_t *)p)[idx] = (uint32_t)v; break;
        default: ((uint64_t *)p)[idx] = v; break;
    }
}

/* src/impl/number.h:44-51 (parityU64) */
static int parity64(uint64_t x)
{
    return __builtin_parityll(x);
}

/* ------------------------------------------------------------------------ */
/* threading (src/cpu/config.cpp:6-38)                                       */
/* ------------------------------------------------------------------------ */

static int g_threads = 0;

int oracle_setNumThreads(int num)
{
    if (num <= 0) return -2;
    g_threads = num;
    return 0;
}

int oracle_getNumThreads(void)
{
    if (g_threads > 0) return g_threads;
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* PRG: seed-keyed AES-128 on little-endian (u64,u64) counters               */
/* (src/impl/dcf.h:197-198,384-387,423-434; dpf.h:183,302,331-339)           */
/* ------------------------------------------------------------------------ */

typedef struct
{
    uint64_t lo, hi; /* 16 bytes, lo = bytes 0..7 little-endian */
} blk_t;

#define LSB_CLEAR 0xFFFFFFFFFFFFFFFEULL /* MASK_MSB63, src/impl/dcf.h:197 */

static blk_t blk_load(const uint8_t *p)
{
    blk_t b;
    memcpy(&b.lo, p, 8);
    memcpy(&b.hi, p + 8, 8);
    return b;
}

static void blk_store(uint8_t *p, blk_t b)
{
    memcpy(p, &b.lo, 8);
    memcpy(p + 8, &b.hi, 8);
}

/* encrypt `n` counter blocks {2k, 2k+1} (k = first..first+n-1) under key `s` */
static void prg_blocks(blk_t s, int first, int n, blk_t *out)
{
    uint8_t key[16], pt[16 * 8], ct[16 * 8];
    blk_store(key, s);
    for (int k = 0; k < n; k++)
    {
        blk_t p = {(uint64_t)(2 * (first + k)), (uint64_t)(2 * (first + k) + 1)};
        blk_store(pt + 16 * k, p);
    }
    oracle_aes128_encrypt_blocks(key, pt, ct, (size_t)n);
    for (int k = 0; k < n; k++) out[k] = blk_load(ct + 16 * k);
}

/* split child block into (seed with bit 0 cleared, t = bit 0) (Q: src/impl/dcf.h:431-434) */
static int take_t(blk_t *b)
{
    int t = (int)(b->lo & 1);
    b->lo &= LSB_CLEAR;
    return t;
}

/* ------------------------------------------------------------------------ */
/* value conversion (DcfConvertCtx src/impl/dcf.h:69-137; DpfConvertCtx      */
/* src/impl/dpf.h:63-128 has the same behaviour)                             */
/* ------------------------------------------------------------------------ */

typedef struct
{
    uint8_t key[16];
    uint8_t buf[16];
    int     bytes;  /* ceil(bitWidth/8) */
    int     count;  /* groupSize as passed */
    int     pos;    /* elements produced so far */
    int     direct; /* bytes*count <= 16: read the seed bytes directly */
} conv_t;

static void conv_init(conv_t *c, blk_t seed, size_t groupSize, size_t bitWidth)
{
    c->bytes  = (int)((bitWidth + 7) / 8);
    c->count  = (int)groupSize;
    c->direct = (c->bytes * c->count <= 16);
    c->pos    = 0;
    if (c->direct)
        blk_store(c->buf, seed);
    else
        blk_store(c->key, seed);
}

static uint64_t conv_next(conv_t *c)
{
    uint64_t r   = 0;
    int      off = c->pos * c->bytes;
    for (int i = 0; i < c->bytes; i++)
    {
        int idx = (off + i) % 16;
        if (!c->direct && idx == 0)
        {
            /* refill with AES_key(u64 0 || u64 element-index) */
            blk_t   p = {0, (uint64_t)c->pos};
            uint8_t pt[16];
            blk_store(pt, p);
            oracle_aes128_encrypt_blocks(c->key, pt, c->buf, 1);
        }
        if (i < 8) r |= (uint64_t)c->buf[idx] << (8 * i);
    }
    c->pos++;
    return r;
}

/* ------------------------------------------------------------------------ */
/* sizes                                                                     */
/* ------------------------------------------------------------------------ */

/* src/impl/dcf.h:38-50 */
static size_t dcf_key_size(size_t n, size_t g, size_t es, size_t N)
{
    return N * (16 * n + es * n * g + es + es + es * g);
}

/* src/impl/dcf.h:61-67 */
static size_t dcf_cache_size(size_t n, size_t g, size_t es, size_t N)
{
    return N * (16 * n) + N * es * n * g;
}

/* src/impl/dpf.h:35-47,57-61 */
static size_t dpf_key_size(size_t n, size_t g, size_t es, size_t N)
{
    return N * (16 * n + es + es + es * g);
}

static size_t dpf_cache_size(size_t n, size_t N)
{
    return N * 16 * n;
}

/* src/impl/grotto.h:37-44,81-85 */
static size_t grotto_key_size(size_t n, size_t es, size_t N)
{
    return N * (16 * (n - 6) + 8 + es + es);
}

static size_t grotto_cache_size(size_t n, size_t N)
{
    return N * (16 * (n - 6) + (n - 6));
}

/* src/impl/spline.h:27-48 */
static size_t spline_cache_size(size_t n, size_t intervals, size_t degree, size_t es, size_t N)
{
    size_t g = intervals * (degree + 1);
    return dcf_cache_size(n, g, es, N) + N * es * (2 * g + degree + 1);
}

/* src/dcf.cpp:36-124 */
int oracle_dcfGetKeyDataSize(size_t *out, size_t n, size_t m, size_t g, size_t es, size_t N)
{
    if (!(0 < n && n <= es * 8)) return E_BITWIDTH;
    if (!(0 < m && m <= es * 8)) return E_BITWIDTH;
    if (!es_ok(es)) return E_ELEM_SIZE;
    *out = dcf_key_size(n, g, es, N);
    return E_OK;
}

int oracle_dcfGetCacheDataSize(size_t *out, size_t n, size_t m, size_t g, size_t es, size_t N)
{
    if (!(0 < n && n <= es * 8)) return E_BITWIDTH;
    if (!(0 < m && m <= es * 8)) return E_BITWIDTH;
    if (!es_ok(es)) return E_ELEM_SIZE;
    *out = dcf_cache_size(n, g, es, N);
    return E_OK;
}

/* src/dpf.cpp:36-106: only the upper bound of the widths is checked */
int oracle_dpfGetKeyDataSize(size_t *out, size_t n, size_t m, size_t g, size_t es, size_t N)
{
    if (n > es * 8 || m > es * 8) return E_BITWIDTH;
    if (!es_ok(es)) return E_ELEM_SIZE;
    *out = dpf_key_size(n, g, es, N);
    return E_OK;
}

int oracle_dpfGetCacheDataSize(size_t *out, size_t n, size_t es, size_t N)
{
    if (n > es * 8) return E_BITWIDTH;
    if (!es_ok(es)) return E_ELEM_SIZE;
    *out = dpf_cache_size(n, N);
    return E_OK;
}

/* src/grotto.cpp:36-83: an unsupported element size reports size 0 and success */
int oracle_grottoGetKeyDataSize(size_t *out, size_t n, size_t es, size_t N)
{
    if (!(6 <= n && n <= es * 8)) return E_BITWIDTH;
    *out = es_ok(es) ? grotto_key_size(n, es, N) : 0;
    return E_OK;
}

int oracle_grottoGetCacheDataSize(size_t *out, size_t n, size_t es, size_t N)
{
    if (!(6 <= n && n <= es * 8)) return E_BITWIDTH;
    *out = es_ok(es) ? grotto_cache_size(n, N) : 0;
    return E_OK;
}

/* src/mic.cpp:36-77: no width checks; unsupported element size -> 0 */
int oracle_dcfMICGetKeyDataSize(size_t *out, size_t n, size_t m, size_t es, size_t N)
{
    *out = es_ok(es) ? dcf_key_size(n, 1, es, N) : 0;
    (void)m;
    return E_OK;
}

int oracle_dcfMICGetCacheDataSize(size_t *out, size_t n, size_t m, size_t es, size_t N)
{
    *out = es_ok(es) ? dcf_cache_size(n, 1, es, N) : 0;
    (void)m;
    return E_OK;
}

/* src/spline.cpp:36-81 */
int oracle_dcfSplineGetKeyDataSize(size_t *out, size_t d, size_t I, size_t n, size_t m, size_t es, size_t N)
{
    *out = es_ok(es) ? dcf_key_size(n, I * (d + 1), es, N) : 0;
    (void)m;
    return E_OK;
}

int oracle_dcfSplineGetCacheDataSize(size_t *out, size_t d, size_t I, size_t n, size_t m, size_t es, size_t N)
{
    *out = es_ok(es) ? spline_cache_size(n, I, d, es, N) : 0;
    (void)m;
    return E_OK;
}

/* src/ottt.cpp:10-20 */
int oracle_otttGetKeyDataSize(size_t *out, size_t n, size_t N)
{
    if (!(3 <= n)) return E_BITWIDTH;
    *out = N * (1ULL << n) / 8;
    return E_OK;
}

/* ------------------------------------------------------------------------ */
/* DCF                                                                       */
/* ------------------------------------------------------------------------ */

typedef struct
{
    const uint8_t *sCW;    /* n x 16 bytes */
    uint8_t       *sCWw;   /* writable alias (keyGen) */
    void          *tail;   /* vCW[n*g] | tLCW | tRCW | lastCW[g] */
    size_t         n, g, es;
} dcf_key_t;

/* src/impl/dcf.h:139-167: [sCW: N x n x 16][per element: vCW n*g | tLCW | tRCW | lastCW g] */
static dcf_key_t dcf_key_at(const void *key, size_t n, size_t g, size_t es, size_t i, size_t N)
{
    dcf_key_t k;
    k.sCW  = (const uint8_t *)key + i * 16 * n;
    k.sCWw = (uint8_t *)k.sCW;
    k.tail = (uint8_t *)key + N * 16 * n + i * (es * n * g + es + es + es * g);
    k.n = n, k.g = g, k.es = es;
    return k;
}

#define VCW(k, lvl, j) ((lvl) * (k)->g + (j))
#define TLCW(k) ((k)->n * (k)->g)
#define TRCW(k) ((k)->n * (k)->g + 1)
#define LASTCW(k, j) ((k)->n * (k)->g + 2 + (j))

/* src/impl/dcf.h:187-368 */
static void dcf_keygen_one(dcf_key_t *k, uint64_t alpha, const uint64_t *beta, const uint8_t *seed0,
                           const uint8_t *seed1, size_t m)
{
    const size_t n = k->n, g = k->g, es = k->es;
    const uint64_t M = ge_mask(es);
    uint64_t *vAlpha = (uint64_t *)calloc(g, sizeof(uint64_t));
    uint64_t *vcw    = (uint64_t *)calloc(g, sizeof(uint64_t));

    blk_t s0 = blk_load(seed0), s1 = blk_load(seed1);
    s0.lo &= LSB_CLEAR;
    s1.lo &= LSB_CLEAR;
    int t0 = 0, t1 = 1;
    uint64_t tlBits = 0, trBits = 0;

    for (size_t i = 0; i < n; i++)
    {
        blk_t c0[4], c1[4]; /* sL vL sR vR */
        prg_blocks(s0, 0, 4, c0);
        prg_blocks(s1, 0, 4, c1);
        int tL0 = take_t(&c0[0]), tR0 = take_t(&c0[2]);
        int tL1 = take_t(&c1[0]), tR1 = take_t(&c1[2]);

        int keep = (int)((alpha >> (n - 1 - i)) & 1); /* 1 = right child */
        int lose = 1 - keep;
        blk_t sLose0 = lose ? c0[2] : c0[0], sLose1 = lose ? c1[2] : c1[0];
        blk_t vLose0 = lose ? c0[3] : c0[1], vLose1 = lose ? c1[3] : c1[1];
        blk_t sKeep0 = keep ? c0[2] : c0[0], sKeep1 = keep ? c1[2] : c1[0];
        blk_t vKeep0 = keep ? c0[3] : c0[1], vKeep1 = keep ? c1[3] : c1[1];

        blk_t sCW = {sLose0.lo ^ sLose1.lo, sLose0.hi ^ sLose1.hi};
        uint64_t sign1 = t1 ? (uint64_t)-1 : 1;

        conv_t cv;
        conv_init(&cv, vLose1, g, m);
        for (size_t j = 0; j < g; j++) vcw[j] = conv_next(&cv);
        conv_init(&cv, vLose0, g, m);
        for (size_t j = 0; j < g; j++) vcw[j] -= conv_next(&cv);
        for (size_t j = 0; j < g; j++) vcw[j] = ((vcw[j] - vAlpha[j]) * sign1) & M;
        if (lose == 0)
            for (size_t j = 0; j < g; j++) vcw[j] = (vcw[j] + sign1 * beta[j]) & M;

        conv_init(&cv, vKeep1, g, m);
        for (size_t j = 0; j < g; j++) vAlpha[j] -= conv_next(&cv);
        conv_init(&cv, vKeep0, g, m);
        for (size_t j = 0; j < g; j++) vAlpha[j] += conv_next(&cv);
        for (size_t j = 0; j < g; j++) vAlpha[j] = (vAlpha[j] + sign1 * vcw[j]) & M;

        int tLCW = tL0 ^ tL1 ^ keep ^ 1;
        int tRCW = tR0 ^ tR1 ^ keep;
        blk_store(k->sCWw + 16 * i, sCW);
        for (size_t j = 0; j < g; j++) st(k->tail, es, VCW(k, i, j), vcw[j]);
        tlBits |= (uint64_t)tLCW << i;
        trBits |= (uint64_t)tRCW << i;

        if (t0) sKeep0.lo ^= sCW.lo, sKeep0.hi ^= sCW.hi;
        if (t1) sKeep1.lo ^= sCW.lo, sKeep1.hi ^= sCW.hi;
        s0 = sKeep0;
        s1 = sKeep1;
        int tKeepCW = keep ? tRCW : tLCW;
        int tk0 = keep ? tR0 : tL0, tk1 = keep ? tR1 : tL1;
        t0 = tk0 ^ (t0 & tKeepCW);
        t1 = tk1 ^ (t1 & tKeepCW);
    }
    st(k->tail, es, TLCW(k), tlBits);
    st(k->tail, es, TRCW(k), trBits);

    uint64_t sign1 = t1 ? (uint64_t)-1 : 1;
    conv_t cv;
    for (size_t j = 0; j < g; j++) vcw[j] = 0 - vAlpha[j];
    conv_init(&cv, s1, g, m);
    for (size_t j = 0; j < g; j++) vcw[j] += conv_next(&cv);
    conv_init(&cv, s0, g, m);
    for (size_t j = 0; j < g; j++) vcw[j] -= conv_next(&cv);
    for (size_t j = 0; j < g; j++) st(k->tail, es, LASTCW(k, j), (vcw[j] * sign1) & M);
    free(vAlpha);
    free(vcw);
}

/* src/impl/dcf.h:370-491 (without the prefix cache) */
static void dcf_eval_one(uint64_t *out, const dcf_key_t *k, uint64_t x, const uint8_t *seed, int party, size_t m)
{
    const size_t n = k->n, g = k->g, es = k->es;
    const uint64_t M    = ge_mask(es);
    const uint64_t sign = party ? (uint64_t)-1 : 1;
    blk_t s = blk_load(seed);
    s.lo &= LSB_CLEAR;
    int t = party;
    x     = mod_bits(x, n);
    for (size_t j = 0; j < g; j++) out[j] = 0;
    const uint64_t tl = ld(k->tail, es, TLCW(k)), tr = ld(k->tail, es, TRCW(k));

    for (size_t i = 0; i < n; i++)
    {
        blk_t c[4];
        prg_blocks(s, 0, 4, c);
        int tL = take_t(&c[0]), tR = take_t(&c[2]);
        if (t)
        {
            blk_t cw = blk_load(k->sCW + 16 * i);
            c[0].lo ^= cw.lo, c[0].hi ^= cw.hi;
            c[2].lo ^= cw.lo, c[2].hi ^= cw.hi;
            tL ^= (int)((tl >> i) & 1);
            tR ^= (int)((tr >> i) & 1);
        }
        int   xi = (int)((x >> (n - 1 - i)) & 1);
        blk_t v  = xi ? c[3] : c[1];
        conv_t cv;
        conv_init(&cv, v, g, m);
        for (size_t j = 0; j < g; j++)
        {
            uint64_t term = conv_next(&cv) + (t ? ld(k->tail, es, VCW(k, i, j)) : 0);
            out[j]        = (out[j] + sign * term) & M;
        }
        s = xi ? c[2] : c[0];
        t = xi ? tR : tL;
    }
    conv_t cv;
    conv_init(&cv, s, g, m);
    for (size_t j = 0; j < g; j++)
    {
        uint64_t term = conv_next(&cv) + (t ? ld(k->tail, es, LASTCW(k, j)) : 0);
        out[j]        = (out[j] + sign * term) & M;
    }
}

/* src/kernel/dcf.h:31-69 */
int oracle_dcfKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize, const void *beta,
                     size_t betaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1,
                     size_t seedDataSize1, size_t n, size_t m, size_t g, size_t es, size_t N)
{
    if (!es_ok(es)) return E_ELEM_SIZE;
    if (!(0 < n && n <= es * 8)) return E_BITWIDTH;
    if (!(0 < m && m <= es * 8)) return E_BITWIDTH;
    if (keyDataSize != dcf_key_size(n, g, es, N)) return E_KEY_SIZE;
    if (g == 0) return E_GROUP_SIZE;
    if (alphaDataSize != N * es) return E_ALPHA_SIZE;
    if (betaDataSize != 0 && betaDataSize != N * es * g) return E_BETA_SIZE;
    if (seedDataSize0 != N * 16 || seedDataSize1 != N * 16) return E_SEED_SIZE;

#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t    i = (size_t)ii;
        uint64_t *b = (uint64_t *)malloc(g * sizeof(uint64_t));
        for (size_t j = 0; j < g; j++)
            /* beta == NULL means beta = 1 (src/kernel/dcf.h:75-88) */
            b[j] = beta ? ld(beta, es, g * i + j) : (j == 0 ? 1 : 0);
        dcf_key_t k = dcf_key_at(key, n, g, es, i, N);
        dcf_keygen_one(&k, ld(alpha, es, i), b, (const uint8_t *)seed0 + 16 * i, (const uint8_t *)seed1 + 16 * i,
                       m);
        free(b);
    }
    return E_OK;
}

/* src/kernel/dcf.h:115-163 */
int oracle_dcfEval(void *out, size_t outSize, const void *x, size_t xSize, const void *key, size_t keyDataSize,
                   const void *seed, size_t seedSize, int party, size_t n, size_t m, size_t g, size_t es, size_t N,
                   void *cache, size_t cacheSize)
{
    (void)cache;
    if (!es_ok(es)) return E_ELEM_SIZE;
    if (!(0 < n && n <= es * 8)) return E_BITWIDTH;
    if (!(0 < m && m <= es * 8)) return E_BITWIDTH;
    if (party != 0 && party != 1) return E_PARTY;
    if (g == 0) return E_GROUP_SIZE;
    if (keyDataSize != dcf_key_size(n, g, es, N)) return E_KEY_SIZE;
    if (outSize != N * es * g) return E_OUT_SIZE;
    if (xSize != N * es) return E_X_SIZE;
    if (seedSize != N * 16) return E_SEED_SIZE;
    if (cacheSize != 0 && cacheSize != dcf_cache_size(n, g, es, N)) return E_CACHE_SIZE;

#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t    i = (size_t)ii;
        uint64_t  small[8];
        uint64_t *o = g <= 8 ? small : (uint64_t *)malloc(g * sizeof(uint64_t));
        dcf_key_t k = dcf_key_at(key, n, g, es, i, N);
        dcf_eval_one(o, &k, ld(x, es, i), (const uint8_t *)seed + 16 * i, party, m);
        for (size_t j = 0; j < g; j++) st(out, es, g * i + j, o[j]);
        if (o != small) free(o);
    }
    return E_OK;
}

/* ------------------------------------------------------------------------ */
/* DPF                                                                       */
/* ------------------------------------------------------------------------ */

typedef struct
{
    const uint8_t *sCW;
    uint8_t       *sCWw;
    void          *tail; /* tLCW | tRCW | lastCW[g] */
    size_t         n, g, es;
} dpf_key_t;

/* src/impl/dpf.h:130-156 */
static dpf_key_t dpf_key_at(const void *key, size_t n, size_t g, size_t es, size_t i, size_t N)
{
    dpf_key_t k;
    k.sCW  = (const uint8_t *)key + i * 16 * n;
    k.sCWw = (uint8_t *)k.sCW;
    k.tail = (uint8_t *)key + N * 16 * n + i * (es + es + es * g);
    k.n = n, k.g = g, k.es = es;
    return k;
}

/*
 * Descend one level of a 2-block (DPF/Grotto) tree for both parties, writing
 * the level's correction words.  Shared by dpfKeyGen (src/impl/dpf.h:201-272)
 * and grottoKeyGen (src/impl/grotto.h:137-201), whose per-level updates are
 * the same computation.
 */
static void two_block_keygen_level(blk_t *s0, blk_t *s1, int *t0, int *t1, int keep, blk_t *sCWout, int *tLCWout,
                                   int *tRCWout)
{
    blk_t c0[2], c1[2];
    prg_blocks(*s0, 0, 2, c0);
    prg_blocks(*s1, 0, 2, c1);
    int tL0 = take_t(&c0[0]), tR0 = take_t(&c0[1]);
    int tL1 = take_t(&c1[0]), tR1 = take_t(&c1[1]);
    int lose    = 1 - keep;
    blk_t sLose0 = lose ? c0[1] : c0[0], sLose1 = lose ? c1[1] : c1[0];
    blk_t sKeep0 = keep ? c0[1] : c0[0], sKeep1 = keep ? c1[1] : c1[0];
    blk_t sCW    = {sLose0.lo ^ sLose1.lo, sLose0.hi ^ sLose1.hi};
    int tLCW = tL0 ^ tL1 ^ keep ^ 1;
    int tRCW = tR0 ^ tR1 ^ keep;
    if (*t0) sKeep0.lo ^= sCW.lo, sKeep0.hi ^= sCW.hi;
    if (*t1) sKeep1.lo ^= sCW.lo, sKeep1.hi ^= sCW.hi;
    int tKeepCW = keep ? tRCW : tLCW;
    int tk0 = keep ? tR0 : tL0, tk1 = keep ? tR1 : tL1;
    *t0 = tk0 ^ (*t0 & tKeepCW);
    *t1 = tk1 ^ (*t1 & tKeepCW);
    *s0 = sKeep0;
    *s1 = sKeep1;
    *sCWout  = sCW;
    *tLCWout = tLCW;
    *tRCWout = tRCW;
}

/* src/impl/dpf.h:172-288 */
static void dpf_keygen_one(dpf_key_t *k, uint64_t alpha, const uint64_t *beta, const uint8_t *seed0,
                           const uint8_t *seed1, size_t m)
{
    const size_t n = k->n, g = k->g, es = k->es;
    const uint64_t M = ge_mask(es);
    blk_t s0 = blk_load(seed0), s1 = blk_load(seed1);
    s0.lo &= LSB_CLEAR;
    s1.lo &= LSB_CLEAR;
    int t0 = 0, t1 = 1;
    uint64_t tl = 0, tr = 0;
    for (size_t i = 0; i < n; i++)
    {
        blk_t sCW;
        int   tLCW, tRCW;
        int   keep = (int)((alpha >> (n - 1 - i)) & 1);
        two_block_keygen_level(&s0, &s1, &t0, &t1, keep, &sCW, &tLCW, &tRCW);
        blk_store(k->sCWw + 16 * i, sCW);
        tl |= (uint64_t)tLCW << i;
        tr |= (uint64_t)tRCW << i;
    }
    st(k->tail, es, 0, tl);
    st(k->tail, es, 1, tr);
    /* the reference passes (bitWidthOut, groupSize) into a (groupSize, bitWidth)
     * parameter list (src/impl/dpf.h:74,278,283); reproduced deliberately */
    uint64_t sign1 = t1 ? (uint64_t)-1 : 1;
    conv_t   c0, c1;
    conv_init(&c0, s0, m, g);
    conv_init(&c1, s1, m, g);
    for (size_t j = 0; j < g; j++)
    {
        uint64_t a = conv_next(&c0);
        uint64_t b = conv_next(&c1);
        st(k->tail, es, 2 + j, (sign1 * (beta[j] - a + b)) & M);
    }
}

/* src/impl/dpf.h:290-380 (without the prefix cache) */
static void dpf_eval_one(uint64_t *out, const dpf_key_t *k, uint64_t x, const uint8_t *seed, int party, size_t m)
{
    const size_t n = k->n, g = k->g, es = k->es;
    const uint64_t M = ge_mask(es);
    blk_t s = blk_load(seed);
    s.lo &= LSB_CLEAR;
    int t = party;
    x     = mod_bits(x, n);
    const uint64_t tl = ld(k->tail, es, 0), tr = ld(k->tail, es, 1);
    for (size_t i = 0; i < n; i++)
    {
        blk_t c[2];
        prg_blocks(s, 0, 2, c);
        int tL = take_t(&c[0]), tR = take_t(&c[1]);
        if (t)
        {
            blk_t cw = blk_load(k->sCW + 16 * i);
            c[0].lo ^= cw.lo, c[0].hi ^= cw.hi;
            c[1].lo ^= cw.lo, c[1].hi ^= cw.hi;
            tL ^= (int)((tl >> i) & 1);
            tR ^= (int)((tr >> i) & 1);
        }
        int xi = (int)((x >> (n - 1 - i)) & 1);
        s      = xi ? c[1] : c[0];
        t      = xi ? tR : tL;
    }
    const uint64_t sign = party ? (uint64_t)-1 : 1;
    conv_t cv;
    conv_init(&cv, s, m, g); /* argument swap, as in keyGen */
    for (size_t j = 0; j < g; j++)
    {
        uint64_t term = conv_next(&cv) + (t ? ld(k->tail, es, 2 + j) : 0);
        out[j]        = (sign * term) & M;
    }
}

/* src/kernel/dpf.h:33-71 */
int oracle_dpfKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize, const void *beta,
                     size_t betaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1,
                     size_t seedDataSize1, size_t n, size_t m, size_t g, size_t es, size_t N)
{
    if (!es_ok(es)) return E_ELEM_SIZE;
    if (!(0 < n && n <= es * 8)) return E_BITWIDTH;
    if (!(0 < m && m <= es * 8)) return E_BITWIDTH;
    if (g == 0) return E_GROUP_SIZE;
    if (keyDataSize != dpf_key_size(n, g, es, N)) return E_KEY_SIZE;
    if (alphaDataSize != N * es) return E_ALPHA_SIZE;
    if (betaDataSize != 0)
    {
        if (betaDataSize != N * es * g) return E_BETA_SIZE;
    }
    else if (g != 1)
        return E_BETA_SIZE;
    if (seedDataSize0 != N * 16 || seedDataSize1 != N * 16) return E_SEED_SIZE;

#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t    i = (size_t)ii;
        uint64_t *b = (uint64_t *)malloc(g * sizeof(uint64_t));
        for (size_t j = 0; j < g; j++) b[j] = beta ? ld(beta, es, g * i + j) : 1;
        dpf_key_t k = dpf_key_at(key, n, g, es, i, N);
        dpf_keygen_one(&k, ld(alpha, es, i), b, (const uint8_t *)seed0 + 16 * i, (const uint8_t *)seed1 + 16 * i,
                       m);
        free(b);
    }
    return E_OK;
}

/* common leading checks of the DPF eval tasks (src/kernel/dpf.h:119-133) */
static int dpf_eval_check(int party, size_t n, size_t m, size_t g, size_t es)
{
    if (!es_ok(es)) return E_ELEM_SIZE;
    if (!(0 < n && n <= es * 8)) return E_BITWIDTH;
    if (!(0 < m && m <= es * 8)) return E_BITWIDTH;
    if (g == 0) return E_GROUP_SIZE;
    if (party != 0 && party != 1) return E_PARTY;
    return E_OK;
}

/* src/kernel/dpf.h:115-154 */
int oracle_dpfEval(void *out, size_t outSize, const void *x, size_t xSize, const void *key, size_t keyDataSize,
                   const void *seed, size_t seedSize, int party, size_t n, size_t m, size_t g, size_t es, size_t N,
                   void *cache, size_t cacheSize)
{
    (void)cache;
    int r = dpf_eval_check(party, n, m, g, es);
    if (r) return r;
    if (keyDataSize != dpf_key_size(n, g, es, N)) return E_KEY_SIZE;
    if (seedSize != N * 16) return E_SEED_SIZE;
    if (xSize != N * es) return E_X_SIZE;
    if (outSize != N * es * g) return E_OUT_SIZE;
    if (cacheSize != 0 && cacheSize != dpf_cache_size(n, N)) return E_CACHE_SIZE;

#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t    i = (size_t)ii;
        uint64_t  small[8];
        uint64_t *o = g <= 8 ? small : (uint64_t *)malloc(g * sizeof(uint64_t));
        dpf_key_t k = dpf_key_at(key, n, g, es, i, N);
        dpf_eval_one(o, &k, ld(x, es, i), (const uint8_t *)seed + 16 * i, party, m);
        for (size_t j = 0; j < g; j++) st(out, es, g * i + j, o[j]);
        if (o != small) free(o);
    }
    return E_OK;
}

/* src/kernel/dpf.h:201-276 */
int oracle_dpfEvalAll(void *out, size_t outSize, const void *x, size_t xSize, const void *key, size_t keyDataSize,
                      const void *seed, size_t seedSize, int party, size_t n, size_t m, size_t g, size_t es,
                      size_t N, void *cache, size_t cacheSize)
{
    (void)cache;
    int r = dpf_eval_check(party, n, m, g, es);
    if (r) return r;
    if (keyDataSize != dpf_key_size(n, g, es, N)) return E_KEY_SIZE;
    if (n >= 64 || outSize != N * es * (1ULL << n) * g) return E_OUT_SIZE;
    if (xSize != N * es) return E_X_SIZE;
    if (seedSize != N * 16) return E_SEED_SIZE;
    if (cacheSize != 0 && cacheSize != dpf_cache_size(n, N)) return E_CACHE_SIZE;

    const size_t size = (size_t)1 << n;
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t    i = (size_t)ii;
        uint64_t *o = (uint64_t *)malloc(g * sizeof(uint64_t));
        dpf_key_t k = dpf_key_at(key, n, g, es, i, N);
        uint64_t xi = ld(x, es, i);
        for (size_t j = 0; j < size; j++)
        {
            size_t idx = (size_t)(xi - j) % size; /* out[(x - j) mod 2^n] = Eval(j) */
            dpf_eval_one(o, &k, j, (const uint8_t *)seed + 16 * i, party, m);
            for (size_t q = 0; q < g; q++) st(out, es, size * i * g + g * idx + q, o[q]);
        }
        free(o);
    }
    return E_OK;
}

/* src/kernel/dpf.h:383-534 */
int oracle_dpfEvalMulti(void *out, size_t outSize, const void *x, size_t xSize, const void *key,
                        size_t keyDataSize, const void *seed, size_t seedSize, int party, const void *point,
                        size_t pointSize, size_t n, size_t m, size_t g, size_t es, size_t N, void *cache,
                        size_t cacheSize)
{
    (void)cache;
    int r = dpf_eval_check(party, n, m, g, es);
    if (r) return r;
    if (pointSize % es != 0) return E_POINT_SIZE;
    if (keyDataSize != dpf_key_size(n, g, es, N)) return E_KEY_SIZE;
    const size_t P = pointSize / es;
    if (outSize != N * es * P * g) return E_OUT_SIZE;
    if (xSize != N * es) return E_X_SIZE;
    if (seedSize != N * 16) return E_SEED_SIZE;
    if (cacheSize != 0 && cacheSize != dpf_cache_size(n, N)) return E_CACHE_SIZE;

    const uint64_t M = ge_mask(es);
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t    i = (size_t)ii;
        uint64_t *o = (uint64_t *)malloc(g * sizeof(uint64_t));
        dpf_key_t k = dpf_key_at(key, n, g, es, i, N);
        for (size_t j = 0; j < P; j++)
        {
            uint64_t xj = (ld(x, es, i) - ld(point, es, j)) & M;
            dpf_eval_one(o, &k, xj, (const uint8_t *)seed + 16 * i, party, m);
            for (size_t q = 0; q < g; q++) st(out, es, P * i * g + j * g + q, o[q]);
        }
        free(o);
    }
    return E_OK;
}

/* ------------------------------------------------------------------------ */
/* Grotto                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct
{
    const uint8_t *sCW;    /* (n-6) x 16 */
    uint8_t       *sCWw;
    uint8_t       *lastCW; /* 8 bytes */
    void          *tlr;    /* tLCW, tRCW */
    size_t         n, es;
} grotto_key_t;

/* src/impl/grotto.h:53-79: [sCW N x (n-6) x 16][lastCW N x 8][tLCW,tRCW N x 2 GE] */
static grotto_key_t grotto_key_at(const void *key, size_t n, size_t es, size_t i, size_t N)
{
    grotto_key_t k;
    k.sCW    = (const uint8_t *)key + i * 16 * (n - 6);
    k.sCWw   = (uint8_t *)k.sCW;
    k.lastCW = (uint8_t *)key + N * 16 * (n - 6) + 8 * i;
    k.tlr    = (uint8_t *)key + N * 16 * (n - 6) + N * 8 + i * es * 2;
    k.n = n, k.es = es;
    return k;
}

/* src/impl/grotto.h:111-205 */
static void grotto_keygen_one(grotto_key_t *k, uint64_t alpha, const uint8_t *seed0, const uint8_t *seed1)
{
    const size_t n = k->n;
    blk_t s0 = blk_load(seed0), s1 = blk_load(seed1);
    s0.lo &= LSB_CLEAR;
    s1.lo &= LSB_CLEAR;
    int t0 = 0, t1 = 1;
    uint64_t tl = 0, tr = 0;
    for (size_t i = 0; i < n - 6; i++)
    {
        blk_t sCW;
        int   tLCW, tRCW;
        int   keep = (int)((alpha >> (n - 1 - i)) & 1);
        two_block_keygen_level(&s0, &s1, &t0, &t1, keep, &sCW, &tLCW, &tRCW);
        blk_store(k->sCWw + 16 * i, sCW);
        tl |= (uint64_t)tLCW << i;
        tr |= (uint64_t)tRCW << i;
    }
    st(k->tlr, k->es, 0, tl);
    st(k->tlr, k->es, 1, tr);
    uint64_t leaf = s0.hi ^ s1.hi ^ (1ULL << (alpha & 63));
    memcpy(k->lastCW, &leaf, 8);
}

/* one Grotto level on the path: a single AES block chosen by the input bit
 * (src/impl/grotto.h:336-362) */
static void grotto_step(const grotto_key_t *k, size_t i, int bit, blk_t *s, int *t, uint64_t tl, uint64_t tr)
{
    blk_t c;
    prg_blocks(*s, bit, 1, &c);
    int ti = take_t(&c);
    if (*t)
    {
        blk_t cw = blk_load(k->sCW + 16 * i);
        c.lo ^= cw.lo, c.hi ^= cw.hi;
        ti ^= (int)((((bit ? tr : tl)) >> i) & 1);
    }
    *s = c;
    *t = ti;
}

/* src/impl/grotto.h:207-282 */
static uint64_t grotto_eq_eval_one(const grotto_key_t *k, uint64_t x, const uint8_t *seed, int party)
{
    const size_t n = k->n;
    blk_t s = blk_load(seed);
    s.lo &= LSB_CLEAR;
    int t = party;
    x     = mod_bits(x, n);
    const uint64_t tl = ld(k->tlr, k->es, 0), tr = ld(k->tlr, k->es, 1);
    for (size_t i = 0; i < n - 6; i++) grotto_step(k, i, (int)((x >> (n - 1 - i)) & 1), &s, &t, tl, tr);
    uint64_t lcw;
    memcpy(&lcw, k->lastCW, 8);
    uint64_t u = s.hi ^ (t ? lcw : 0);
    return (u >> (x & 63)) & 1;
}

/* src/impl/grotto.h:284-391 */
static uint64_t grotto_eval_one(const grotto_key_t *k, uint64_t x, const uint8_t *seed, int party, int equalBound)
{
    const size_t n = k->n;
    x = mod_bits(x, n);
    blk_t s = blk_load(seed);
    s.lo &= LSB_CLEAR;
    int t = party, parity = 0;
    const uint64_t tl = ld(k->tlr, k->es, 0), tr = ld(k->tlr, k->es, 1);
    for (size_t i = 0; i < n - 6; i++)
    {
        /* early exit when the remaining low n-i bits are all zero */
        if (equalBound && (x & mod_bits(~0ULL, n - i)) == 0) return (uint64_t)(parity ^ party);
        int bit  = (int)((x >> (n - 1 - i)) & 1);
        int prev = t;
        grotto_step(k, i, bit, &s, &t, tl, tr);
        if (bit) parity ^= prev ^ t;
    }
    uint64_t lcw;
    memcpy(&lcw, k->lastCW, 8);
    uint64_t u = t ? (s.hi ^ lcw) : s.hi;
    unsigned p = (unsigned)(x & 63);
    if (equalBound)
    {
        if (p != 0) parity ^= parity64(u & (~0ULL >> (64 - p)));
    }
    else
        parity ^= parity64(u & (~0ULL >> (63 - p)));
    return (uint64_t)(parity ^ party);
}

static int grotto_common_check(size_t n, size_t es, size_t N, size_t xSize, size_t seedSize, int party,
                               size_t keyDataSize, size_t cacheSize)
{
    if (party != 0 && party != 1) return E_PARTY;
    if (keyDataSize != grotto_key_size(n, es, N)) return E_KEY_SIZE;
    if (cacheSize != 0 && cacheSize != grotto_cache_size(n, N)) return E_CACHE_SIZE;
    (void)xSize, (void)seedSize;
    return E_OK;
}

/* src/kernel/grotto.h:26-46 */
int oracle_grottoKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize, const void *seed0,
                        size_t seedDataSize0, const void *seed1, size_t seedDataSize1, size_t n, size_t es, size_t N)
{
    if (!es_ok(es)) return E_ELEM_SIZE;
    if (!(6 <= n && n <= es * 8)) return E_BITWIDTH;
    if (alphaDataSize != N * es) return E_ALPHA_SIZE;
    if (seedDataSize0 != N * 16 || seedDataSize1 != N * 16) return E_SEED_SIZE;
    if (keyDataSize != grotto_key_size(n, es, N)) return E_KEY_SIZE;
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t       i = (size_t)ii;
        grotto_key_t k = grotto_key_at(key, n, es, i, N);
        grotto_keygen_one(&k, ld(alpha, es, i), (const uint8_t *)seed0 + 16 * i, (const uint8_t *)seed1 + 16 * i);
    }
    return E_OK;
}

/* src/kernel/grotto.h:77-112 */
int oracle_grottoEqEval(void *out, size_t outSize, const void *x, size_t xSize, const void *key, size_t keyDataSize,
                        const void *seed, size_t seedSize, int party, size_t n, size_t es, size_t N, void *cache,
                        size_t cacheSize)
{
    (void)cache;
    if (!es_ok(es)) return E_ELEM_SIZE;
    if (!(6 <= n && n <= es * 8)) return E_BITWIDTH;
    if (outSize != N * es) return E_OUT_SIZE;
    if (xSize != N * es) return E_X_SIZE;
    if (seedSize != N * 16) return E_SEED_SIZE;
    int r = grotto_common_check(n, es, N, xSize, seedSize, party, keyDataSize, cacheSize);
    if (r) return r;
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t       i = (size_t)ii;
        grotto_key_t k = grotto_key_at(key, n, es, i, N);
        st(out, es, i, grotto_eq_eval_one(&k, ld(x, es, i), (const uint8_t *)seed + 16 * i, party));
    }
    return E_OK;
}

/* src/kernel/grotto.h:153-188 */
int oracle_grottoEval(void *out, size_t outSize, const void *x, size_t xSize, const void *key, size_t keyDataSize,
                      const void *seed, size_t seedSize, bool equalBound, int party, size_t n, size_t es, size_t N,
                      void *cache, size_t cacheSize)
{
    (void)cache;
    if (!es_ok(es)) return E_ELEM_SIZE;
    if (!(6 <= n && n <= es * 8)) return E_BITWIDTH;
    if (outSize != N * es) return E_OUT_SIZE;
    if (xSize != N * es) return E_X_SIZE;
    if (seedSize != N * 16) return E_SEED_SIZE;
    int r = grotto_common_check(n, es, N, xSize, seedSize, party, keyDataSize, cacheSize);
    if (r) return r;
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t       i = (size_t)ii;
        grotto_key_t k = grotto_key_at(key, n, es, i, N);
        st(out, es, i, grotto_eval_one(&k, ld(x, es, i), (const uint8_t *)seed + 16 * i, party, equalBound ? 1 : 0));
    }
    return E_OK;
}

/* interval share with reuse of the previous right endpoint (src/impl/grotto.h:393-447) */
static void grotto_mic_one(uint64_t *share, const grotto_key_t *k, uint64_t x, const uint8_t *seed, int party,
                           const void *left, const void *right, size_t I, size_t es)
{
    const size_t n = k->n;
    uint64_t     sp = 0, sq = 0;
    const uint64_t M = ge_mask(es);
    for (size_t i = 0; i < I; i++)
    {
        uint64_t L  = ld(left, es, i), R = ld(right, es, i);
        uint64_t xP = mod_bits((x - (L - 1)) & M, n);
        uint64_t xQ = mod_bits((x - R) & M, n);
        if (i > 0 && L == mod_bits((ld(right, es, i - 1) + 1) & M, n))
            sp = sq;
        else
            sp = grotto_eval_one(k, xP, seed, party, 1);
        sq       = grotto_eval_one(k, xQ, seed, party, 1);
        share[i] = (sp ^ sq) ^ ((xQ >= xP) ? (uint64_t)party : 0);
    }
}

/* src/kernel/grotto.h:234-297 */
int oracle_grottoMICEval(void *out, size_t outSize, const void *x, size_t xSize, const void *key,
                         size_t keyDataSize, const void *seed, size_t seedSize, int party, const void *left,
                         size_t leftSize, const void *right, size_t rightSize, size_t n, size_t es, size_t N,
                         void *cache, size_t cacheSize)
{
    (void)cache;
    if (!es_ok(es)) return E_ELEM_SIZE;
    if (!(6 <= n && n <= es * 8)) return E_BITWIDTH;
    size_t I = leftSize / es;
    if (I * es != leftSize || I * es != rightSize) return E_BOUNDARY_SIZE;
    if (outSize != N * es * I) return E_OUT_SIZE;
    if (xSize != N * es) return E_X_SIZE;
    if (seedSize != N * 16) return E_SEED_SIZE;
    int r = grotto_common_check(n, es, N, xSize, seedSize, party, keyDataSize, cacheSize);
    if (r) return r;
    if (I == 0) return E_OK;
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t       i  = (size_t)ii;
        uint64_t    *sh = (uint64_t *)malloc(I * sizeof(uint64_t));
        grotto_key_t k  = grotto_key_at(key, n, es, i, N);
        grotto_mic_one(sh, &k, ld(x, es, i), (const uint8_t *)seed + 16 * i, party, left, right, I, es);
        for (size_t q = 0; q < I; q++) st(out, es, I * i + q, sh[q]);
        free(sh);
    }
    return E_OK;
}

/* src/impl/grotto.h:449-545; src/kernel/grotto.h:327-379 */
int oracle_grottoIntervalLutEval(void *outE, size_t outESize, void *outT, size_t outTSize, const void *x,
                                 size_t xSize, const void *key, size_t keyDataSize, const void *seed,
                                 size_t seedSize, int party, const void *left, size_t leftSize, const void *right,
                                 size_t rightSize, const void *lut, size_t lutSize, size_t n, size_t m, size_t es,
                                 size_t N, void *cache, size_t cacheSize)
{
    (void)cache, (void)m;
    if (!es_ok(es)) return E_ELEM_SIZE;
    if (!(6 <= n && n <= es * 8)) return E_BITWIDTH;
    if (!(leftSize == rightSize && leftSize % es == 0)) return E_BOUNDARY_SIZE;
    size_t I = leftSize / es;
    if (I == 0 || lutSize % (es * I) != 0) return E_LUT_SIZE;
    size_t lutNum = lutSize / (es * I);
    if (outESize != N * es) return E_OUT_SIZE;
    if (outTSize != N * es * lutNum) return E_OUT_SIZE;
    if (xSize != N * es) return E_X_SIZE;
    if (seedSize != N * 16) return E_SEED_SIZE;
    int r = grotto_common_check(n, es, N, xSize, seedSize, party, keyDataSize, cacheSize);
    if (r) return r;
    const uint64_t M = ge_mask(es);
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t       i  = (size_t)ii;
        uint64_t    *sh = (uint64_t *)malloc(I * sizeof(uint64_t));
        grotto_key_t k  = grotto_key_at(key, n, es, i, N);
        grotto_mic_one(sh, &k, ld(x, es, i), (const uint8_t *)seed + 16 * i, party, left, right, I, es);
        uint64_t e = 0;
        for (size_t q = 0; q < I; q++) e += sh[q] & 1;
        for (size_t j = 0; j < lutNum; j++)
        {
            uint64_t tsum = 0;
            for (size_t q = 0; q < I; q++) tsum += (sh[q] & 1) * ld(lut, es, q + j * I);
            if (party) tsum = 0 - tsum;
            st(outT, es, i * lutNum + j, tsum & M);
        }
        if (party) e = 0 - e;
        st(outE, es, i, e & M);
        free(sh);
    }
    return E_OK;
}

/* ------------------------------------------------------------------------ */
/* DCF-based MIC                                                             */
/* ------------------------------------------------------------------------ */

/* src/kernel/mic.h:33-66 */
int oracle_dcfMICKeyGen(void *key, size_t keyDataSize, void *z, size_t zSize, const void *alpha, size_t alphaSize,
                        const void *seed0, size_t seedSize0, const void *seed1, size_t seedSize1, const void *left,
                        size_t leftSize, const void *right, size_t rightSize, size_t n, size_t m, size_t es,
                        size_t N)
{
    if (!es_ok(es)) return E_ELEM_SIZE;
    const size_t I = leftSize / es;
    if (!(0 < n && n <= es * 8)) return E_BITWIDTH;
    if (!(0 < m && m <= es * 8)) return E_BITWIDTH;
    if (keyDataSize != dcf_key_size(n, 1, es, N)) return E_KEY_SIZE;
    if (zSize != N * es * I) return E_Z_SIZE;
    if (alphaSize != N * es) return E_ALPHA_SIZE;
    if (seedSize0 != N * 16 || seedSize1 != N * 16) return E_SEED_SIZE;
    if (leftSize != I * es || rightSize != I * es) return E_BOUNDARY_SIZE;
    const uint64_t M   = ge_mask(es);
    const uint64_t MAX = mod_bits(M, n);
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t    i  = (size_t)ii;
        uint64_t  a  = ld(alpha, es, i);
        uint64_t  one = 1;
        dcf_key_t k  = dcf_key_at(key, n, 1, es, i, N);
        /* DCF on alpha-1 with beta = 1 (src/impl/mic.h:22) */
        dcf_keygen_one(&k, (a - 1) & M, &one, (const uint8_t *)seed0 + 16 * i, (const uint8_t *)seed1 + 16 * i, m);
        for (size_t q = 0; q < I; q++)
        {
            uint64_t L = ld(left, es, q), R = ld(right, es, q);
            uint64_t qPrime  = mod_bits((R + 1) & M, n);
            uint64_t alphaP  = mod_bits((L + a) & M, n);
            uint64_t alphaQ  = mod_bits((R + a) & M, n);
            uint64_t alphaQP = mod_bits((R + 1 + a) & M, n);
            uint64_t p       = mod_bits(L, n);
            int64_t  zv = (int64_t)(alphaP > alphaQ) - (int64_t)(alphaP > p) + (int64_t)(alphaQP > qPrime) +
                         (int64_t)(alphaQ == MAX);
            st(z, es, I * i + q, (uint64_t)zv & M);
        }
    }
    return E_OK;
}

/* src/impl/mic.h:43-109 */
static void dcf_mic_eval_one(uint64_t *out, uint64_t x, const dcf_key_t *k, const void *zs, size_t zoff,
                             const uint8_t *seed, int party, const void *left, const void *right, size_t I, size_t n,
                             size_t m, size_t es)
{
    const uint64_t M = ge_mask(es);
    x                = mod_bits(x, n);
    uint64_t sp = 0, sq = 0;
    for (size_t i = 0; i < I; i++)
    {
        uint64_t L = ld(left, es, i), R = ld(right, es, i);
        uint64_t qPrime  = mod_bits((R + 1) & M, n);
        uint64_t xP      = mod_bits((x - 1 - L) & M, n);
        uint64_t xQPrime = mod_bits((x - 1 - ((R + 1) & M)) & M, n);
        uint64_t p       = mod_bits(L, n);
        if (i > 0 && L == mod_bits((ld(right, es, i - 1) + 1) & M, n))
            sp = sq;
        else
            dcf_eval_one(&sp, k, xP, seed, party, m);
        dcf_eval_one(&sq, k, xQPrime, seed, party, m);
        uint64_t o = (sq - sp + ld(zs, es, zoff + i)) & M;
        if (party == 1) o = (o + (uint64_t)(int64_t)((int)(x > p) - (int)(x > qPrime))) & M;
        out[i] = o;
    }
}

/* src/kernel/mic.h:112-159 */
int oracle_dcfMICEval(void *out, size_t outSize, const void *x, size_t xSize, const void *key, size_t keyDataSize,
                      const void *z, size_t zSize, const void *seed, size_t seedSize, int party, const void *left,
                      size_t leftSize, const void *right, size_t rightSize, size_t n, size_t m, size_t es, size_t N,
                      void *cache, size_t cacheSize)
{
    (void)cache;
    if (!es_ok(es)) return E_ELEM_SIZE;
    const size_t I = leftSize / es;
    if (!(0 < n && n <= es * 8)) return E_BITWIDTH;
    if (!(0 < m && m <= es * 8)) return E_BITWIDTH;
    if (party != 0 && party != 1) return E_PARTY;
    if (keyDataSize != dcf_key_size(n, 1, es, N)) return E_KEY_SIZE;
    if (outSize != N * es * I) return E_OUT_SIZE;
    if (xSize != N * es) return E_X_SIZE;
    if (zSize != N * es * I) return E_Z_SIZE;
    if (seedSize != N * 16) return E_SEED_SIZE;
    if (leftSize != I * es || rightSize != I * es) return E_BOUNDARY_SIZE;
    if (cacheSize != 0 && cacheSize != dcf_cache_size(n, 1, es, N)) return E_CACHE_SIZE;
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t    i  = (size_t)ii;
        uint64_t *o  = (uint64_t *)malloc((I ? I : 1) * sizeof(uint64_t));
        dcf_key_t k  = dcf_key_at(key, n, 1, es, i, N);
        dcf_mic_eval_one(o, ld(x, es, i), &k, z, I * i, (const uint8_t *)seed + 16 * i, party, left, right, I, n, m,
                         es);
        for (size_t q = 0; q < I; q++) st(out, es, I * i + q, o[q]);
        free(o);
    }
    return E_OK;
}

/* ------------------------------------------------------------------------ */
/* DCF-based spline                                                          */
/* ------------------------------------------------------------------------ */

/* coefficients of f(x - shift), reduced mod 2^m (src/impl/spline.h:89-117) */
static void spline_shift(uint64_t *outc, const uint64_t *c, uint64_t shift, size_t d, size_t m, uint64_t M)
{
    uint64_t neg = mod_bits((0 - shift) & M, m);
    for (size_t i = 0; i <= d; i++) outc[i] = 0;
    for (size_t i = d + 1; i > 0; --i)
    {
        for (size_t j = d; j > 0; --j) outc[j] = (outc[j - 1] + outc[j] * neg) & M;
        outc[0] = (outc[0] * neg + c[i - 1]) & M;
    }
    for (size_t i = 0; i <= d; i++) outc[i] = mod_bits(outc[i], m);
}

/* src/kernel/spline.h:41-88 */
int oracle_dcfSplineKeyGen(void *key, size_t keyDataSize, void *e, size_t eSize, void *beta, size_t betaSize,
                           const void *alpha, size_t alphaSize, const void *seed0, size_t seedSize0,
                           const void *seed1, size_t seedSize1, const void *coeff, size_t coeffSize, size_t d,
                           const void *left, size_t leftSize, const void *right, size_t rightSize, size_t I, size_t n,
                           size_t m, size_t es, size_t N)
{
    if (!es_ok(es)) return E_ELEM_SIZE;
    const size_t C = d + 1, g = I * C;
    if (!(0 < n && n <= es * 8)) return E_BITWIDTH;
    if (!(0 < m && m <= es * 8)) return E_BITWIDTH;
    if (I == 0) return E_BOUNDARY_SIZE;
    if (keyDataSize != dcf_key_size(n, g, es, N)) return E_KEY_SIZE;
    if (eSize != N * es * g) return E_Z_SIZE;
    if (betaSize != N * es * g) return E_BETA_SIZE;
    if (alphaSize != N * es) return E_ALPHA_SIZE;
    if (seedSize0 != N * 16 || seedSize1 != N * 16) return E_SEED_SIZE;
    if (leftSize != I * es) return E_BOUNDARY_SIZE;
    if (rightSize != I * es) return E_BOUNDARY_SIZE;
    if (coeffSize != I * C * es) return E_BETA_SIZE;
    const uint64_t M    = ge_mask(es);
    const uint64_t maxv = mod_bits(M, n);
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t    i = (size_t)ii;
        uint64_t  a = ld(alpha, es, i);
        uint64_t *b = (uint64_t *)malloc(g * sizeof(uint64_t));
        uint64_t *c = (uint64_t *)malloc(C * sizeof(uint64_t));
        for (size_t q = 0; q < I; q++)
        {
            for (size_t j = 0; j < C; j++) c[j] = ld(coeff, es, q * C + j);
            spline_shift(b + q * C, c, a, d, m, M);
        }
        for (size_t j = 0; j < g; j++) st(beta, es, g * i + j, b[j]);
        dcf_key_t k = dcf_key_at(key, n, g, es, i, N);
        dcf_keygen_one(&k, (a - 1) & M, b, (const uint8_t *)seed0 + 16 * i, (const uint8_t *)seed1 + 16 * i, m);
        for (size_t q = 0; q < I; q++)
        {
            uint64_t Lq = mod_bits(ld(left, es, q), n), Rq = mod_bits(ld(right, es, q), n);
            uint64_t Rp = mod_bits(Rq + 1, n);
            uint64_t aL = mod_bits((Lq + a) & M, n), aR = mod_bits((Rq + a) & M, n), aRp = mod_bits((Rp + a) & M, n);
            int64_t  cr = (int64_t)(aL > aR) - (int64_t)(aL > Lq) + (int64_t)(aRp > Rp) + (int64_t)(aR == maxv);
            for (size_t j = 0; j < C; j++)
                st(e, es, g * i + q * C + j, mod_bits(((uint64_t)cr * b[q * C + j]) & M, m));
        }
        free(b);
        free(c);
    }
    return E_OK;
}

/* src/impl/spline.h:186-253 */
static uint64_t spline_eval_one(const dcf_key_t *k, uint64_t x, const void *sb, const void *se, size_t off,
                                const uint8_t *seed, int party, const void *left, const void *right, size_t I,
                                size_t d, size_t n, size_t m, size_t es, uint64_t *cur, uint64_t *nxt)
{
    const size_t   C = d + 1, g = I * C;
    const uint64_t M = ge_mask(es);
    x                = mod_bits(x, n);
    uint64_t acc[64];
    uint64_t *w = C <= 64 ? acc : (uint64_t *)malloc(C * sizeof(uint64_t));
    for (size_t j = 0; j < C; j++) w[j] = 0;
    for (size_t q = 0; q < I; q++)
    {
        uint64_t L  = mod_bits(ld(left, es, q), n), R = mod_bits(ld(right, es, q), n);
        uint64_t Rp = mod_bits(R + 1, n);
        if (q > 0 && mod_bits((ld(right, es, q - 1) + 1) & M, n) == L)
            memcpy(cur, nxt, g * sizeof(uint64_t));
        else
            dcf_eval_one(cur, k, (x - 1 - L) & M, seed, party, m);
        dcf_eval_one(nxt, k, (x - 1 - Rp) & M, seed, party, m);
        uint64_t cx = (uint64_t)(int64_t)((int)(x > L) - (int)(x > Rp));
        for (size_t j = 0; j < C; j++)
        {
            uint64_t bj = ld(sb, es, off + q * C + j), ej = ld(se, es, off + q * C + j);
            w[j]        = (w[j] + cx * bj - cur[q * C + j] + nxt[q * C + j] + ej) & M;
        }
    }
    /* sum_i w_i x^i mod 2^m (src/impl/spline.h:119-135) */
    uint64_t r = 0, pw = 1;
    for (size_t j = 0; j <= d; j++)
    {
        r  = (r + w[j] * pw) & M;
        pw = (pw * x) & M;
    }
    if (w != acc) free(w);
    return mod_bits(r, m);
}

/* src/kernel/spline.h:148-207 */
int oracle_dcfSplineEval(void *out, size_t outSize, const void *x, size_t xSize, const void *key,
                         size_t keyDataSize, const void *se, size_t seSize, const void *sb, size_t sbSize,
                         const void *seed, size_t seedSize, int party, const void *left, size_t leftSize,
                         const void *right, size_t rightSize, size_t I, size_t d, size_t n, size_t m, size_t es,
                         size_t N, void *cache, size_t cacheSize)
{
    if (!es_ok(es)) return E_ELEM_SIZE;
    const size_t C = d + 1, g = I * C;
    if (!(0 < n && n <= es * 8)) return E_BITWIDTH;
    if (!(0 < m && m <= es * 8)) return E_BITWIDTH;
    if (party != 0 && party != 1) return E_PARTY;
    if (I == 0) return E_BOUNDARY_SIZE;
    if (keyDataSize != dcf_key_size(n, g, es, N)) return E_KEY_SIZE;
    if (outSize != N * es) return E_OUT_SIZE;
    if (xSize != N * es) return E_X_SIZE;
    if (seSize != N * es * g) return E_Z_SIZE;
    if (sbSize != N * es * g) return E_BETA_SIZE;
    if (seedSize != N * 16) return E_SEED_SIZE;
    if (leftSize != I * es) return E_BOUNDARY_SIZE;
    if (rightSize != I * es) return E_BOUNDARY_SIZE;
    if (cache == NULL || cacheSize != spline_cache_size(n, I, d, es, N)) return E_CACHE_SIZE;
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t    i   = (size_t)ii;
        uint64_t *cur = (uint64_t *)malloc(g * sizeof(uint64_t));
        uint64_t *nxt = (uint64_t *)malloc(g * sizeof(uint64_t));
        dcf_key_t k   = dcf_key_at(key, n, g, es, i, N);
        uint64_t  v   = spline_eval_one(&k, ld(x, es, i), sb, se, g * i, (const uint8_t *)seed + 16 * i, party, left,
                                      right, I, d, n, m, es, cur, nxt);
        st(out, es, i, v);
        free(cur);
        free(nxt);
    }
    return E_OK;
}

/* ------------------------------------------------------------------------ */
/* OTTT                                                                      */
/* ------------------------------------------------------------------------ */

/* src/kernel/ottt.h:24-40; src/impl/ottt.h:15-24 */
int oracle_otttKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaSize, size_t n, size_t es,
                      size_t N)
{
    if (!es_ok(es)) return E_ELEM_SIZE;
    if (!(3 <= n && n <= es * 8)) return E_BITWIDTH;
    if (keyDataSize != N * (1ULL << n) / 8) return E_KEY_SIZE;
    if (alphaSize != N * es) return E_ALPHA_SIZE;
    const size_t per = (1ULL << n) / 8;
    for (size_t i = 0; i < N; i++)
    {
        uint64_t a = mod_bits(ld(alpha, es, i), n);
        ((uint8_t *)key)[i * per + a / 8] ^= (uint8_t)(1u << (a % 8));
    }
    return E_OK;
}

/* src/kernel/ottt.h:63-95; src/impl/ottt.h:26-192 */
int oracle_otttLutEval(void *outE, size_t outESize, void *outT, size_t outTSize, const void *x, size_t xSize,
                       const void *key, size_t keyDataSize, int party, const void *lut, size_t lutSize, size_t n,
                       size_t es, size_t N)
{
    if (!es_ok(es)) return E_ELEM_SIZE;
    if (!(3 <= n && n <= es * 8)) return E_BITWIDTH;
    if (party != 0 && party != 1) return E_PARTY;
    if (keyDataSize != N * (1ULL << n) / 8) return E_KEY_SIZE;
    const size_t each = es * (1ULL << n);
    /* the reference reports a malformed table with the seed-size code (src/kernel/ottt.h:75-78) */
    if (lutSize % each != 0) return E_SEED_SIZE;
    const size_t lutNum = lutSize / each;
    if (outESize != N * es) return E_OUT_SIZE;
    if (outTSize != N * es * lutNum) return E_OUT_SIZE;
    if (xSize != N * es) return E_X_SIZE;
    const uint64_t M     = ge_mask(es);
    const size_t   total = (size_t)1 << n, per = total / 8;
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long ii = 0; ii < (long long)N; ii++)
    {
        size_t         i  = (size_t)ii;
        const uint8_t *kb = (const uint8_t *)key + i * per;
        uint64_t       xv = ld(x, es, i);
        uint64_t       e  = 0;
        for (size_t j = 0; j < lutNum; j++) st(outT, es, i * lutNum + j, 0);
        uint64_t *tacc = (uint64_t *)calloc(lutNum ? lutNum : 1, sizeof(uint64_t));
        /* key bit b (little-endian bit order) selects LUT[(x - b) mod 2^n] */
        for (size_t b = 0; b < total; b++)
        {
            if (!((kb[b / 8] >> (b % 8)) & 1)) continue;
            e++;
            size_t idx = (size_t)(xv - b) & (total - 1);
            for (size_t j = 0; j < lutNum; j++) tacc[j] += ld(lut, es, idx + j * total);
        }
        if (party == 0)
        {
            e = 0 - e;
            for (size_t j = 0; j < lutNum; j++) tacc[j] = 0 - tacc[j];
        }
        st(outE, es, i, e & M);
        for (size_t j = 0; j < lutNum; j++) st(outT, es, i * lutNum + j, tacc[j] & M);
        free(tacc);
    }
    return E_OK;
}

/* ------------------------------------------------------------------------ */
/* AES-CTR PRNG (src/cpu/prng.cpp:20-148)                                    */
/* ------------------------------------------------------------------------ */

typedef struct
{
    uint8_t seed[16];
    uint8_t counter[16];
} prng_t;

void *oracle_prngInit(void)
{
    return calloc(1, sizeof(prng_t));
}

void oracle_prngRelease(void *p)
{
    free(p);
}

int oracle_prngSetCurrentSeed(void *p, const void *seed, const void *ctr)
{
    if (!p || !seed) return E_PRNG_ARG;
    prng_t *g = (prng_t *)p;
    memcpy(g->seed, seed, 16);
    if (ctr)
        memcpy(g->counter, ctr, 16);
    else
        memset(g->counter, 0, 16);
    return E_OK;
}

int oracle_prngGetCurrentSeed(const void *p, void *seed, void *ctr)
{
    if (!p || !seed) return E_PRNG_ARG;
    const prng_t *g = (const prng_t *)p;
    memcpy(seed, g->seed, 16);
    if (ctr) memcpy(ctr, g->counter, 16);
    return E_OK;
}

/* block k of the stream = AES_seed(counter + k), counter a 128-bit LE integer */
int oracle_prngGen(void *p, void *dst, size_t bw, size_t es, size_t N)
{
    if (!p || !dst) return E_PRNG_ARG;
    if (bw == 0 || bw > es * 8) return E_PRNG_BW;
    prng_t  *g     = (prng_t *)p;
    size_t   bytes = es * N;
    size_t   nblk  = (bytes + 15) / 16;
    uint64_t c0, c1;
    memcpy(&c0, g->counter, 8);
    memcpy(&c1, g->counter + 8, 8);
    const size_t CH = 4096;
#pragma omp parallel for schedule(static) num_threads(oracle_getNumThreads())
    for (long long cc = 0; cc < (long long)((nblk + CH - 1) / CH); cc++)
    {
        size_t   b0 = (size_t)cc * CH, b1 = b0 + CH < nblk ? b0 + CH : nblk;
        uint8_t *pt = (uint8_t *)malloc(16 * CH), *ct = (uint8_t *)malloc(16 * CH);
        for (size_t b = b0; b < b1; b++)
        {
            uint64_t lo = c0 + b, hi = c1 + (lo < c0 ? 1 : 0);
            memcpy(pt + 16 * (b - b0), &lo, 8);
            memcpy(pt + 16 * (b - b0) + 8, &hi, 8);
        }
        oracle_aes128_encrypt_blocks(g->seed, pt, ct, b1 - b0);
        size_t start = 16 * b0, end = 16 * b1 < bytes ? 16 * b1 : bytes;
        memcpy((uint8_t *)dst + start, ct, end - start);
        free(pt);
        free(ct);
    }
    uint64_t lo = c0 + nblk, hi = c1 + (lo < c0 ? 1 : 0);
    memcpy(g->counter, &lo, 8);
    memcpy(g->counter + 8, &hi, 8);
    return E_OK;
}
