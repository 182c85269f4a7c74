/*
 * Memory and undefined-behaviour check of the oracle (test infrastructure).
 *
 * Built by `make -f oracle/sanitize.mk` with AddressSanitizer and
 * UndefinedBehaviorSanitizer (gcc, host code only) and run by
 * tests/test_oracle_sanitizers.py.  Every buffer is its own exact-size heap
 * block so an access one byte past any input, key, cache or output trips
 * ASan.  Each primitive runs keyGen and both parties' evaluation over a grid
 * of element sizes, widths and group sizes, once with AES-NI (when present)
 * and once with the portable AES, and checks the reconstruction property of
 * each primitive so the run also fails on wrong results.  Exit status 0 means
 * every case passed.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fss_oracle.h"

static uint64_t g_state = 0x9E3779B97F4A7C15ull;

static uint64_t rnd(void)
{
    /* splitmix64 */
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint64_t mask_of(size_t bits)
{
    return bits >= 64 ? ~0ull : ((1ull << bits) - 1);
}

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p)
    {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

static void put(void *buf, size_t es, size_t i, uint64_t v)
{
    memcpy((uint8_t *)buf + i * es, &v, es); /* little-endian host */
}

static uint64_t get(const void *buf, size_t es, size_t i)
{
    uint64_t v = 0;
    memcpy(&v, (const uint8_t *)buf + i * es, es);
    return v;
}

static void *rand_elems(size_t n, size_t bits, size_t es)
{
    void *p = xmalloc(n * es);
    for (size_t i = 0; i < n; i++) put(p, es, i, rnd() & mask_of(bits));
    return p;
}

static uint8_t *rand_bytes(size_t n)
{
    uint8_t *p = xmalloc(n);
    for (size_t i = 0; i < n; i++) p[i] = (uint8_t)rnd();
    return p;
}

static int g_fail = 0;

#define CHECK(cond, ...)                                                                                              \
    do                                                                                                                \
    {                                                                                                                 \
        if (!(cond))                                                                                                  \
        {                                                                                                             \
            fprintf(stderr, __VA_ARGS__);                                                                             \
            fprintf(stderr, "\n");                                                                                    \
            g_fail = 1;                                                                                               \
        }                                                                                                             \
    } while (0)

static const size_t kSizes[] = {1, 2, 4, 8};

static void widths_for(size_t es, size_t w[3])
{
    w[0] = 1, w[1] = 8 * es - 1, w[2] = 8 * es;
}

/* DCF and DPF: beta * [x < alpha] and beta * [x == alpha] */
static void check_dcf_dpf(int dpf)
{
    for (int a = 0; a < 4; a++)
    {
        const size_t es = kSizes[a];
        size_t       W[3];
        widths_for(es, W);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (size_t g = 1; g <= 5; g += 2)
                {
                    const size_t n = W[i], m = W[j], N = 7;
                    size_t       ks = 0;
                    int          rc = dpf ? oracle_dpfGetKeyDataSize(&ks, n, m, g, es, N)
                                          : oracle_dcfGetKeyDataSize(&ks, n, m, g, es, N);
                    CHECK(rc == 0, "size query %d", rc);
                    void    *alpha = rand_elems(N, n, es), *beta = rand_elems(N * g, m, es);
                    void    *x     = rand_elems(N, n, es);
                    uint8_t *s0 = rand_bytes(16 * N), *s1 = rand_bytes(16 * N), *key = xmalloc(ks);
                    put(x, es, 0, get(alpha, es, 0));
                    void *y0 = xmalloc(N * g * es), *y1 = xmalloc(N * g * es);
                    if (dpf)
                    {
                        rc = oracle_dpfKeyGen(key, ks, alpha, N * es, beta, N * g * es, s0, 16 * N, s1, 16 * N, n, m,
                                              g, es, N);
                        rc |= oracle_dpfEval(y0, N * g * es, x, N * es, key, ks, s0, 16 * N, 0, n, m, g, es, N, NULL,
                                             0);
                        rc |= oracle_dpfEval(y1, N * g * es, x, N * es, key, ks, s1, 16 * N, 1, n, m, g, es, N, NULL,
                                             0);
                    }
                    else
                    {
                        size_t cs = 0;
                        oracle_dcfGetCacheDataSize(&cs, n, m, g, es, N);
                        uint8_t *cache = xmalloc(cs);
                        rc = oracle_dcfKeyGen(key, ks, alpha, N * es, beta, N * g * es, s0, 16 * N, s1, 16 * N, n, m,
                                              g, es, N);
                        rc |= oracle_dcfEval(y0, N * g * es, x, N * es, key, ks, s0, 16 * N, 0, n, m, g, es, N, NULL,
                                             0);
                        rc |= oracle_dcfEval(y1, N * g * es, x, N * es, key, ks, s1, 16 * N, 1, n, m, g, es, N, cache,
                                             cs);
                        free(cache);
                    }
                    CHECK(rc == 0, "%s rc %d n=%zu m=%zu g=%zu es=%zu", dpf ? "dpf" : "dcf", rc, n, m, g, es);
                    for (size_t e = 0; e < N; e++)
                    {
                        const uint64_t xe = get(x, es, e), ae = get(alpha, es, e);
                        const int      hit = dpf ? (xe == ae) : (xe < ae);
                        for (size_t k = 0; k < g; k++)
                        {
                            const uint64_t r = (get(y0, es, e * g + k) + get(y1, es, e * g + k)) & mask_of(m);
                            const uint64_t w = hit ? (get(beta, es, e * g + k) & mask_of(m)) : 0;
                            CHECK(r == w, "%s reconstruction n=%zu m=%zu g=%zu es=%zu e=%zu", dpf ? "dpf" : "dcf", n,
                                  m, g, es, e);
                        }
                    }
                    if (dpf && n <= 8)
                    {
                        const size_t D = (size_t)1 << n, P = 3;
                        void        *all0 = xmalloc(N * D * g * es), *all1 = xmalloc(N * D * g * es);
                        void        *pts = rand_elems(P, n, es), *m0 = xmalloc(N * P * g * es);
                        rc = oracle_dpfEvalAll(all0, N * D * g * es, x, N * es, key, ks, s0, 16 * N, 0, n, m, g, es, N,
                                               NULL, 0);
                        rc |= oracle_dpfEvalAll(all1, N * D * g * es, x, N * es, key, ks, s1, 16 * N, 1, n, m, g, es,
                                                N, NULL, 0);
                        rc |= oracle_dpfEvalMulti(m0, N * P * g * es, x, N * es, key, ks, s0, 16 * N, 0, pts, P * es,
                                                  n, m, g, es, N, NULL, 0);
                        CHECK(rc == 0, "dpf evalAll/multi rc %d", rc);
                        free(all0), free(all1), free(pts), free(m0);
                    }
                    free(alpha), free(beta), free(x), free(s0), free(s1), free(key), free(y0), free(y1);
                }
    }
}

/* Grotto: XOR of the shares == [x < alpha], [x <= alpha], [x == alpha]; MIC and LUT run to completion */
static void check_grotto(void)
{
    for (int a = 0; a < 4; a++)
    {
        const size_t es = kSizes[a];
        for (size_t n = 6; n <= 8 * es; n += (8 * es > 6 ? 8 * es - 6 : 1))
        {
            const size_t N = 9, I = 3;
            size_t       ks = 0;
            oracle_grottoGetKeyDataSize(&ks, n, es, N);
            void    *alpha = rand_elems(N, n, es), *x = rand_elems(N, n, es);
            uint8_t *s0 = rand_bytes(16 * N), *s1 = rand_bytes(16 * N), *key = xmalloc(ks);
            put(x, es, 0, get(alpha, es, 0));
            put(x, es, 1, 0);
            int rc = oracle_grottoKeyGen(key, ks, alpha, N * es, s0, 16 * N, s1, 16 * N, n, es, N);
            for (int mode = 0; mode < 3; mode++)
            {
                void *y0 = xmalloc(N * es), *y1 = xmalloc(N * es);
                if (mode == 2)
                {
                    rc |= oracle_grottoEqEval(y0, N * es, x, N * es, key, ks, s0, 16 * N, 0, n, es, N, NULL, 0);
                    rc |= oracle_grottoEqEval(y1, N * es, x, N * es, key, ks, s1, 16 * N, 1, n, es, N, NULL, 0);
                }
                else
                {
                    rc |= oracle_grottoEval(y0, N * es, x, N * es, key, ks, s0, 16 * N, mode == 1, 0, n, es, N, NULL,
                                            0);
                    rc |= oracle_grottoEval(y1, N * es, x, N * es, key, ks, s1, 16 * N, mode == 1, 1, n, es, N, NULL,
                                            0);
                }
                for (size_t e = 0; e < N; e++)
                {
                    const uint64_t xe = get(x, es, e), ae = get(alpha, es, e);
                    const int      want = mode == 0 ? xe < ae : (mode == 1 ? xe <= ae : xe == ae);
                    CHECK(((get(y0, es, e) ^ get(y1, es, e)) & 1) == (uint64_t)want, "grotto mode %d n=%zu es=%zu",
                          mode, n, es);
                }
                free(y0), free(y1);
            }
            void *left = rand_elems(I, n, es), *right = rand_elems(I, n, es), *lut = rand_elems(2 * I, 8 * es, es);
            void *mo = xmalloc(N * I * es), *oe = xmalloc(N * es), *ot = xmalloc(N * 2 * es);
            rc |= oracle_grottoMICEval(mo, N * I * es, x, N * es, key, ks, s0, 16 * N, 0, left, I * es, right, I * es,
                                       n, es, N, NULL, 0);
            rc |= oracle_grottoIntervalLutEval(oe, N * es, ot, N * 2 * es, x, N * es, key, ks, s0, 16 * N, 0, left,
                                               I * es, right, I * es, lut, 2 * I * es, n, 8 * es, es, N, NULL, 0);
            CHECK(rc == 0, "grotto rc %d n=%zu es=%zu", rc, n, es);
            free(alpha), free(x), free(s0), free(s1), free(key), free(left), free(right), free(lut), free(mo),
                free(oe), free(ot);
        }
    }
}

/* MIC (contiguous unsigned intervals inside the positive half), spline, OTTT */
static void check_mic_spline_ottt(void)
{
    for (int a = 1; a < 4; a++)
    {
        const size_t es = kSizes[a], n = 8 * es - 2, m = 8 * es, N = 5, I = 3, d = 2, g = I * (d + 1);
        void        *alpha = rand_elems(N, n, es), *x = rand_elems(N, n, es);
        uint8_t     *s0 = rand_bytes(16 * N), *s1 = rand_bytes(16 * N);
        const uint64_t q = (1ull << (n - 2)) / I;
        void          *left = xmalloc(I * es), *right = xmalloc(I * es);
        for (size_t i = 0; i < I; i++) put(left, es, i, i * q), put(right, es, i, (i + 1) * q - 1);
        size_t ks = 0;
        oracle_dcfMICGetKeyDataSize(&ks, n, m, es, N);
        uint8_t *key = xmalloc(ks);
        void    *z = xmalloc(N * I * es), *y0 = xmalloc(N * I * es), *y1 = xmalloc(N * I * es);
        void    *z0 = rand_elems(N * I, m, es), *z1 = xmalloc(N * I * es);
        int      rc = oracle_dcfMICKeyGen(key, ks, z, N * I * es, alpha, N * es, s0, 16 * N, s1, 16 * N, left,
                                          I * es, right, I * es, n, m, es, N);
        for (size_t i = 0; i < N * I; i++) put(z1, es, i, (get(z, es, i) - get(z0, es, i)) & mask_of(m));
        void *mx = xmalloc(N * es);
        for (size_t e = 0; e < N; e++) put(mx, es, e, (get(x, es, e) + get(alpha, es, e)) & mask_of(n));
        rc |= oracle_dcfMICEval(y0, N * I * es, mx, N * es, key, ks, z0, N * I * es, s0, 16 * N, 0, left, I * es,
                                right, I * es, n, m, es, N, NULL, 0);
        rc |= oracle_dcfMICEval(y1, N * I * es, mx, N * es, key, ks, z1, N * I * es, s1, 16 * N, 1, left, I * es,
                                right, I * es, n, m, es, N, NULL, 0);
        CHECK(rc == 0, "mic rc %d es=%zu", rc, es);
        for (size_t e = 0; e < N; e++)
            for (size_t i = 0; i < I; i++)
            {
                const uint64_t xe = get(x, es, e);
                const uint64_t want = xe >= get(left, es, i) && xe <= get(right, es, i);
                const uint64_t r    = (get(y0, es, e * I + i) + get(y1, es, e * I + i)) & mask_of(m);
                CHECK(r == want, "mic reconstruction es=%zu e=%zu i=%zu", es, e, i);
            }
        free(key), free(z), free(y0), free(y1), free(z0), free(z1);

        /* spline: contiguous intervals covering [0, 2^n) */
        for (size_t i = 0; i < I; i++)
            put(left, es, i, i * ((1ull << n) / I)), put(right, es, i, i + 1 == I ? mask_of(n) : (i + 1) * ((1ull << n) / I) - 1);
        void *coeff = rand_elems(g, m, es);
        oracle_dcfSplineGetKeyDataSize(&ks, d, I, n, m, es, N);
        size_t cs = 0;
        oracle_dcfSplineGetCacheDataSize(&cs, d, I, n, m, es, N);
        key         = xmalloc(ks);
        uint8_t *ca = xmalloc(cs);
        void    *e_ = xmalloc(N * g * es), *b_ = xmalloc(N * g * es), *so = xmalloc(N * es);
        rc = oracle_dcfSplineKeyGen(key, ks, e_, N * g * es, b_, N * g * es, alpha, N * es, s0, 16 * N, s1, 16 * N,
                                    coeff, g * es, d, left, I * es, right, I * es, I, n, m, es, N);
        rc |= oracle_dcfSplineEval(so, N * es, mx, N * es, key, ks, e_, N * g * es, b_, N * g * es, s0, 16 * N, 0,
                                   left, I * es, right, I * es, I, d, n, m, es, N, ca, cs);
        CHECK(rc == 0, "spline rc %d es=%zu", rc, es);
        free(coeff), free(key), free(ca), free(e_), free(b_), free(so);

        /* OTTT */
        const size_t nb = 8, lutNum = 2;
        oracle_otttGetKeyDataSize(&ks, nb, N);
        uint8_t *k0 = rand_bytes(ks), *k1 = xmalloc(ks);
        memcpy(k1, k0, ks);
        void *al = rand_elems(N, nb, es), *xx = rand_elems(N, nb, es), *lut = rand_elems(lutNum << nb, m, es);
        void *E = xmalloc(N * es), *T = xmalloc(N * lutNum * es);
        rc = oracle_otttKeyGen(k1, ks, al, N * es, nb, es, N);
        rc |= oracle_otttLutEval(E, N * es, T, N * lutNum * es, xx, N * es, k1, ks, 1, lut, (lutNum << nb) * es, nb,
                                 es, N);
        CHECK(rc == 0, "ottt rc %d es=%zu", rc, es);
        free(k0), free(k1), free(al), free(xx), free(lut), free(E), free(T);
        free(alpha), free(x), free(s0), free(s1), free(left), free(right), free(mx);
    }
}

static void check_prng_and_errors(void)
{
    void   *g = oracle_prngInit();
    uint8_t seed[16], ctr[16];
    for (int i = 0; i < 16; i++) seed[i] = (uint8_t)i, ctr[i] = (uint8_t)(255 - i);
    CHECK(oracle_prngSetCurrentSeed(g, seed, ctr) == 0, "prng seed");
    for (size_t es = 1; es <= 8; es *= 2)
    {
        void *buf = xmalloc(37 * es);
        CHECK(oracle_prngGen(g, buf, 8 * es, es, 37) == 0, "prng gen es=%zu", es);
        free(buf);
    }
    oracle_prngRelease(g);
    /* a wrong size is rejected before any buffer is touched */
    uint8_t small[4];
    CHECK(oracle_dcfEval(small, 4, small, 4, small, 4, small, 4, 0, 32, 32, 1, 4, 1, NULL, 0) < 0, "dcf size check");
    CHECK(oracle_dpfEval(small, 4, small, 4, small, 4, small, 4, 2, 32, 32, 1, 4, 1, NULL, 0) < 0, "dpf party check");
}

int main(void)
{
    for (int portable = 0; portable < 2; portable++)
    {
        oracle_aes_force_portable(portable);
        check_dcf_dpf(0);
        check_dcf_dpf(1);
        check_grotto();
        check_mic_spline_ottt();
        check_prng_and_errors();
    }
    if (g_fail) return 1;
    printf("sanitized oracle run: all checks passed\n");
    return 0;
}
