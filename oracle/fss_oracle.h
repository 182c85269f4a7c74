/*
 * FSS oracle: a plain-C restatement of the reference CPU path of
 * oldprincess/FastFunctionSecretSharing (src/impl/<prim>.h, src/kernel/<prim>.h,
 * src/cpu/<prim>.cpp, src/<prim>.cpp).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  Only the test suite, the smoke check and the CPU-baseline
 * leg of the benchmark use it, and only as the checker / the timed CPU port.
 *
 * Every entry point takes exactly the arguments of the matching reference
 * `FastFss_cpu_<op>` function (include/FastFss/cpu/<prim>.h) and returns the same
 * FastFssError codes (include/FastFss/errors.h), so a test can call the
 * oracle and the device library with one argument list.
 *
 * Group elements of 1, 2, 4 and 8 bytes are supported.  Wider elements
 * (reference wideint types, src/impl/def.h:42-44) return
 * FAST_FSS_INVALID_ELEMENT_SIZE_ERROR here.
 *
 * The `cache` arguments are validated exactly as the reference does and
 * otherwise unused: the reference's prefix cache (src/impl/dcf.h:396-412) is a
 * pure re-use of identical intermediate values and does not change outputs.
 */
#ifndef FSS_ORACLE_H
#define FSS_ORACLE_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- AES-128 (aes.c) ---- */
uint8_t oracle_aes_sbox(uint8_t x);
void    oracle_aes128_expand_portable(const uint8_t key[16], uint8_t rk[176]);
void    oracle_aes128_encrypt_portable(const uint8_t rk[176], const uint8_t in[16], uint8_t out[16]);
void    oracle_aes128_encrypt_blocks(const uint8_t key[16], const uint8_t *in, uint8_t *out, size_t nblocks);
void    oracle_aes_force_portable(int on);
int     oracle_aes_uses_aesni(void);

/* ---- threading ---- */
int oracle_setNumThreads(int num);
int oracle_getNumThreads(void);

/* ---- sizes (src/dcf.cpp, dpf.cpp, grotto.cpp, mic.cpp, spline.cpp, ottt.cpp) ---- */
int oracle_dcfGetKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize,
                             size_t elementSize, size_t elementNum);
int oracle_dcfGetCacheDataSize(size_t *cacheDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize,
                               size_t elementSize, size_t elementNum);
int oracle_dpfGetKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize,
                             size_t elementSize, size_t elementNum);
int oracle_dpfGetCacheDataSize(size_t *cacheDataSize, size_t bitWidthIn, size_t elementSize, size_t elementNum);
int oracle_grottoGetKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t elementSize, size_t elementNum);
int oracle_grottoGetCacheDataSize(size_t *cacheDataSize, size_t bitWidthIn, size_t elementSize, size_t elementNum);
int oracle_dcfMICGetKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize,
                                size_t elementNum);
int oracle_dcfMICGetCacheDataSize(size_t *cacheDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize,
                                  size_t elementNum);
int oracle_dcfSplineGetKeyDataSize(size_t *keyDataSize, size_t degree, size_t intervalNum, size_t bitWidthIn,
                                   size_t bitWidthOut, size_t elementSize, size_t elementNum);
int oracle_dcfSplineGetCacheDataSize(size_t *cacheDataSize, size_t degree, size_t intervalNum, size_t bitWidthIn,
                                     size_t bitWidthOut, size_t elementSize, size_t elementNum);
int oracle_otttGetKeyDataSize(size_t *keyDataSize, size_t bitWidthIn, size_t elementNum);

/* ---- DCF (src/impl/dcf.h, src/kernel/dcf.h) ---- */
int oracle_dcfKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize, const void *beta,
                     size_t betaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1,
                     size_t seedDataSize1, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize,
                     size_t elementSize, size_t elementNum);
int oracle_dcfEval(void *sharedOut, size_t sharedOutSize, const void *maskedX, size_t maskedXDataSize,
                   const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId,
                   size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum,
                   void *cache, size_t cacheDataSize);

/* ---- DPF (src/impl/dpf.h, src/kernel/dpf.h) ---- */
int oracle_dpfKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize, const void *beta,
                     size_t betaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1,
                     size_t seedDataSize1, size_t bitWidthIn, size_t bitWidthOut, size_t groupSize,
                     size_t elementSize, size_t elementNum);
int oracle_dpfEval(void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize,
                   const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId,
                   size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize, size_t elementNum,
                   void *cache, size_t cacheDataSize);
int oracle_dpfEvalAll(void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize,
                      const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId,
                      size_t bitWidthIn, size_t bitWidthOut, size_t groupSize, size_t elementSize,
                      size_t elementNum, void *cache, size_t cacheDataSize);
int oracle_dpfEvalMulti(void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize,
                        const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId,
                        const void *point, size_t pointDataSize, size_t bitWidthIn, size_t bitWidthOut,
                        size_t groupSize, size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize);

/* ---- Grotto (src/impl/grotto.h, src/kernel/grotto.h) ---- */
int oracle_grottoKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize, const void *seed0,
                        size_t seedDataSize0, const void *seed1, size_t seedDataSize1, size_t bitWidthIn,
                        size_t elementSize, size_t elementNum);
int oracle_grottoEqEval(void *sharedBooleanOut, size_t sharedOutDataSize, const void *maskedX,
                        size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed,
                        size_t seedDataSize, int partyId, size_t bitWidthIn, size_t elementSize, size_t elementNum,
                        void *cache, size_t cacheDataSize);
int oracle_grottoEval(void *sharedBooleanOut, size_t sharedOutDataSize, const void *maskedX,
                      size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed,
                      size_t seedDataSize, bool equalBound, int partyId, size_t bitWidthIn, size_t elementSize,
                      size_t elementNum, void *cache, size_t cacheDataSize);
int oracle_grottoMICEval(void *sharedBooleanOut, size_t sharedOutDataSize, const void *maskedX,
                         size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed,
                         size_t seedDataSize, int partyId, const void *leftEndpoints, size_t leftEndpointsDataSize,
                         const void *rightEndpoints, size_t rightEndpointsDataSize, size_t bitWidthIn,
                         size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize);
int oracle_grottoIntervalLutEval(void *sharedOutE, size_t sharedOutEDataSize, void *sharedOutT,
                                 size_t sharedOutTDataSize, const void *maskedX, size_t maskedXDataSize,
                                 const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize,
                                 int partyId, const void *leftEndpoints, size_t leftEndpointsDataSize,
                                 const void *rightEndpoints, size_t rightEndpointsDataSize, const void *lookUpTable,
                                 size_t lookUpTableDataSize, size_t bitWidthIn, size_t bitWidthOut,
                                 size_t elementSize, size_t elementNum, void *cache, size_t cacheDataSize);

/* ---- DCF-based MIC (src/impl/mic.h, src/kernel/mic.h) ---- */
int oracle_dcfMICKeyGen(void *key, size_t keyDataSize, void *z, size_t zDataSize, const void *alpha,
                        size_t alphaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1,
                        size_t seedDataSize1, const void *leftEndpoints, size_t leftEndpointsDataSize,
                        const void *rightEndpoints, size_t rightEndpointsDataSize, size_t bitWidthIn,
                        size_t bitWidthOut, size_t elementSize, size_t elementNum);
int oracle_dcfMICEval(void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize,
                      const void *key, size_t keyDataSize, const void *sharedZ, size_t sharedZDataSize,
                      const void *seed, size_t seedDataSize, int partyId, const void *leftEndpoints,
                      size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize,
                      size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum, void *cache,
                      size_t cacheDataSize);

/* ---- DCF-based spline (src/impl/spline.h, src/kernel/spline.h) ---- */
int oracle_dcfSplineKeyGen(void *key, size_t keyDataSize, void *e, size_t eDataSize, void *beta,
                           size_t betaDataSize, const void *alpha, size_t alphaDataSize, const void *seed0,
                           size_t seedDataSize0, const void *seed1, size_t seedDataSize1, const void *coefficients,
                           size_t coefficientsDataSize, size_t degree, const void *leftEndpoints,
                           size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize,
                           size_t intervalNum, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize,
                           size_t elementNum);
int oracle_dcfSplineEval(void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize,
                         const void *key, size_t keyDataSize, const void *sharedE, size_t sharedEDataSize,
                         const void *sharedBeta, size_t sharedBetaDataSize, const void *seed, size_t seedDataSize,
                         int partyId, const void *leftEndpoints, size_t leftEndpointsDataSize,
                         const void *rightEndpoints, size_t rightEndpointsDataSize, size_t intervalNum,
                         size_t degree, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize,
                         size_t elementNum, void *cache, size_t cacheDataSize);

/* ---- OTTT (src/impl/ottt.h, src/kernel/ottt.h) ---- */
int oracle_otttKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize, size_t bitWidthIn,
                      size_t elementSize, size_t elementNum);
int oracle_otttLutEval(void *sharedOutE, size_t sharedOutEDataSize, void *sharedOutT, size_t sharedOutTDataSize,
                       const void *maskedX, size_t maskedXDataSize, const void *key, size_t keyDataSize,
                       int partyId, const void *lookUpTable, size_t lookUpTableDataSize, size_t bitWidthIn,
                       size_t elementSize, size_t elementNum);

/* ---- AES-CTR PRNG (src/cpu/prng.cpp) ---- */
void *oracle_prngInit(void);
void  oracle_prngRelease(void *prng);
int   oracle_prngSetCurrentSeed(void *prng, const void *seed128bit, const void *counter128bit);
int   oracle_prngGetCurrentSeed(const void *prng, void *seed128bit, void *counter128bit);
int   oracle_prngGen(void *prng, void *dst, size_t bitWidth, size_t elementSize, size_t elementNum);

#ifdef __cplusplus
}
#endif

#endif
