// The reference's host-buffer API (FastFss_cpu_*, include/FastFss/cpu/*.h),
// served on the GPU: each call stages the caller's host buffers in device
// memory on the calling thread's default stream, runs the FastFss_cuda_*
// entry point of the same name (so validation, error codes and results are
// those of the device path), copies the outputs back and returns once they
// are in place.  Nothing here evaluates on the host.
#include <hip/hip_runtime.h>

#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "FastFss/cpu/config.h"
#include "FastFss/cpu/dcf.h"
#include "FastFss/cpu/dpf.h"
#include "FastFss/cpu/grotto.h"
#include "FastFss/cpu/mic.h"
#include "FastFss/cpu/ottt.h"
#include "FastFss/cpu/prng.h"
#include "FastFss/cpu/spline.h"
#include "FastFss/cuda/dcf.h"
#include "FastFss/cuda/dpf.h"
#include "FastFss/cuda/grotto.h"
#include "FastFss/cuda/mic.h"
#include "FastFss/cuda/ottt.h"
#include "FastFss/cuda/prng.h"
#include "FastFss/cuda/spline.h"
#include "FastFss/errors.h"

namespace {

// Device copies of one call's host buffers, all on the per-thread default
// stream.  A host pointer that is NULL with a non-zero size fails the call
// before anything is launched.
class Staging
{
public:
    hipStream_t stream = hipStreamPerThread;

    const void *in(const void *host, size_t bytes) { return put(const_cast<void *>(host), bytes, true, false); }
    void       *out(void *host, size_t bytes) { return put(host, bytes, false, true); }
    void       *inout(void *host, size_t bytes) { return put(host, bytes, true, true); }
    void       *scratch(void *host, size_t bytes) { return put(host, bytes, false, false); }
    bool        ok() const { return !failed_; }
    void       *stream_ptr() { return &stream; }

    // Copies the outputs back if rc is success, frees the device buffers and
    // waits for the stream.
    int finish(int rc)
    {
        if (rc == FAST_FSS_SUCCESS)
            for (const Buf &b : bufs_)
                if (b.back && hipMemcpyAsync(b.host, b.dev, b.bytes, hipMemcpyDeviceToHost, stream) != hipSuccess)
                    rc = FAST_FSS_RUNTIME_ERROR;
        for (const Buf &b : bufs_) (void)hipFreeAsync(b.dev, stream);
        if (hipStreamSynchronize(stream) != hipSuccess && rc == FAST_FSS_SUCCESS) rc = FAST_FSS_RUNTIME_ERROR;
        bufs_.clear();
        return rc;
    }

    ~Staging()
    {
        if (!bufs_.empty()) (void)finish(FAST_FSS_RUNTIME_ERROR);
    }

private:
    struct Buf
    {
        void  *host;
        void  *dev;
        size_t bytes;
        bool   back;
    };
    std::vector<Buf> bufs_;
    bool             failed_ = false;

    void *put(void *host, size_t bytes, bool copyIn, bool copyBack)
    {
        if (bytes == 0 || failed_) return nullptr;
        if (host == nullptr)
        {
            failed_ = true;
            return nullptr;
        }
        void *dev = nullptr;
        if (hipMallocAsync(&dev, bytes, stream) != hipSuccess)
        {
            failed_ = true;
            return nullptr;
        }
        bufs_.push_back({host, dev, bytes, copyBack});
        if (copyIn && hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, stream) != hipSuccess) failed_ = true;
        return dev;
    }
};

// Runs `call()` once every buffer is staged; its result goes through finish().
template <typename F>
int staged(Staging &s, F call)
{
    if (!s.ok()) return s.finish(FAST_FSS_RUNTIME_ERROR);
    return s.finish(call());
}

int        g_num_threads = (int)std::thread::hardware_concurrency();
std::mutex g_num_threads_mu;

// A host-API generator: the device-side generator it drives and that generator's device.
struct HostPrng
{
    void *dev_prng;
    int   device;
};

// Makes the generator's device current for the scope (restoring the previous one).
class DeviceScope
{
public:
    explicit DeviceScope(int dev)
    {
        if (hipGetDevice(&prev_) != hipSuccess)
        {
            ok_ = false;
            return;
        }
        if (prev_ != dev && hipSetDevice(dev) != hipSuccess) ok_ = false;
    }
    ~DeviceScope()
    {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev_) (void)hipSetDevice(prev_);
    }
    bool ok() const { return ok_; }

private:
    int  prev_ = 0;
    bool ok_   = true;
};

} // namespace

// ------------------------------------------------------------------ config ----

extern "C" FAST_FSS_API int FastFss_cpu_setNumThreads(int num)
{
    if (num <= 0) return -2;
    std::lock_guard<std::mutex> lock(g_num_threads_mu);
    g_num_threads = num;
    return 0;
}

extern "C" FAST_FSS_API int FastFss_cpu_getNumThreads(void)
{
    std::lock_guard<std::mutex> lock(g_num_threads_mu);
    return g_num_threads;
}

// --------------------------------------------------------------------- DCF ----

extern "C" FAST_FSS_API int FastFss_cpu_dcfKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize,
                                                  const void *beta, size_t betaDataSize, const void *seed0,
                                                  size_t seedDataSize0, const void *seed1, size_t seedDataSize1,
                                                  size_t bitWidthIn, size_t bitWidthOut, size_t groupSize,
                                                  size_t elementSize, size_t elementNum)
{
    Staging s;
    void       *k  = s.out(key, keyDataSize);
    const void *a  = s.in(alpha, alphaDataSize);
    const void *b  = s.in(beta, betaDataSize);
    const void *d0 = s.in(seed0, seedDataSize0);
    const void *d1 = s.in(seed1, seedDataSize1);
    return staged(s, [&] {
        return FastFss_cuda_dcfKeyGen(k, keyDataSize, a, alphaDataSize, b, betaDataSize, d0, seedDataSize0, d1,
                                      seedDataSize1, bitWidthIn, bitWidthOut, groupSize, elementSize, elementNum,
                                      s.stream_ptr());
    });
}

extern "C" FAST_FSS_API int FastFss_cpu_dcfEval(void *sharedOut, size_t sharedOutSize, const void *maskedX,
                                                size_t maskedXDataSize, const void *key, size_t keyDataSize,
                                                const void *seed, size_t seedDataSize, int partyId, size_t bitWidthIn,
                                                size_t bitWidthOut, size_t groupSize, size_t elementSize,
                                                size_t elementNum, void *cache, size_t cacheDataSize)
{
    Staging s;
    void       *o  = s.out(sharedOut, sharedOutSize);
    const void *x  = s.in(maskedX, maskedXDataSize);
    const void *k  = s.in(key, keyDataSize);
    const void *sd = s.in(seed, seedDataSize);
    void       *c  = s.scratch(cache, cacheDataSize);
    return staged(s, [&] {
        return FastFss_cuda_dcfEval(o, sharedOutSize, x, maskedXDataSize, k, keyDataSize, sd, seedDataSize, partyId,
                                    bitWidthIn, bitWidthOut, groupSize, elementSize, elementNum, c, cacheDataSize,
                                    s.stream_ptr());
    });
}

// --------------------------------------------------------------------- DPF ----

extern "C" FAST_FSS_API int FastFss_cpu_dpfKeyGen(void *key, size_t keyDataSize, const void *alpha, size_t alphaDataSize,
                                                  const void *beta, size_t betaDataSize, const void *seed0,
                                                  size_t seedDataSize0, const void *seed1, size_t seedDataSize1,
                                                  size_t bitWidthIn, size_t bitWidthOut, size_t groupSize,
                                                  size_t elementSize, size_t elementNum)
{
    Staging s;
    void       *k  = s.out(key, keyDataSize);
    const void *a  = s.in(alpha, alphaDataSize);
    const void *b  = s.in(beta, betaDataSize);
    const void *d0 = s.in(seed0, seedDataSize0);
    const void *d1 = s.in(seed1, seedDataSize1);
    return staged(s, [&] {
        return FastFss_cuda_dpfKeyGen(k, keyDataSize, a, alphaDataSize, b, betaDataSize, d0, seedDataSize0, d1,
                                      seedDataSize1, bitWidthIn, bitWidthOut, groupSize, elementSize, elementNum,
                                      s.stream_ptr());
    });
}

#define FSS_CPU_DPF_EVAL(NAME, CUDA_NAME)                                                                        \
    extern "C" FAST_FSS_API int NAME(void *sharedOut, size_t sharedOutDataSize, const void *maskedX,               \
                                     size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed, \
                                     size_t seedDataSize, int partyId, size_t bitWidthIn, size_t bitWidthOut,      \
                                     size_t groupSize, size_t elementSize, size_t elementNum, void *cache,         \
                                     size_t cacheDataSize)                                                         \
    {                                                                                                              \
        Staging s;                                                                                                 \
        void       *o  = s.out(sharedOut, sharedOutDataSize);                                                     \
        const void *x  = s.in(maskedX, maskedXDataSize);                                                           \
        const void *k  = s.in(key, keyDataSize);                                                                   \
        const void *sd = s.in(seed, seedDataSize);                                                                 \
        void       *c  = s.scratch(cache, cacheDataSize);                                                          \
        return staged(s, [&] {                                                                                     \
            return CUDA_NAME(o, sharedOutDataSize, x, maskedXDataSize, k, keyDataSize, sd, seedDataSize, partyId,  \
                             bitWidthIn, bitWidthOut, groupSize, elementSize, elementNum, c, cacheDataSize,        \
                             s.stream_ptr());                                                                      \
        });                                                                                                        \
    }

FSS_CPU_DPF_EVAL(FastFss_cpu_dpfEval, FastFss_cuda_dpfEval)
FSS_CPU_DPF_EVAL(FastFss_cpu_dpfEvalAll, FastFss_cuda_dpfEvalAll)
#undef FSS_CPU_DPF_EVAL

extern "C" FAST_FSS_API int FastFss_cpu_dpfEvalMulti(void *sharedOut, size_t sharedOutDataSize, const void *maskedX,
                                                     size_t maskedXDataSize, const void *key, size_t keyDataSize,
                                                     const void *seed, size_t seedDataSize, int partyId,
                                                     const void *point, size_t pointDataSize, size_t bitWidthIn,
                                                     size_t bitWidthOut, size_t groupSize, size_t elementSize,
                                                     size_t elementNum, void *cache, size_t cacheDataSize)
{
    Staging s;
    void       *o  = s.out(sharedOut, sharedOutDataSize);
    const void *x  = s.in(maskedX, maskedXDataSize);
    const void *k  = s.in(key, keyDataSize);
    const void *sd = s.in(seed, seedDataSize);
    const void *p  = s.in(point, pointDataSize);
    void       *c  = s.scratch(cache, cacheDataSize);
    return staged(s, [&] {
        return FastFss_cuda_dpfEvalMulti(o, sharedOutDataSize, x, maskedXDataSize, k, keyDataSize, sd, seedDataSize,
                                         partyId, p, pointDataSize, bitWidthIn, bitWidthOut, groupSize, elementSize,
                                         elementNum, c, cacheDataSize, s.stream_ptr());
    });
}

// ------------------------------------------------------------------ Grotto ----

extern "C" FAST_FSS_API int FastFss_cpu_grottoKeyGen(void *key, size_t keyDataSize, const void *alpha,
                                                     size_t alphaDataSize, const void *seed0, size_t seedDataSize0,
                                                     const void *seed1, size_t seedDataSize1, size_t bitWidthIn,
                                                     size_t elementSize, size_t elementNum)
{
    Staging s;
    void       *k  = s.out(key, keyDataSize);
    const void *a  = s.in(alpha, alphaDataSize);
    const void *d0 = s.in(seed0, seedDataSize0);
    const void *d1 = s.in(seed1, seedDataSize1);
    return staged(s, [&] {
        return FastFss_cuda_grottoKeyGen(k, keyDataSize, a, alphaDataSize, d0, seedDataSize0, d1, seedDataSize1,
                                         bitWidthIn, elementSize, elementNum, s.stream_ptr());
    });
}

extern "C" FAST_FSS_API int FastFss_cpu_grottoEqEval(void *sharedBooleanOut, size_t sharedOutDataSize,
                                                     const void *maskedX, size_t maskedXDataSize, const void *key,
                                                     size_t keyDataSize, const void *seed, size_t seedDataSize,
                                                     int partyId, size_t bitWidthIn, size_t elementSize,
                                                     size_t elementNum, void *cache, size_t cacheDataSize)
{
    Staging s;
    void       *o  = s.out(sharedBooleanOut, sharedOutDataSize);
    const void *x  = s.in(maskedX, maskedXDataSize);
    const void *k  = s.in(key, keyDataSize);
    const void *sd = s.in(seed, seedDataSize);
    void       *c  = s.scratch(cache, cacheDataSize);
    return staged(s, [&] {
        return FastFss_cuda_grottoEqEval(o, sharedOutDataSize, x, maskedXDataSize, k, keyDataSize, sd, seedDataSize,
                                         partyId, bitWidthIn, elementSize, elementNum, c, cacheDataSize,
                                         s.stream_ptr());
    });
}

extern "C" FAST_FSS_API int FastFss_cpu_grottoEval(void *sharedBooleanOut, size_t sharedOutDataSize,
                                                   const void *maskedX, size_t maskedXDataSize, const void *key,
                                                   size_t keyDataSize, const void *seed, size_t seedDataSize,
                                                   bool equalBound, int partyId, size_t bitWidthIn,
                                                   size_t elementSize, size_t elementNum, void *cache,
                                                   size_t cacheDataSize)
{
    Staging s;
    void       *o  = s.out(sharedBooleanOut, sharedOutDataSize);
    const void *x  = s.in(maskedX, maskedXDataSize);
    const void *k  = s.in(key, keyDataSize);
    const void *sd = s.in(seed, seedDataSize);
    void       *c  = s.scratch(cache, cacheDataSize);
    return staged(s, [&] {
        return FastFss_cuda_grottoEval(o, sharedOutDataSize, x, maskedXDataSize, k, keyDataSize, sd, seedDataSize,
                                       equalBound, partyId, bitWidthIn, elementSize, elementNum, c, cacheDataSize,
                                       s.stream_ptr());
    });
}

extern "C" FAST_FSS_API int FastFss_cpu_grottoMICEval(void *sharedBooleanOut, size_t sharedOutDataSize,
                                                      const void *maskedX, size_t maskedXDataSize, const void *key,
                                                      size_t keyDataSize, const void *seed, size_t seedDataSize,
                                                      int partyId, const void *leftEndpoints,
                                                      size_t leftEndpointsDataSize, const void *rightEndpoints,
                                                      size_t rightEndpointsDataSize, size_t bitWidthIn,
                                                      size_t elementSize, size_t elementNum, void *cache,
                                                      size_t cacheDataSize)
{
    Staging s;
    void       *o  = s.out(sharedBooleanOut, sharedOutDataSize);
    const void *x  = s.in(maskedX, maskedXDataSize);
    const void *k  = s.in(key, keyDataSize);
    const void *sd = s.in(seed, seedDataSize);
    const void *l  = s.in(leftEndpoints, leftEndpointsDataSize);
    const void *r  = s.in(rightEndpoints, rightEndpointsDataSize);
    void       *c  = s.scratch(cache, cacheDataSize);
    return staged(s, [&] {
        return FastFss_cuda_grottoMICEval(o, sharedOutDataSize, x, maskedXDataSize, k, keyDataSize, sd, seedDataSize,
                                          partyId, l, leftEndpointsDataSize, r, rightEndpointsDataSize, bitWidthIn,
                                          elementSize, elementNum, c, cacheDataSize, s.stream_ptr());
    });
}

extern "C" FAST_FSS_API int FastFss_cpu_grottoIntervalLutEval(
    void *sharedOutE, size_t sharedOutEDataSize, void *sharedOutT, size_t sharedOutTDataSize, const void *maskedX,
    size_t maskedXDataSize, const void *key, size_t keyDataSize, const void *seed, size_t seedDataSize, int partyId,
    const void *leftEndpoints, size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize,
    const void *lookUpTable, size_t lookUpTableDataSize, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize,
    size_t elementNum, void *cache, size_t cacheDataSize)
{
    Staging s;
    void       *oe = s.out(sharedOutE, sharedOutEDataSize);
    void       *ot = s.out(sharedOutT, sharedOutTDataSize);
    const void *x  = s.in(maskedX, maskedXDataSize);
    const void *k  = s.in(key, keyDataSize);
    const void *sd = s.in(seed, seedDataSize);
    const void *l  = s.in(leftEndpoints, leftEndpointsDataSize);
    const void *r  = s.in(rightEndpoints, rightEndpointsDataSize);
    const void *t  = s.in(lookUpTable, lookUpTableDataSize);
    void       *c  = s.scratch(cache, cacheDataSize);
    return staged(s, [&] {
        return FastFss_cuda_grottoIntervalLutEval(oe, sharedOutEDataSize, ot, sharedOutTDataSize, x, maskedXDataSize,
                                                  k, keyDataSize, sd, seedDataSize, partyId, l, leftEndpointsDataSize,
                                                  r, rightEndpointsDataSize, t, lookUpTableDataSize, bitWidthIn,
                                                  bitWidthOut, elementSize, elementNum, c, cacheDataSize,
                                                  s.stream_ptr());
    });
}

// --------------------------------------------------------------------- MIC ----

extern "C" FAST_FSS_API int FastFss_cpu_dcfMICKeyGen(void *key, size_t keyDataSize, void *z, size_t zDataSize,
                                                     const void *alpha, size_t alphaDataSize, const void *seed0,
                                                     size_t seedDataSize0, const void *seed1, size_t seedDataSize1,
                                                     const void *leftEndpoints, size_t leftEndpointsDataSize,
                                                     const void *rightEndpoints, size_t rightEndpointsDataSize,
                                                     size_t bitWidthIn, size_t bitWidthOut, size_t elementSize,
                                                     size_t elementNum)
{
    Staging s;
    void       *k  = s.out(key, keyDataSize);
    void       *zz = s.out(z, zDataSize);
    const void *a  = s.in(alpha, alphaDataSize);
    const void *d0 = s.in(seed0, seedDataSize0);
    const void *d1 = s.in(seed1, seedDataSize1);
    const void *l  = s.in(leftEndpoints, leftEndpointsDataSize);
    const void *r  = s.in(rightEndpoints, rightEndpointsDataSize);
    return staged(s, [&] {
        return FastFss_cuda_dcfMICKeyGen(k, keyDataSize, zz, zDataSize, a, alphaDataSize, d0, seedDataSize0, d1,
                                         seedDataSize1, l, leftEndpointsDataSize, r, rightEndpointsDataSize,
                                         bitWidthIn, bitWidthOut, elementSize, elementNum, s.stream_ptr());
    });
}

extern "C" FAST_FSS_API int FastFss_cpu_dcfMICEval(void *sharedOut, size_t sharedOutDataSize, const void *maskedX,
                                                   size_t maskedXDataSize, const void *key, size_t keyDataSize,
                                                   const void *sharedZ, size_t sharedZDataSize, const void *seed,
                                                   size_t seedDataSize, int partyId, const void *leftEndpoints,
                                                   size_t leftEndpointsDataSize, const void *rightEndpoints,
                                                   size_t rightEndpointsDataSize, size_t bitWidthIn,
                                                   size_t bitWidthOut, size_t elementSize, size_t elementNum,
                                                   void *cache, size_t cacheDataSize)
{
    Staging s;
    void       *o  = s.out(sharedOut, sharedOutDataSize);
    const void *x  = s.in(maskedX, maskedXDataSize);
    const void *k  = s.in(key, keyDataSize);
    const void *zz = s.in(sharedZ, sharedZDataSize);
    const void *sd = s.in(seed, seedDataSize);
    const void *l  = s.in(leftEndpoints, leftEndpointsDataSize);
    const void *r  = s.in(rightEndpoints, rightEndpointsDataSize);
    void       *c  = s.scratch(cache, cacheDataSize);
    return staged(s, [&] {
        return FastFss_cuda_dcfMICEval(o, sharedOutDataSize, x, maskedXDataSize, k, keyDataSize, zz, sharedZDataSize,
                                       sd, seedDataSize, partyId, l, leftEndpointsDataSize, r, rightEndpointsDataSize,
                                       bitWidthIn, bitWidthOut, elementSize, elementNum, c, cacheDataSize,
                                       s.stream_ptr());
    });
}

// -------------------------------------------------------------------- OTTT ----

extern "C" FAST_FSS_API int FastFss_cpu_otttKeyGen(void *key, size_t keyDataSize, const void *alpha,
                                                   size_t alphaDataSize, size_t bitWidthIn, size_t elementSize,
                                                   size_t elementNum)
{
    Staging s;
    void       *k = s.inout(key, keyDataSize); // keyGen XORs bit alpha into the caller's random key
    const void *a = s.in(alpha, alphaDataSize);
    return staged(s, [&] {
        return FastFss_cuda_otttKeyGen(k, keyDataSize, a, alphaDataSize, bitWidthIn, elementSize, elementNum,
                                       s.stream_ptr());
    });
}

extern "C" FAST_FSS_API int FastFss_cpu_otttLutEval(void *sharedOutE, size_t sharedOutEDataSize, void *sharedOutT,
                                                    size_t sharedOutTDataSize, const void *maskedX,
                                                    size_t maskedXDataSize, const void *key, size_t keyDataSize,
                                                    int partyId, const void *lookUpTable, size_t lookUpTableDataSize,
                                                    size_t bitWidthIn, size_t elementSize, size_t elementNum)
{
    Staging s;
    void       *oe = s.out(sharedOutE, sharedOutEDataSize);
    void       *ot = s.out(sharedOutT, sharedOutTDataSize);
    const void *x  = s.in(maskedX, maskedXDataSize);
    const void *k  = s.in(key, keyDataSize);
    const void *t  = s.in(lookUpTable, lookUpTableDataSize);
    return staged(s, [&] {
        return FastFss_cuda_otttLutEval(oe, sharedOutEDataSize, ot, sharedOutTDataSize, x, maskedXDataSize, k,
                                        keyDataSize, partyId, t, lookUpTableDataSize, bitWidthIn, elementSize,
                                        elementNum, s.stream_ptr());
    });
}

// ------------------------------------------------------------------ spline ----

extern "C" FAST_FSS_API int FastFss_cpu_dcfSplineKeyGen(
    void *key, size_t keyDataSize, void *e, size_t eDataSize, void *beta, size_t betaDataSize, const void *alpha,
    size_t alphaDataSize, const void *seed0, size_t seedDataSize0, const void *seed1, size_t seedDataSize1,
    const void *coefficients, size_t coefficientsDataSize, size_t degree, const void *leftEndpoints,
    size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize, size_t intervalNum,
    size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum)
{
    Staging s;
    void       *k  = s.out(key, keyDataSize);
    void       *ee = s.out(e, eDataSize);
    void       *b  = s.out(beta, betaDataSize);
    const void *a  = s.in(alpha, alphaDataSize);
    const void *d0 = s.in(seed0, seedDataSize0);
    const void *d1 = s.in(seed1, seedDataSize1);
    const void *cf = s.in(coefficients, coefficientsDataSize);
    const void *l  = s.in(leftEndpoints, leftEndpointsDataSize);
    const void *r  = s.in(rightEndpoints, rightEndpointsDataSize);
    return staged(s, [&] {
        return FastFss_cuda_dcfSplineKeyGen(k, keyDataSize, ee, eDataSize, b, betaDataSize, a, alphaDataSize, d0,
                                            seedDataSize0, d1, seedDataSize1, cf, coefficientsDataSize, degree, l,
                                            leftEndpointsDataSize, r, rightEndpointsDataSize, intervalNum, bitWidthIn,
                                            bitWidthOut, elementSize, elementNum, s.stream_ptr());
    });
}

extern "C" FAST_FSS_API int FastFss_cpu_dcfSplineEval(
    void *sharedOut, size_t sharedOutDataSize, const void *maskedX, size_t maskedXDataSize, const void *key,
    size_t keyDataSize, const void *sharedE, size_t sharedEDataSize, const void *sharedBeta,
    size_t sharedBetaDataSize, const void *seed, size_t seedDataSize, int partyId, const void *leftEndpoints,
    size_t leftEndpointsDataSize, const void *rightEndpoints, size_t rightEndpointsDataSize, size_t intervalNum,
    size_t degree, size_t bitWidthIn, size_t bitWidthOut, size_t elementSize, size_t elementNum, void *cache,
    size_t cacheDataSize)
{
    Staging s;
    void       *o  = s.out(sharedOut, sharedOutDataSize);
    const void *x  = s.in(maskedX, maskedXDataSize);
    const void *k  = s.in(key, keyDataSize);
    const void *ee = s.in(sharedE, sharedEDataSize);
    const void *b  = s.in(sharedBeta, sharedBetaDataSize);
    const void *sd = s.in(seed, seedDataSize);
    const void *l  = s.in(leftEndpoints, leftEndpointsDataSize);
    const void *r  = s.in(rightEndpoints, rightEndpointsDataSize);
    void       *c  = s.scratch(cache, cacheDataSize);
    return staged(s, [&] {
        return FastFss_cuda_dcfSplineEval(o, sharedOutDataSize, x, maskedXDataSize, k, keyDataSize, ee,
                                          sharedEDataSize, b, sharedBetaDataSize, sd, seedDataSize, partyId, l,
                                          leftEndpointsDataSize, r, rightEndpointsDataSize, intervalNum, degree,
                                          bitWidthIn, bitWidthOut, elementSize, elementNum, c, cacheDataSize,
                                          s.stream_ptr());
    });
}

// -------------------------------------------------------------------- PRNG ----

extern "C" FAST_FSS_API void *FastFss_cpu_prngInit(void)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    void *p = FastFss_cuda_prngInit();
    if (p == nullptr) return nullptr;
    return new (std::nothrow) HostPrng{p, dev};
}

extern "C" FAST_FSS_API void FastFss_cpu_prngRelease(void *prng)
{
    if (prng == nullptr) return;
    HostPrng *h = (HostPrng *)prng;
    {
        DeviceScope scope(h->device);
        FastFss_cuda_prngRelease(h->dev_prng);
    }
    delete h;
}

extern "C" FAST_FSS_API int FastFss_cpu_prngSetCurrentSeed(void *prng, const void *seed128bit,
                                                           const void *counter128bit)
{
    if (prng == nullptr || seed128bit == nullptr) return FAST_FSS_PRNG_INPUT_INVALID_ARGUMENT;
    HostPrng   *h = (HostPrng *)prng;
    DeviceScope scope(h->device);
    if (!scope.ok()) return FAST_FSS_RUNTIME_ERROR;
    return FastFss_cuda_prngSetCurrentSeed(h->dev_prng, seed128bit, counter128bit);
}

extern "C" FAST_FSS_API int FastFss_cpu_prngGetCurrentSeed(const void *prng, void *seed128bit, void *counter128bit)
{
    if (prng == nullptr || seed128bit == nullptr) return FAST_FSS_PRNG_INPUT_INVALID_ARGUMENT;
    const HostPrng *h = (const HostPrng *)prng;
    DeviceScope     scope(h->device);
    if (!scope.ok()) return FAST_FSS_RUNTIME_ERROR;
    return FastFss_cuda_prngGetCurrentSeed(h->dev_prng, seed128bit, counter128bit);
}

extern "C" FAST_FSS_API int FastFss_cpu_prngGen(void *prng, void *dst, size_t bitWidth, size_t elementSize,
                                                size_t elementNum)
{
    // the reference's checks, in its order (src/cpu/prng.cpp:134-147)
    if (prng == nullptr || dst == nullptr) return FAST_FSS_PRNG_INPUT_INVALID_ARGUMENT;
    if (bitWidth == 0 || bitWidth > elementSize * 8) return FAST_FSS_PRNG_INVALID_BIT_WIDTH;
    HostPrng   *h = (HostPrng *)prng;
    DeviceScope scope(h->device);
    if (!scope.ok()) return FAST_FSS_RUNTIME_ERROR;
    const size_t bytes = elementSize * elementNum;
    Staging      s;
    void        *d = s.out(dst, bytes);
    if (bytes == 0) // nothing to fill; the counter does not move either
        return s.finish(FAST_FSS_SUCCESS);
    return staged(s, [&] {
        return FastFss_cuda_prngGen(h->dev_prng, d, bitWidth, elementSize, elementNum, s.stream_ptr());
    });
}
