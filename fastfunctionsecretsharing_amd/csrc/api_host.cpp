// Backend-independent part of the C ABI: size queries, the key zip/unzip
// stubs the reference also leaves unimplemented, launch configuration and the
// per-device CU count used to size persistent grids.
#include <mutex>

#include "FastFss/cuda/config.h"
#include "FastFss/dcf.h"
#include "FastFss/dpf.h"
#include "FastFss/grotto.h"
#include "FastFss/mic.h"
#include "FastFss/ottt.h"
#include "FastFss/spline.h"
#include "fss_host.h"

namespace fss {

int device_cu_count()
{
    static int        cached[64] = {0};
    static std::mutex mu;
    int               dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    std::lock_guard<std::mutex> lock(mu);
    if (cached[dev] == 0)
    {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cached[dev] = cus;
    }
    return cached[dev];
}

} // namespace fss

using namespace fss;

// ---- launch configuration (reference src/cuda/config.cpp:6-70) ----

static int        gGridDim   = 128;
static int        gThreshold = 64;
static std::mutex gConfigMutex;

extern "C" FAST_FSS_API int FastFss_cuda_setGridDim(int dim)
{
    std::lock_guard<std::mutex> lock(gConfigMutex);
    if (dim <= 0) return -2;
    gGridDim = dim;
    return 0;
}

extern "C" FAST_FSS_API int FastFss_cuda_getGridDim(void)
{
    std::lock_guard<std::mutex> lock(gConfigMutex);
    return gGridDim;
}

extern "C" FAST_FSS_API int FastFss_cuda_setFineGrainParallelGridDimThreshold(int dim)
{
    std::lock_guard<std::mutex> lock(gConfigMutex);
    if (dim <= 0) return -2;
    gThreshold = dim;
    return 0;
}

extern "C" FAST_FSS_API int FastFss_cuda_getFineGrainParallelGridDimThreshold(void)
{
    std::lock_guard<std::mutex> lock(gConfigMutex);
    return gThreshold;
}

// ---- DCF sizes (reference src/dcf.cpp:10-124) ----

extern "C" FAST_FSS_API int FastFss_dcfKeyZip(void *, size_t, const void *, size_t, size_t, size_t, size_t, size_t,
                                              size_t)
{
    return FAST_FSS_RUNTIME_ERROR;
}

extern "C" FAST_FSS_API int FastFss_dcfKeyUnzip(void *, size_t, const void *, size_t, size_t, size_t, size_t, size_t,
                                                size_t)
{
    return FAST_FSS_RUNTIME_ERROR;
}

static int dcf_widths(size_t n, size_t m, size_t es)
{
    if (!(0 < n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!(0 < m && m <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_dcfGetKeyDataSize(size_t *out, size_t n, size_t m, size_t g, size_t es, size_t N)
{
    int r = dcf_widths(n, m, es);
    if (r) return r;
    *out = dcf_key_size(n, g, es, N);
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_dcfGetZippedKeyDataSize(size_t *out, size_t n, size_t m, size_t g, size_t es,
                                                            size_t N)
{
    int r = dcf_widths(n, m, es);
    if (r) return r;
    *out = 0; // no compressed format exists (reference src/impl/dcf.h:52-59)
    (void)g, (void)N;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_dcfGetCacheDataSize(size_t *out, size_t n, size_t m, size_t g, size_t es,
                                                        size_t N)
{
    int r = dcf_widths(n, m, es);
    if (r) return r;
    *out = dcf_cache_size(n, g, es, N);
    return FAST_FSS_SUCCESS;
}

// ---- DPF sizes (reference src/dpf.cpp): only upper width bounds are checked ----

extern "C" FAST_FSS_API int FastFss_dpfKeyZip(void *, size_t, const void *, size_t, size_t, size_t, size_t, size_t,
                                              size_t)
{
    return FAST_FSS_RUNTIME_ERROR;
}

extern "C" FAST_FSS_API int FastFss_dpfKeyUnzip(void *, size_t, const void *, size_t, size_t, size_t, size_t, size_t,
                                                size_t)
{
    return FAST_FSS_RUNTIME_ERROR;
}

extern "C" FAST_FSS_API int FastFss_dpfGetKeyDataSize(size_t *out, size_t n, size_t m, size_t g, size_t es, size_t N)
{
    if (n > es * 8 || m > es * 8) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    *out = dpf_key_size(n, g, es, N);
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_dpfGetZippedKeyDataSize(size_t *out, size_t n, size_t m, size_t g, size_t es,
                                                            size_t N)
{
    if (n > es * 8 || m > es * 8) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    *out = 0;
    (void)g, (void)N;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_dpfGetCacheDataSize(size_t *out, size_t n, size_t es, size_t N)
{
    if (n > es * 8) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    *out = dpf_cache_size(n, N);
    return FAST_FSS_SUCCESS;
}

// ---- Grotto sizes (reference src/grotto.cpp): unsupported element sizes report 0 ----

extern "C" FAST_FSS_API int FastFss_grottoKeyZip(void *, size_t, const void *, size_t, size_t, size_t, size_t)
{
    return FAST_FSS_RUNTIME_ERROR;
}

extern "C" FAST_FSS_API int FastFss_grottoKeyUnzip(void *, size_t, const void *, size_t, size_t, size_t, size_t)
{
    return FAST_FSS_RUNTIME_ERROR;
}

extern "C" FAST_FSS_API int FastFss_grottoGetKeyDataSize(size_t *out, size_t n, size_t es, size_t N)
{
    if (!(6 <= n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    *out = elem_size_ok(es) ? grotto_key_size(n, es, N) : 0;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_grottoGetZippedKeyDataSize(size_t *out, size_t n, size_t es, size_t N)
{
    if (!(6 <= n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    // the reference reports (size_t)-1 here (src/impl/grotto.h:46-51)
    *out = elem_size_ok(es) ? (size_t)-1 : 0;
    (void)N;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_grottoGetCacheDataSize(size_t *out, size_t n, size_t es, size_t N)
{
    if (!(6 <= n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    *out = elem_size_ok(es) ? grotto_cache_size(n, N) : 0;
    return FAST_FSS_SUCCESS;
}

// ---- MIC sizes (reference src/mic.cpp): no width checks ----

extern "C" FAST_FSS_API int FastFss_dcfMICKeyZip(void *, size_t, const void *, size_t, size_t, size_t, size_t, size_t)
{
    return FAST_FSS_RUNTIME_ERROR;
}

extern "C" FAST_FSS_API int FastFss_dcfMICKeyUnzip(void *, size_t, const void *, size_t, size_t, size_t, size_t,
                                                   size_t)
{
    return FAST_FSS_RUNTIME_ERROR;
}

extern "C" FAST_FSS_API int FastFss_dcfMICGetCacheDataSize(size_t *out, size_t n, size_t m, size_t es, size_t N)
{
    *out = elem_size_ok(es) ? dcf_cache_size(n, 1, es, N) : 0;
    (void)m;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_dcfMICGetKeyDataSize(size_t *out, size_t n, size_t m, size_t es, size_t N)
{
    *out = elem_size_ok(es) ? dcf_key_size(n, 1, es, N) : 0;
    (void)m;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_dcfMICGetZippedKeyDataSize(size_t *out, size_t n, size_t m, size_t es, size_t N)
{
    *out = 0;
    (void)n, (void)m, (void)es, (void)N;
    return FAST_FSS_SUCCESS;
}

// ---- spline sizes (reference src/spline.cpp) ----

extern "C" FAST_FSS_API int FastFss_dcfSplineKeyZip(void *, size_t, const void *, size_t, size_t, size_t, size_t,
                                                    size_t, size_t, size_t)
{
    return FAST_FSS_RUNTIME_ERROR;
}

extern "C" FAST_FSS_API int FastFss_dcfSplineKeyUnzip(void *, size_t, const void *, size_t, size_t, size_t, size_t,
                                                      size_t, size_t, size_t)
{
    return FAST_FSS_RUNTIME_ERROR;
}

extern "C" FAST_FSS_API int FastFss_dcfSplineGetCacheDataSize(size_t *out, size_t d, size_t I, size_t n, size_t m,
                                                              size_t es, size_t N)
{
    *out = elem_size_ok(es) ? spline_cache_size(n, I, d, es, N) : 0;
    (void)m;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_dcfSplineGetKeyDataSize(size_t *out, size_t d, size_t I, size_t n, size_t m,
                                                            size_t es, size_t N)
{
    *out = elem_size_ok(es) ? dcf_key_size(n, I * (d + 1), es, N) : 0;
    (void)m;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_dcfSplineGetZippedKeyDataSize(size_t *out, size_t d, size_t I, size_t n, size_t m,
                                                                  size_t es, size_t N)
{
    *out = 0;
    (void)d, (void)I, (void)n, (void)m, (void)es, (void)N;
    return FAST_FSS_SUCCESS;
}

// ---- OTTT size (reference src/ottt.cpp:10-20) ----

extern "C" FAST_FSS_API int FastFss_otttGetKeyDataSize(size_t *out, size_t n, size_t N)
{
    if (!(3 <= n)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    *out = N * (1ULL << n) / 8;
    return FAST_FSS_SUCCESS;
}
