// LDS read-throughput probe: the ceiling of the LDS table lookups that bind
// the T-table AES kernels.
//
// A 128-KB LDS table holds rows `pitch` bytes apart; a row holds one W-byte
// slot per lane.  Each lane starts at a random group of K rows; per iteration
// it reads its slot in each of the group's K rows (independent of each other
// and of the data, so K are in flight per wave), then advances by one group,
// wrapping at the last whole group.  The values are folded with XORs and
// written once, so no read is dead.  Every address is g * K * pitch +
// i * pitch + slot with g < groups = 128 KB / (K * pitch), i < K and
// slot + W <= pitch, so no lane reads outside the table.  K is 32 (an AES
// round of two interleaved blocks issues 32 lookups) unless noted.
//
// With OPAQUE each read gets its own address register, built by one VALU add
// that the compiler cannot see through (an empty asm), as each AES table
// lookup builds its address with one v_perm; the compiler then cannot merge
// two reads into one ds_read2 instruction.  Without it the reads share one
// address register at immediate offsets.
//
// Layouts:
//   0  W = 4, 256-byte rows, slot 4 * (L mod 32): lanes L and L + 32 share a
//      bank in different rows, the bank pattern of the replicated AES tables
//   1  W = 4, 256-byte rows, slot 4 * L: 64 distinct banks
//   2  W = 4, rows 1028 bytes apart (not 256-byte aligned), one address
//      register, K = 8: the cost of rows that straddle 256-byte lines
//   3  W = 8, 512-byte rows (ds_read_b64)
//   4  W = 16, 1024-byte rows (ds_read_b128)
//   5  W = 4, 256-byte rows, one address register: the compiler merges the
//      reads two per ds_read2st64_b32
//   6  as 0 with K = 8: fewer reads in flight per wave
//   7  W = 4, 256-byte rows, slot 4 * L, K = 32: one asm block of 32
//      ds_read_b32 at immediate offsets from one address register, which
//      waits for all of them before it ends (its outputs are early-clobber
//      and all consumed), so no VALU work sits between the reads
//   8  as 0, software-pipelined: the reads of group k + 1 are issued before
//      the values of group k are folded, so a wave never drains its reads
//      (K = 16 per group, 32 in flight)
//   9  as 8 without OPAQUE: the compiler merges the reads two per
//      ds_read2st64_b32
//
// fss_lds_probe() runs a layout on `blocks` workgroups of 1024 lanes (one per
// CU at this LDS size) and returns the kernel time and the number of 4-byte
// words its reads delivered.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr uint32_t kBytes = 128 * 1024;

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int W> struct Vec;
template <> struct Vec<4> { using type = uint32_t; };
template <> struct Vec<8> { using type = u32x2; };
template <> struct Vec<16> { using type = u32x4; };

__device__ inline uint32_t fold(uint32_t v) { return v; }
__device__ inline uint32_t fold(u32x2 v) { return v.x ^ v.y; }
__device__ inline uint32_t fold(u32x4 v) { return v.x ^ v.y ^ v.z ^ v.w; }

template <int W, uint32_t PITCH, bool HALF, bool OPAQUE, int K = 32>
__global__ __launch_bounds__(1024) void lds_probe_kernel(uint32_t *out, int iters, uint32_t seed)
{
    using V = typename Vec<W>::type;
    static_assert(64 * W <= PITCH && PITCH % W == 0, "a row holds one slot per lane");
    __shared__ __align__(16) uint32_t lds[kBytes / 4];
    for (uint32_t i = threadIdx.x; i < kBytes / 4; i += blockDim.x) lds[i] = i * 2654435761u ^ seed;
    __syncthreads();
    constexpr uint32_t groups = kBytes / (K * PITCH);
    constexpr uint32_t span   = groups * K * PITCH;
    static_assert(groups >= 1, "a group of K rows fits the table");
    const uint32_t     L      = threadIdx.x & 63;
    const uint32_t     slot   = HALF ? 4 * (L & 31) : W * L;
    const uint32_t     g0     = (threadIdx.x * 2654435761u ^ seed) % groups;
    uint32_t           base   = g0 * K * PITCH + slot;
    uint32_t           acc    = 0;
    for (int it = 0; it < iters; it++)
    {
        V v[K];
        if (OPAQUE)
        {
            uint32_t a[K];
#pragma unroll
            for (int i = 0; i < K; i++)
            {
                a[i] = base + i * PITCH;
                asm volatile("" : "+v"(a[i]));
            }
#pragma unroll
            for (int i = 0; i < K; i++) v[i] = *(const V *)((const char *)lds + a[i]);
        }
        else
        {
            const V *p = (const V *)((const char *)lds + base);
#pragma unroll
            for (int i = 0; i < K; i++) v[i] = p[i * (PITCH / W)];
        }
#pragma unroll
        for (int i = 0; i < K; i++) acc ^= fold(v[i]);
        base += K * PITCH;
        if (base >= span) base -= span;
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// Layouts 8 and 9: groups of K reads, group k + 1 issued before group k's
// values are folded (see above).  Same address bounds as lds_probe_kernel.
template <bool OPAQUE, int K = 16>
__global__ __launch_bounds__(1024) void lds_probe_pipelined_kernel(uint32_t *out, int iters, uint32_t seed)
{
    constexpr uint32_t PITCH = 256;
    __shared__ __align__(16) uint32_t lds[kBytes / 4];
    for (uint32_t i = threadIdx.x; i < kBytes / 4; i += blockDim.x) lds[i] = i * 2654435761u ^ seed;
    __syncthreads();
    constexpr uint32_t groups = kBytes / (K * PITCH);
    constexpr uint32_t span   = groups * K * PITCH;
    const uint32_t     L      = threadIdx.x & 63;
    const uint32_t     slot   = OPAQUE ? 4 * (L & 31) : 4 * L;
    const uint32_t     g0     = (threadIdx.x * 2654435761u ^ seed) % groups;
    uint32_t           base   = g0 * K * PITCH + slot;
    uint32_t           acc    = 0;
    auto               load   = [&](uint32_t v[K]) {
        if (OPAQUE)
        {
            uint32_t a[K];
#pragma unroll
            for (int i = 0; i < K; i++)
            {
                a[i] = base + i * PITCH;
                asm volatile("" : "+v"(a[i]));
            }
#pragma unroll
            for (int i = 0; i < K; i++) v[i] = *(const uint32_t *)((const char *)lds + a[i]);
        }
        else
        {
            const uint32_t *p = (const uint32_t *)((const char *)lds + base);
#pragma unroll
            for (int i = 0; i < K; i++) v[i] = p[i * (PITCH / 4)];
        }
        base += K * PITCH;
        if (base >= span) base -= span;
    };
    uint32_t prev[K];
    load(prev);
    for (int it = 0; it < iters; it++)
    {
        uint32_t cur[K];
        load(cur);
#pragma unroll
        for (int i = 0; i < K; i++) acc ^= prev[i], prev[i] = cur[i];
    }
#pragma unroll
    for (int i = 0; i < K; i++) acc ^= prev[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// Layout 7: the 32 reads of a group as one asm block (see above); the LDS
// address of the table comes from an address-space cast of its pointer.
__global__ __launch_bounds__(1024) void lds_probe_asm_kernel(uint32_t *out, int iters, uint32_t seed)
{
    __shared__ __align__(16) uint32_t lds[kBytes / 4];
    for (uint32_t i = threadIdx.x; i < kBytes / 4; i += blockDim.x) lds[i] = i * 2654435761u ^ seed;
    __syncthreads();
    typedef __attribute__((address_space(3))) uint32_t lds_u32;
    const uint32_t     table  = (uint32_t)(uintptr_t)(lds_u32 *)lds;
    constexpr uint32_t PITCH = 256, K = 32, groups = kBytes / (K * PITCH), span = groups * K * PITCH;
    const uint32_t     L    = threadIdx.x & 63;
    const uint32_t     g0   = (threadIdx.x * 2654435761u ^ seed) % groups;
    uint32_t           base = g0 * K * PITCH + 4 * L;
    uint32_t           acc  = 0;
    for (int it = 0; it < iters; it++)
    {
        const uint32_t a = table + base;
        uint32_t       r[32];
        asm volatile(
            "ds_read_b32 %0, %32 offset:0\n ds_read_b32 %1, %32 offset:256\n"
            "ds_read_b32 %2, %32 offset:512\n ds_read_b32 %3, %32 offset:768\n"
            "ds_read_b32 %4, %32 offset:1024\n ds_read_b32 %5, %32 offset:1280\n"
            "ds_read_b32 %6, %32 offset:1536\n ds_read_b32 %7, %32 offset:1792\n"
            "ds_read_b32 %8, %32 offset:2048\n ds_read_b32 %9, %32 offset:2304\n"
            "ds_read_b32 %10, %32 offset:2560\n ds_read_b32 %11, %32 offset:2816\n"
            "ds_read_b32 %12, %32 offset:3072\n ds_read_b32 %13, %32 offset:3328\n"
            "ds_read_b32 %14, %32 offset:3584\n ds_read_b32 %15, %32 offset:3840\n"
            "ds_read_b32 %16, %32 offset:4096\n ds_read_b32 %17, %32 offset:4352\n"
            "ds_read_b32 %18, %32 offset:4608\n ds_read_b32 %19, %32 offset:4864\n"
            "ds_read_b32 %20, %32 offset:5120\n ds_read_b32 %21, %32 offset:5376\n"
            "ds_read_b32 %22, %32 offset:5632\n ds_read_b32 %23, %32 offset:5888\n"
            "ds_read_b32 %24, %32 offset:6144\n ds_read_b32 %25, %32 offset:6400\n"
            "ds_read_b32 %26, %32 offset:6656\n ds_read_b32 %27, %32 offset:6912\n"
            "ds_read_b32 %28, %32 offset:7168\n ds_read_b32 %29, %32 offset:7424\n"
            "ds_read_b32 %30, %32 offset:7680\n ds_read_b32 %31, %32 offset:7936\n"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3]), "=&v"(r[4]), "=&v"(r[5]), "=&v"(r[6]),
              "=&v"(r[7]), "=&v"(r[8]), "=&v"(r[9]), "=&v"(r[10]), "=&v"(r[11]), "=&v"(r[12]), "=&v"(r[13]),
              "=&v"(r[14]), "=&v"(r[15]), "=&v"(r[16]), "=&v"(r[17]), "=&v"(r[18]), "=&v"(r[19]),
              "=&v"(r[20]), "=&v"(r[21]), "=&v"(r[22]), "=&v"(r[23]), "=&v"(r[24]), "=&v"(r[25]),
              "=&v"(r[26]), "=&v"(r[27]), "=&v"(r[28]), "=&v"(r[29]), "=&v"(r[30]), "=&v"(r[31])
            : "v"(a)
            : "memory");
#pragma unroll
        for (int i = 0; i < 32; i++) acc ^= r[i];
        base += K * PITCH;
        if (base >= span) base -= span;
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

} // namespace

extern "C" __attribute__((visibility("default"))) int fss_lds_probe(int layout, int blocks, int iters, double *ms,
                                                                    double *words)
{
    if (layout < 0 || layout > 9 || blocks <= 0 || iters <= 0) return -1;
    uint32_t *out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * 1024 * sizeof(uint32_t)) != hipSuccess) return -1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1;
    auto launch = [&](int it) {
        switch (layout)
        {
            case 0: lds_probe_kernel<4, 256, true, true><<<blocks, 1024>>>(out, it, 7u); break;
            case 1: lds_probe_kernel<4, 256, false, true><<<blocks, 1024>>>(out, it, 7u); break;
            case 2: lds_probe_kernel<4, 1028, false, false, 8><<<blocks, 1024>>>(out, it, 7u); break;
            case 3: lds_probe_kernel<8, 512, false, true><<<blocks, 1024>>>(out, it, 7u); break;
            case 4: lds_probe_kernel<16, 1024, false, false><<<blocks, 1024>>>(out, it, 7u); break;
            case 5: lds_probe_kernel<4, 256, false, false><<<blocks, 1024>>>(out, it, 7u); break;
            case 7: lds_probe_asm_kernel<<<blocks, 1024>>>(out, it, 7u); break;
            case 8: lds_probe_pipelined_kernel<true><<<blocks, 1024>>>(out, it, 7u); break;
            case 9: lds_probe_pipelined_kernel<false><<<blocks, 1024>>>(out, it, 7u); break;
            default: lds_probe_kernel<4, 256, true, true, 8><<<blocks, 1024>>>(out, it, 7u); break;
        }
    };
    launch(64); // warm-up
    (void)hipEventRecord(e0, nullptr);
    launch(iters);
    (void)hipEventRecord(e1, nullptr);
    const int rc = (hipEventSynchronize(e1) == hipSuccess && hipGetLastError() == hipSuccess) ? 0 : -1;
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    const int w = layout == 3 ? 2 : (layout == 4 ? 4 : 1);
    *ms    = t;
    // reads per lane and iteration: 8 (layouts 2, 6), 16 (8, 9; the pipelined
    // kernels' one extra group is not counted), 32 (the others)
    const double perIter = (layout == 2 || layout == 6) ? 8.0 : (layout >= 8 ? 16.0 : 32.0);
    *words = (double)blocks * 1024.0 * (double)iters * perIter * w;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    return rc;
}
