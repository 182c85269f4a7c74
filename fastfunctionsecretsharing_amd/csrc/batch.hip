// Ragged batched DCF / DPF evaluation: one persistent launch over many key
// groups of mixed element widths (no counterpart in the reference; each
// group's shares equal FastFss_cuda_dcfEval / FastFss_cuda_dpfEval).
//
// Scheduling: one 1024-thread workgroup per CU; each wave repeatedly takes
// the next tile from a device-wide atomic counter.  Tiles are numbered group
// by group (groups sorted by descending tree depth so the long walks start
// first); a wave finds its group by binary search over the groups' first
// tile indices, which live in a small table read through the scalar cache.
// Within a tile lane L walks elements L, L+64, ...
//
// The walk is width-generic: element loads, stores and the 64-bit
// arithmetic are selected at run time by the group's (wave-uniform) element
// size, so one kernel body serves 8..64-bit groups.  Arithmetic is carried in
// 64 bits and truncated to the element width on store, which is the same
// ring arithmetic as the per-width kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "FastFss/hip/batch.h"
#include "dpf_device.h"
#include "fss_host.h"
#include "fss_launch.h"

namespace fss {

struct BatchGroupDev
{
    const uint8_t *key;
    const uint8_t *seed;
    const uint8_t *x;
    uint8_t       *out;
    uint64_t       N;
    uint32_t       tileStart;
    uint32_t       tileElems;
    int32_t        op, es, n, m, party, pad;
};

__device__ __forceinline__ uint64_t ld_ge(const uint8_t *p, int es, size_t i)
{
    switch (es)
    {
        case 1: return p[i];
        case 2: return ((const uint16_t *)p)[i];
        case 4: return ((const uint32_t *)p)[i];
        default: return ((const uint64_t *)p)[i];
    }
}

__device__ __forceinline__ void st_ge(uint8_t *p, int es, size_t i, uint64_t v)
{
    switch (es)
    {
        case 1: p[i] = (uint8_t)v; break;
        case 2: ((uint16_t *)p)[i] = (uint16_t)v; break;
        case 4: ((uint32_t *)p)[i] = (uint32_t)v; break;
        default: ((uint64_t *)p)[i] = v; break;
    }
}

__device__ __forceinline__ uint64_t mask_bits(uint64_t x, int bw)
{
    return bw >= 64 ? x : (x & ((1ull << bw) - 1));
}

constexpr int kBatchCH = 8; // levels per correction-word burst

// DCF share, groupSize 1 (reference src/impl/dcf.h:370-491), width at run time.
__device__ __forceinline__ uint64_t dcf_walk_rt(const AesCtx &c, const uint8_t *sCW, const uint8_t *tail, int es,
                                                uint64_t x, uint32_t s[4], int party, int n, int outBytes)
{
    s[0] &= ~1u;
    uint32_t       t   = (uint32_t)party;
    uint64_t       acc = 0;
    const uint64_t tl = ld_ge(tail, es, n), tr = ld_ge(tail, es, n + 1);
    x = mask_bits(x, n);
    for (int i0 = 0; i0 < n; i0 += kBatchCH)
    {
        uint32_t cw[kBatchCH][4];
        uint64_t vcw[kBatchCH];
#pragma unroll
        for (int k = 0; k < kBatchCH; k++)
            if (i0 + k < n)
            {
                load16(cw[k], sCW + 16 * (i0 + k));
                vcw[k] = ld_ge(tail, es, i0 + k);
            }
#pragma unroll
        for (int k = 0; k < kBatchCH; k++)
        {
            const int i = i0 + k;
            if (i < n)
            {
                const uint32_t xi = (uint32_t)(x >> (n - 1 - i)) & 1u;
                uint32_t       blk[2][4];
                aes_encrypt_dcf_pair(c, s, xi, blk);
                const uint32_t mk = 0u - t;
                uint32_t       tn = blk[0][0] & 1u;
                s[0]              = (blk[0][0] & ~1u) ^ (cw[k][0] & mk);
                s[1]              = blk[0][1] ^ (cw[k][1] & mk);
                s[2]              = blk[0][2] ^ (cw[k][2] & mk);
                s[3]              = blk[0][3] ^ (cw[k][3] & mk);
                tn ^= (uint32_t)(((xi ? tr : tl) >> i) & 1) & t;
                const uint64_t term = low_bytes(blk[1], outBytes) + (t ? vcw[k] : 0ull);
                acc                 = party ? acc - term : acc + term;
                t                   = tn;
            }
        }
    }
    const uint64_t term = low_bytes(s, outBytes) + (t ? ld_ge(tail, es, n + 2) : 0ull);
    return party ? acc - term : acc + term;
}

// DPF share, groupSize 1 (reference src/impl/dpf.h:290-380).  The reference
// converts with (bitWidthOut, groupSize) in the (groupSize, bitWidth) slots:
// one byte per element, read from the seed when bitWidthOut <= 16 and from
// AES_seed(0, 0) otherwise.
__device__ __forceinline__ uint64_t dpf_walk_rt(const AesCtx &c, const uint8_t *sCW, const uint8_t *tail, int es,
                                                uint64_t x, uint32_t s[4], int party, int n, int m)
{
    s[0] &= ~1u;
    uint32_t       t  = (uint32_t)party;
    const uint64_t tl = ld_ge(tail, es, 0), tr = ld_ge(tail, es, 1);
    x = mask_bits(x, n);
    for (int i0 = 0; i0 < n; i0 += kBatchCH)
    {
        uint32_t cw[kBatchCH][4];
#pragma unroll
        for (int k = 0; k < kBatchCH; k++)
            if (i0 + k < n) load16(cw[k], sCW + 16 * (i0 + k));
#pragma unroll
        for (int k = 0; k < kBatchCH; k++)
        {
            const int i = i0 + k;
            if (i < n)
            {
                const uint32_t bit    = (uint32_t)(x >> (n - 1 - i)) & 1u;
                uint32_t       blk[1][4] = {{2u * bit, 0u, 2u * bit + 1u, 0u}};
                aes_encrypt<1>(c, s, blk);
                const uint32_t mk = 0u - t;
                uint32_t       tn = blk[0][0] & 1u;
                s[0]              = (blk[0][0] & ~1u) ^ (cw[k][0] & mk);
                s[1]              = blk[0][1] ^ (cw[k][1] & mk);
                s[2]              = blk[0][2] ^ (cw[k][2] & mk);
                s[3]              = blk[0][3] ^ (cw[k][3] & mk);
                tn ^= (uint32_t)(((bit ? tr : tl) >> i) & 1) & t;
                t = tn;
            }
        }
    }
    uint32_t conv;
    if (m <= 16)
        conv = s[0] & 0xffu;
    else
    {
        uint32_t blk[1][4] = {{0u, 0u, 0u, 0u}};
        aes_encrypt<1>(c, s, blk);
        conv = blk[0][0] & 0xffu;
    }
    const uint64_t term = (uint64_t)conv + (t ? ld_ge(tail, es, 2) : 0ull);
    return party ? 0 - term : term;
}

__global__ void batch_reset_kernel(uint32_t *counter)
{
    *counter = 0;
}

__global__ __launch_bounds__(kBlock) void batch_eval_kernel(const BatchGroupDev *__restrict__ gs, int G,
                                                            uint32_t totalTiles, uint32_t *counter)
{
    __shared__ uint32_t lds[kLdsTableWords];
    aes_fill_lds(lds);
    __syncthreads();
    const AesCtx   c    = aes_ctx(lds);
    const uint32_t lane = threadIdx.x & 63u;
    for (;;)
    {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(counter, 1u);
        tile = __builtin_amdgcn_readfirstlane(tile);
        if (tile >= totalTiles) break;
        int lo = 0, hi = G - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (gs[mid].tileStart <= tile)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BatchGroupDev g    = gs[lo];
        const uint64_t      base = (uint64_t)(tile - g.tileStart) * g.tileElems;
        const uint64_t      end  = (base + g.tileElems < g.N) ? base + g.tileElems : g.N;
        const size_t        sw   = 16 * (size_t)g.n;
        const uint8_t      *tails = g.key + g.N * sw;
        for (uint64_t e = base + lane; e < end; e += 64)
        {
            uint32_t s[4];
            load16(s, g.seed + 16 * e);
            const uint64_t xv = ld_ge(g.x, g.es, e);
            uint64_t       v;
            if (g.op == FAST_FSS_BATCH_DCF_EVAL)
                v = dcf_walk_rt(c, g.key + e * sw, tails + e * (size_t)g.es * (g.n + 3), g.es, xv, s, g.party, g.n,
                                (g.m + 7) / 8);
            else
                v = dpf_walk_rt(c, g.key + e * sw, tails + e * (size_t)g.es * 3, g.es, xv, s, g.party, g.n, g.m);
            st_ge(g.out, g.es, e, v);
        }
    }
}

struct BatchPlan
{
    BatchGroupDev  *groups  = nullptr; // device
    uint32_t       *counter = nullptr; // device
    int             G       = 0;
    uint32_t        tiles   = 0;
    int             device  = 0;
    hipGraph_t      graph   = nullptr;
    hipGraphExec_t  exec    = nullptr;
};

static std::mutex g_tile_mu;
// default elements per tile, [op][elementSize]
static uint32_t g_tile[2][9] = {{0, 256, 256, 0, 128, 0, 0, 0, 64}, {0, 512, 256, 0, 256, 0, 0, 0, 128}};
static std::atomic<int> g_blocks_per_cu{1};

static int validate(const FastFssBatchGroup &q)
{
    const size_t n = q.bitWidthIn, m = q.bitWidthOut, es = q.elementSize, N = q.elementNum;
    if (q.op != FAST_FSS_BATCH_DCF_EVAL && q.op != FAST_FSS_BATCH_DPF_EVAL) return FAST_FSS_RUNTIME_ERROR;
    if (!elem_size_narrow(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(0 < n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!(0 < m && m <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (q.partyId != 0 && q.partyId != 1) return FAST_FSS_INVALID_PARTY_ID_ERROR;
    const size_t need = q.op == FAST_FSS_BATCH_DCF_EVAL ? dcf_key_size(n, 1, es, N) : dpf_key_size(n, 1, es, N);
    if (q.keyDataSize != need) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (q.tileElems % 64 != 0) return FAST_FSS_RUNTIME_ERROR;
    if (N && (!q.sharedOut || !q.maskedX || !q.key || !q.seed)) return FAST_FSS_RUNTIME_ERROR;
    if (N && (!aligned(q.key, 8) || !aligned(q.seed, 8))) return FAST_FSS_RUNTIME_ERROR;
    return FAST_FSS_SUCCESS;
}

static void plan_free(BatchPlan *p)
{
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->groups) (void)hipFree(p->groups);
    if (p->counter) (void)hipFree(p->counter);
    delete p;
}

static int plan_enqueue(BatchPlan *p, hipStream_t st)
{
    batch_reset_kernel<<<1, 1, 0, st>>>(p->counter);
    if (p->tiles == 0) return launch_status();
    const int cus  = device_cu_count();
    size_t    grid = (size_t)cus * (size_t)g_blocks_per_cu.load(std::memory_order_relaxed);
    // never launch more workgroups than 16 tiles each can keep busy
    const size_t need = (p->tiles + 15) / 16;
    if (grid > need) grid = need < 1 ? 1 : need;
    batch_eval_kernel<<<(unsigned)grid, kBlock, 0, st>>>(p->groups, p->G, p->tiles, p->counter);
    return launch_status();
}

} // namespace fss

using namespace fss;

extern "C" FAST_FSS_API int FastFss_hip_batchSetDefaultTile(int op, size_t elementSize, uint32_t tileElems)
{
    if ((op != 0 && op != 1) || !elem_size_narrow(elementSize) || tileElems == 0 || tileElems % 64 != 0)
        return FAST_FSS_RUNTIME_ERROR;
    std::lock_guard<std::mutex> lock(g_tile_mu);
    g_tile[op][elementSize] = tileElems;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API uint32_t FastFss_hip_batchGetDefaultTile(int op, size_t elementSize)
{
    if ((op != 0 && op != 1) || !elem_size_narrow(elementSize)) return 0;
    std::lock_guard<std::mutex> lock(g_tile_mu);
    return g_tile[op][elementSize];
}

extern "C" FAST_FSS_API int FastFss_hip_batchSetBlocksPerCu(int blocks)
{
    if (blocks < 1 || blocks > 8) return FAST_FSS_RUNTIME_ERROR;
    g_blocks_per_cu.store(blocks, std::memory_order_relaxed);
    return FAST_FSS_SUCCESS;
}

namespace fss {

// Validate, order and number the groups' tiles (host side of a plan).
static int build_table(const FastFssBatchGroup *groups, size_t groupNum, std::vector<BatchGroupDev> &host,
                       uint32_t &tilesOut)
{
    if (groupNum && !groups) return FAST_FSS_RUNTIME_ERROR;
    if (groupNum > (1u << 30)) return FAST_FSS_RUNTIME_ERROR;
    for (size_t i = 0; i < groupNum; i++)
    {
        int r = validate(groups[i]);
        if (r) return r;
    }
    // long walks first: sort by tree depth, DCF (two blocks a level) before DPF
    std::vector<size_t> order(groupNum);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const size_t ca = groups[a].bitWidthIn * (groups[a].op == FAST_FSS_BATCH_DCF_EVAL ? 2 : 1);
        const size_t cb = groups[b].bitWidthIn * (groups[b].op == FAST_FSS_BATCH_DCF_EVAL ? 2 : 1);
        return ca > cb;
    });
    host.clear();
    host.reserve(groupNum);
    uint64_t                    tiles = 0;
    std::lock_guard<std::mutex> lock(g_tile_mu);
    for (size_t k = 0; k < groupNum; k++)
    {
        const FastFssBatchGroup &q = groups[order[k]];
        if (q.elementNum == 0) continue;
        BatchGroupDev d{};
        d.key       = (const uint8_t *)q.key;
        d.seed      = (const uint8_t *)q.seed;
        d.x         = (const uint8_t *)q.maskedX;
        d.out       = (uint8_t *)q.sharedOut;
        d.N         = q.elementNum;
        d.tileElems = q.tileElems ? q.tileElems : g_tile[q.op][q.elementSize];
        d.tileStart = (uint32_t)tiles;
        d.op = q.op, d.es = (int)q.elementSize, d.n = (int)q.bitWidthIn, d.m = (int)q.bitWidthOut;
        d.party = q.partyId;
        tiles += (q.elementNum + d.tileElems - 1) / d.tileElems;
        if (tiles >= 0xFFFFFFFFull) return FAST_FSS_RUNTIME_ERROR;
        host.push_back(d);
    }
    tilesOut = (uint32_t)tiles;
    return FAST_FSS_SUCCESS;
}

size_t batch_table_entry_bytes()
{
    return sizeof(BatchGroupDev);
}

int batch_enqueue_owned(const FastFssBatchGroup *groups, size_t groupNum, void *hostTab, void *devTab,
                        uint32_t *devCounter, size_t cap, hipStream_t st)
{
    std::vector<BatchGroupDev> host;
    uint32_t                   tiles = 0;
    int                        rc    = build_table(groups, groupNum, host, tiles);
    if (rc) return rc;
    if (host.size() > cap || !hostTab || !devTab || !devCounter) return FAST_FSS_RUNTIME_ERROR;
    if (tiles == 0) return FAST_FSS_SUCCESS;
    const size_t bytes = sizeof(BatchGroupDev) * host.size();
    std::memcpy(hostTab, host.data(), bytes);
    // pinned source: the copy is asynchronous; the caller reuses hostTab only
    // after work recorded behind this launch has completed
    if (hipMemcpyAsync(devTab, hostTab, bytes, hipMemcpyHostToDevice, st) != hipSuccess)
        return FAST_FSS_RUNTIME_ERROR;
    BatchPlan p;
    p.groups  = (BatchGroupDev *)devTab;
    p.counter = devCounter;
    p.G       = (int)host.size();
    p.tiles   = tiles;
    return plan_enqueue(&p, st);
}

} // namespace fss

extern "C" FAST_FSS_API int FastFss_hip_batchPlanCreate(void **plan, const FastFssBatchGroup *groups, size_t groupNum)
{
    if (!plan) return FAST_FSS_RUNTIME_ERROR;
    std::vector<BatchGroupDev> host;
    uint32_t                   tiles = 0;
    int                        rc    = build_table(groups, groupNum, host, tiles);
    if (rc) return rc;
    BatchPlan *p = new BatchPlan();
    p->G         = (int)host.size();
    p->tiles     = tiles;
    if (hipGetDevice(&p->device) != hipSuccess ||
        hipMalloc(&p->groups, sizeof(BatchGroupDev) * (host.size() ? host.size() : 1)) != hipSuccess ||
        hipMalloc(&p->counter, sizeof(uint32_t)) != hipSuccess ||
        (host.size() &&
         hipMemcpy(p->groups, host.data(), sizeof(BatchGroupDev) * host.size(), hipMemcpyHostToDevice) != hipSuccess))
    {
        plan_free(p);
        return FAST_FSS_RUNTIME_ERROR;
    }
    *plan = p;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_hip_batchPlanLaunch(void *plan, void *streamPtr)
{
    if (!plan) return FAST_FSS_RUNTIME_ERROR;
    return plan_enqueue((BatchPlan *)plan, resolve_stream(streamPtr));
}

extern "C" FAST_FSS_API int FastFss_hip_batchPlanCapture(void *plan, void *streamPtr)
{
    BatchPlan *p = (BatchPlan *)plan;
    if (!p) return FAST_FSS_RUNTIME_ERROR;
    if (p->exec) return FAST_FSS_SUCCESS;
    hipStream_t st = resolve_stream(streamPtr);
    if (st == hipStreamPerThread || st == nullptr) return FAST_FSS_RUNTIME_ERROR; // capture needs a real stream
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return FAST_FSS_RUNTIME_ERROR;
    int rc = plan_enqueue(p, st);
    if (hipStreamEndCapture(st, &p->graph) != hipSuccess || rc != FAST_FSS_SUCCESS) return FAST_FSS_RUNTIME_ERROR;
    if (hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess) return FAST_FSS_RUNTIME_ERROR;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_hip_batchPlanReplay(void *plan, void *streamPtr)
{
    BatchPlan *p = (BatchPlan *)plan;
    if (!p || !p->exec) return FAST_FSS_RUNTIME_ERROR;
    return hipGraphLaunch(p->exec, resolve_stream(streamPtr)) == hipSuccess ? FAST_FSS_SUCCESS
                                                                             : FAST_FSS_RUNTIME_ERROR;
}

extern "C" FAST_FSS_API size_t FastFss_hip_batchPlanTileCount(const void *plan)
{
    return plan ? ((const BatchPlan *)plan)->tiles : 0;
}

extern "C" FAST_FSS_API void FastFss_hip_batchPlanDestroy(void *plan)
{
    if (!plan) return;
    (void)hipDeviceSynchronize();
    plan_free((BatchPlan *)plan);
}

extern "C" FAST_FSS_API int FastFss_hip_batchEval(const FastFssBatchGroup *groups, size_t groupNum, void *streamPtr)
{
    std::vector<BatchGroupDev> host;
    uint32_t                   tiles = 0;
    int                        rc    = build_table(groups, groupNum, host, tiles);
    if (rc || tiles == 0) return rc;
    hipStream_t st = resolve_stream(streamPtr);
    BatchPlan   p;
    p.G     = (int)host.size();
    p.tiles = tiles;
    // stream-ordered tables: allocated, filled, used and released in order on `st`
    if (hipMallocAsync((void **)&p.groups, sizeof(BatchGroupDev) * host.size(), st) != hipSuccess)
        return FAST_FSS_RUNTIME_ERROR;
    if (hipMallocAsync((void **)&p.counter, sizeof(uint32_t), st) != hipSuccess)
    {
        (void)hipFreeAsync(p.groups, st);
        return FAST_FSS_RUNTIME_ERROR;
    }
    rc = FAST_FSS_SUCCESS;
    if (hipMemcpyAsync(p.groups, host.data(), sizeof(BatchGroupDev) * host.size(), hipMemcpyHostToDevice, st) !=
        hipSuccess)
        rc = FAST_FSS_RUNTIME_ERROR;
    // the pageable source must outlive the copy
    if (rc == FAST_FSS_SUCCESS && hipStreamSynchronize(st) != hipSuccess) rc = FAST_FSS_RUNTIME_ERROR;
    if (rc == FAST_FSS_SUCCESS) rc = plan_enqueue(&p, st);
    (void)hipFreeAsync(p.groups, st);
    (void)hipFreeAsync(p.counter, st);
    p.groups = nullptr, p.counter = nullptr;
    return rc;
}
