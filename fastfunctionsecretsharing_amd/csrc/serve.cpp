// Continuous-batching request server (see include/FastFss/hip/serve.h).
//
// One scheduler thread owns all device work.  Each pass it
//   1. queries the completion events of the launches in flight (no lock),
//   2. under the mutex: retires the requests of finished launches and admits
//      queued requests in arrival order under the in-flight and HBM-budget
//      limits,
//   3. without the mutex: launches what it admitted.  DCF and DPF requests
//      admitted in the same pass share one persistent batch launch (the
//      ragged-batch kernel); MIC and spline requests, and a lone DCF or DPF
//      request, get a launch each.  Each launch goes to the stream with the
//      fewest launches in flight and is followed by a completion event.
// Client threads only touch the queue and the request records, under the
// mutex, so a submit never waits behind a kernel launch.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "FastFss/hip/batch.h"
#include "FastFss/hip/serve.h"
#include "fss_host.h"
#include "fss_launch.h"

namespace fss {

using Clock = std::chrono::steady_clock;

struct ServeItem
{
    FastFssServeRequest req;
    uint64_t            id;
    size_t              bytes;
    Clock::time_point   tSubmit, tDone;
    int                 status = 1; // 1 queued, 2 running, 0 done, <0 failed
};

// One launch in flight: the requests it serves and its completion event.
struct Launch
{
    std::vector<ServeItem *> items;
    hipEvent_t               ev     = nullptr;
    int                      stream = 0;
    int                      slot   = -1; // fused-table slot, -1 for a single request
};

// Tables of one fused batch launch: the group table in pinned host memory,
// its device copy and the work-queue counter.
struct FusedSlot
{
    void     *host    = nullptr;
    void     *dev     = nullptr;
    uint32_t *counter = nullptr;
    bool      busy    = false;
};

struct Server
{
    // shared with client threads, under mu
    std::mutex                                               mu;
    std::condition_variable                                  cvWork, cvDone;
    std::deque<ServeItem *>                                  queue;
    std::unordered_map<uint64_t, std::unique_ptr<ServeItem>> items;
    size_t                                                   inflightBytes = 0;
    int                                                      inflightItems = 0;
    uint64_t                                                 nextId        = 1;
    bool                                                     stop          = false;
    FastFssServeStats                                        stats{};
    Clock::time_point                                        busySince;
    // set at creation
    size_t budget;
    int    maxInflight, launchMode, device = 0;
    // owned by the scheduler thread
    std::vector<hipStream_t> streams;
    std::vector<int>         streamLoad;
    int                      nextStream = 0;
    std::vector<Launch>      running;
    std::vector<FusedSlot>   slots;
    std::vector<hipEvent_t>  eventPool;
    uint64_t                 newLaunches = 0, newFused = 0; // added to stats at the next locked step
    std::thread              th;
};

static size_t request_bytes(const FastFssServeRequest &r)
{
    const size_t es = r.elementSize, N = r.elementNum;
    size_t       b  = r.keyDataSize + 16 * N + es * N;
    if (r.op == FAST_FSS_SERVE_MIC_EVAL)
        b += 2 * es * N * r.intervalNum;
    else if (r.op == FAST_FSS_SERVE_SPLINE_EVAL)
        b += es * N + 2 * es * N * r.intervalNum * (r.degree + 1);
    else
        b += es * N;
    return b;
}

static int validate(const FastFssServeRequest &r)
{
    const size_t n = r.bitWidthIn, m = r.bitWidthOut, es = r.elementSize, N = r.elementNum;
    if (!elem_size_narrow(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
    if (!(0 < n && n <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (!(0 < m && m <= es * 8)) return FAST_FSS_INVALID_BITWIDTH_ERROR;
    if (r.partyId != 0 && r.partyId != 1) return FAST_FSS_INVALID_PARTY_ID_ERROR;
    size_t need = 0;
    switch (r.op)
    {
        case FAST_FSS_SERVE_DCF_EVAL: need = dcf_key_size(n, 1, es, N); break;
        case FAST_FSS_SERVE_DPF_EVAL: need = dpf_key_size(n, 1, es, N); break;
        case FAST_FSS_SERVE_MIC_EVAL:
            if (r.intervalNum == 0 || !r.sharedZ || !r.leftEndpoints || !r.rightEndpoints)
                return FAST_FSS_INVALID_BOUNDARY_DATA_SIZE_ERROR;
            need = dcf_key_size(n, 1, es, N);
            break;
        case FAST_FSS_SERVE_SPLINE_EVAL:
            if (r.intervalNum == 0 || !r.sharedZ || !r.sharedBeta || !r.leftEndpoints || !r.rightEndpoints)
                return FAST_FSS_INVALID_BOUNDARY_DATA_SIZE_ERROR;
            if (r.degree > 7) return FAST_FSS_RUNTIME_ERROR;
            need = dcf_key_size(n, r.intervalNum * (r.degree + 1), es, N);
            break;
        default: return FAST_FSS_RUNTIME_ERROR;
    }
    if (r.keyDataSize != need) return FAST_FSS_INVALID_KEY_DATA_SIZE_ERROR;
    if (N && (!r.sharedOut || !r.maskedX || !r.key || !r.seed)) return FAST_FSS_RUNTIME_ERROR;
    if (N && (!aligned(r.key, 8) || !aligned(r.seed, 8))) return FAST_FSS_RUNTIME_ERROR;
    return FAST_FSS_SUCCESS;
}

static int launch(Server *s, ServeItem *it, hipStream_t st)
{
    const FastFssServeRequest &r = it->req;
    if (r.elementNum == 0) return FAST_FSS_SUCCESS;
    size_t work = r.elementNum;
    if (r.op == FAST_FSS_SERVE_MIC_EVAL) work *= r.intervalNum > 4 ? 4 : r.intervalNum;
    // mode 0 gives a small spline call the per-call entry point's interval split
    if (r.op == FAST_FSS_SERVE_SPLINE_EVAL && s->launchMode == 0) work *= spline_blocks_per_elem(work, r.intervalNum);
    // mode 1: one full workgroup per 1024 elements; mode 0: the per-call shape
    LaunchShape ls = s->launchMode == 1 ? LaunchShape{(int)((work + kBlock - 1) / kBlock), kBlock} : shape_for(work);
    if (ls.grid > 65535) ls.grid = 65535;
    if (ls.grid < 1) ls.grid = 1;
    const int grid = ls.grid;
    const int n = (int)r.bitWidthIn, m = (int)r.bitWidthOut;
    switch (r.op)
    {
        case FAST_FSS_SERVE_DCF_EVAL:
            return launch_dcf_eval_g1_rt(r.elementSize, grid, st, r.sharedOut, r.maskedX, r.key, r.seed, r.partyId, n,
                                         m, r.elementNum, ls.block);
        case FAST_FSS_SERVE_DPF_EVAL:
            return launch_dpf_eval_g1_rt(r.elementSize, grid, st, r.sharedOut, r.maskedX, r.key, r.seed, r.partyId, n,
                                         m, r.elementNum, ls.block);
        default: break;
    }
    // MIC and spline: mode 0 takes the per-call entry points' work-queue
    // choice (MIC from kMicQueueMinElems elements, the spline at every size);
    // without a counter (mode 1, or none to be had) the same kernels run on
    // the static grid split, which gives the same outputs
    const bool wantQueue = s->launchMode == 0 &&
                           (r.op == FAST_FSS_SERVE_SPLINE_EVAL || r.elementNum >= kMicQueueMinElems);
    uint32_t  *counter   = wantQueue ? queue_counter_acquire(st) : nullptr;
    int        rc;
    if (r.op == FAST_FSS_SERVE_MIC_EVAL)
        rc = launch_mic_eval_rt(r.elementSize, grid, st, r.sharedOut, r.maskedX, r.key, r.sharedZ, r.seed, r.partyId,
                                r.leftEndpoints, r.rightEndpoints, r.intervalNum, n, m, r.elementNum, counter,
                                ls.block);
    else
        rc = launch_spline_eval_rt(r.elementSize, grid, st, r.sharedOut, r.maskedX, r.key, r.sharedZ, r.sharedBeta,
                                   r.seed, r.partyId, r.leftEndpoints, r.rightEndpoints, r.intervalNum, (int)r.degree,
                                   n, m, r.elementNum, counter, ls.block);
    if (counter) queue_counter_release(st, counter);
    return rc;
}

static int pick_stream(Server *s)
{
    const int k    = (int)s->streams.size();
    int       best = s->nextStream;
    for (int j = 1; j < k; j++)
    {
        const int c = (s->nextStream + j) % k;
        if (s->streamLoad[c] < s->streamLoad[best]) best = c;
    }
    s->nextStream = (best + 1) % k;
    return best;
}

static hipEvent_t take_event(Server *s)
{
    if (!s->eventPool.empty())
    {
        hipEvent_t e = s->eventPool.back();
        s->eventPool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    return hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess ? e : nullptr;
}

// Enqueue launch L (its items and slot already set) on a stream and record
// its completion event.  Returns a FastFssError.
static int start_launch(Server *s, Launch &L)
{
    L.stream       = pick_stream(s);
    hipStream_t st = s->streams[L.stream];
    int         rc = FAST_FSS_SUCCESS;
    if (L.slot >= 0)
    {
        std::vector<FastFssBatchGroup> gs(L.items.size());
        for (size_t k = 0; k < L.items.size(); k++)
        {
            const FastFssServeRequest &r = L.items[k]->req;
            FastFssBatchGroup         &g = gs[k];
            g.op          = r.op == FAST_FSS_SERVE_DCF_EVAL ? FAST_FSS_BATCH_DCF_EVAL : FAST_FSS_BATCH_DPF_EVAL;
            g.partyId     = r.partyId;
            g.bitWidthIn  = r.bitWidthIn;
            g.bitWidthOut = r.bitWidthOut;
            g.elementSize = r.elementSize;
            g.elementNum  = r.elementNum;
            g.sharedOut   = r.sharedOut;
            g.maskedX     = r.maskedX;
            g.key         = r.key;
            g.seed        = r.seed;
            g.keyDataSize = r.keyDataSize;
            g.tileElems   = 0;
        }
        const FusedSlot &f = s->slots[L.slot];
        rc = batch_enqueue_owned(gs.data(), gs.size(), f.host, f.dev, f.counter, (size_t)s->maxInflight, st);
    }
    else
        rc = launch(s, L.items[0], st);
    if (rc != FAST_FSS_SUCCESS) return rc;
    L.ev = take_event(s);
    if (!L.ev) return FAST_FSS_RUNTIME_ERROR;
    if (hipEventRecord(L.ev, st) != hipSuccess)
    {
        s->eventPool.push_back(L.ev);
        L.ev = nullptr;
        return FAST_FSS_RUNTIME_ERROR;
    }
    return FAST_FSS_SUCCESS;
}

// Group the requests admitted in one pass into launches and start them;
// requests whose launch failed go to `failed` with their error code.
static void launch_admitted(Server *s, const std::vector<ServeItem *> &admitted,
                            std::vector<std::pair<ServeItem *, int>> &failed)
{
    std::vector<ServeItem *> fuse;
    std::vector<Launch>      launches;
    for (ServeItem *it : admitted)
    {
        const int op = it->req.op;
        if ((op == FAST_FSS_SERVE_DCF_EVAL || op == FAST_FSS_SERVE_DPF_EVAL) && it->req.elementNum > 0)
            fuse.push_back(it);
        else
        {
            Launch L;
            L.items.push_back(it);
            launches.push_back(std::move(L));
        }
    }
    int slot = -1;
    if (fuse.size() >= 2)
        for (int k = 0; k < (int)s->slots.size() && slot < 0; k++)
            if (!s->slots[k].busy) slot = k;
    if (slot >= 0)
    {
        Launch L;
        L.items = fuse;
        L.slot  = slot;
        launches.insert(launches.begin(), std::move(L));
    }
    else
        for (ServeItem *it : fuse)
        {
            Launch L;
            L.items.push_back(it);
            launches.insert(launches.begin(), std::move(L));
        }
    for (Launch &L : launches)
    {
        const int rc = start_launch(s, L);
        if (rc != FAST_FSS_SUCCESS)
        {
            for (ServeItem *it : L.items) failed.emplace_back(it, rc);
            continue;
        }
        if (L.slot >= 0)
        {
            s->slots[L.slot].busy = true;
            s->newFused += L.items.size();
        }
        s->newLaunches++;
        s->streamLoad[L.stream]++;
        s->running.push_back(std::move(L));
    }
}

// Mark `it` finished with `status` (under the mutex).
static void retire(Server *s, ServeItem *it, int status, Clock::time_point now)
{
    it->tDone  = now;
    it->status = status;
    s->inflightBytes -= it->bytes;
    s->inflightItems--;
    if (status == 0)
    {
        s->stats.completed++;
        s->stats.elementsDone += it->req.elementNum;
    }
    else
        s->stats.failed++;
    if (s->inflightItems == 0) s->stats.busySeconds += std::chrono::duration<double>(now - s->busySince).count();
}

static void scheduler(Server *s)
{
    (void)hipSetDevice(s->device);
    std::vector<ServeItem *>                 admitted;
    std::vector<std::pair<ServeItem *, int>> finished, failed;
    while (true)
    {
        // 1. completions (scheduler-owned state, no lock)
        finished.clear();
        for (size_t i = 0; i < s->running.size();)
        {
            Launch          &L = s->running[i];
            const hipError_t q = hipEventQuery(L.ev);
            if (q == hipErrorNotReady)
            {
                i++;
                continue;
            }
            const int status = q == hipSuccess ? 0 : FAST_FSS_RUNTIME_ERROR;
            for (ServeItem *it : L.items) finished.emplace_back(it, status);
            s->eventPool.push_back(L.ev);
            s->streamLoad[L.stream]--;
            if (L.slot >= 0) s->slots[L.slot].busy = false;
            s->running[i] = std::move(s->running.back());
            s->running.pop_back();
        }
        // 2. retire and admit
        admitted.clear();
        {
            std::unique_lock<std::mutex> lock(s->mu);
            s->stats.launches += s->newLaunches;
            s->stats.fusedRequests += s->newFused;
            s->newLaunches = s->newFused = 0;
            if (!finished.empty() || !failed.empty())
            {
                const auto now = Clock::now();
                for (auto &f : finished) retire(s, f.first, f.second, now);
                for (auto &f : failed) retire(s, f.first, f.second, now);
                failed.clear();
                s->cvDone.notify_all();
            }
            while (!s->queue.empty() && s->inflightItems < s->maxInflight)
            {
                ServeItem *it = s->queue.front();
                if (s->inflightItems > 0 && s->inflightBytes + it->bytes > s->budget) break;
                s->queue.pop_front();
                if (s->inflightItems == 0) s->busySince = Clock::now();
                it->status = 2;
                s->inflightItems++;
                s->inflightBytes += it->bytes;
                admitted.push_back(it);
            }
            if ((uint64_t)s->inflightItems > s->stats.maxInflightSeen) s->stats.maxInflightSeen = s->inflightItems;
            if (s->inflightItems == 0)
            {
                if (s->stop && s->queue.empty()) break;
                s->cvWork.wait(lock, [&] { return s->stop || !s->queue.empty(); });
                continue;
            }
        }
        // 3. launches (no lock)
        if (!admitted.empty())
            launch_admitted(s, admitted, failed);
        else
            // poll completions at a fine grain while work is in flight
            std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

} // namespace fss

using namespace fss;

static void server_free(Server *s)
{
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    for (auto e : s->eventPool) (void)hipEventDestroy(e);
    for (auto &f : s->slots)
    {
        if (f.host) (void)hipHostFree(f.host);
        if (f.dev) (void)hipFree(f.dev);
        if (f.counter) (void)hipFree(f.counter);
    }
    delete s;
}

extern "C" FAST_FSS_API int FastFss_hip_serverCreate(void **server, size_t hbmBudgetBytes, int maxInflight,
                                                     int numStreams, int launchMode)
{
    if (!server || maxInflight < 1 || maxInflight > 4096 || numStreams < 1 || numStreams > 64 ||
        (launchMode != 0 && launchMode != 1))
        return FAST_FSS_RUNTIME_ERROR;
    Server *s      = new Server();
    s->budget      = hbmBudgetBytes;
    s->maxInflight = maxInflight;
    s->launchMode  = launchMode;
    bool ok        = hipGetDevice(&s->device) == hipSuccess;
    for (int i = 0; ok && i < numStreams; i++)
    {
        hipStream_t st;
        ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
        if (ok) s->streams.push_back(st);
    }
    s->streamLoad.assign(s->streams.size(), 0);
    // a fused launch serves two or more requests, so maxInflight / 2 of them
    // can be in flight at once
    if (maxInflight >= 2) s->slots.resize((size_t)maxInflight / 2);
    const size_t tab = batch_table_entry_bytes() * (size_t)maxInflight;
    for (auto &f : s->slots)
    {
        if (!ok) break;
        ok = hipHostMalloc(&f.host, tab, hipHostMallocDefault) == hipSuccess &&
             hipMalloc(&f.dev, tab) == hipSuccess && hipMalloc((void **)&f.counter, sizeof(uint32_t)) == hipSuccess;
    }
    if (!ok)
    {
        server_free(s);
        return FAST_FSS_RUNTIME_ERROR;
    }
    s->th   = std::thread(scheduler, s);
    *server = s;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_hip_serverSubmit(void *server, const FastFssServeRequest *req, uint64_t *ticket)
{
    Server *s = (Server *)server;
    if (!s || !req || !ticket) return FAST_FSS_RUNTIME_ERROR;
    int rc = validate(*req);
    if (rc) return rc;
    auto it     = std::make_unique<ServeItem>();
    it->req     = *req;
    it->bytes   = request_bytes(*req);
    it->tSubmit = Clock::now();
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->stop) return FAST_FSS_RUNTIME_ERROR;
    it->id  = s->nextId++;
    *ticket = it->id;
    s->queue.push_back(it.get());
    s->items.emplace(it->id, std::move(it));
    s->stats.submitted++;
    s->cvWork.notify_one();
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_hip_serverQuery(void *server, uint64_t ticket)
{
    Server *s = (Server *)server;
    if (!s) return FAST_FSS_RUNTIME_ERROR;
    std::lock_guard<std::mutex> lock(s->mu);
    auto                        f = s->items.find(ticket);
    if (f == s->items.end()) return FAST_FSS_RUNTIME_ERROR;
    const int st = f->second->status;
    return st == 0 ? 1 : (st < 0 ? st : 0);
}

extern "C" FAST_FSS_API int FastFss_hip_serverWait(void *server, uint64_t ticket, double *latencySeconds)
{
    Server *s = (Server *)server;
    if (!s) return FAST_FSS_RUNTIME_ERROR;
    std::unique_lock<std::mutex> lock(s->mu);
    auto                         f = s->items.find(ticket);
    if (f == s->items.end()) return FAST_FSS_RUNTIME_ERROR;
    ServeItem *it = f->second.get();
    s->cvDone.wait(lock, [&] { return it->status <= 0; });
    if (latencySeconds) *latencySeconds = std::chrono::duration<double>(it->tDone - it->tSubmit).count();
    return it->status;
}

extern "C" FAST_FSS_API int FastFss_hip_serverRelease(void *server, uint64_t ticket)
{
    Server *s = (Server *)server;
    if (!s) return FAST_FSS_RUNTIME_ERROR;
    std::lock_guard<std::mutex> lock(s->mu);
    auto                        f = s->items.find(ticket);
    if (f == s->items.end() || f->second->status > 0) return FAST_FSS_RUNTIME_ERROR;
    s->items.erase(f);
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_hip_serverDrain(void *server)
{
    Server *s = (Server *)server;
    if (!s) return FAST_FSS_RUNTIME_ERROR;
    std::unique_lock<std::mutex> lock(s->mu);
    s->cvDone.wait(lock, [&] { return s->queue.empty() && s->inflightItems == 0; });
    return s->stats.failed ? FAST_FSS_RUNTIME_ERROR : FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_hip_serverStats(void *server, FastFssServeStats *stats)
{
    Server *s = (Server *)server;
    if (!s || !stats) return FAST_FSS_RUNTIME_ERROR;
    std::lock_guard<std::mutex> lock(s->mu);
    *stats = s->stats;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API void FastFss_hip_serverDestroy(void *server)
{
    Server *s = (Server *)server;
    if (!s) return;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
        s->cvWork.notify_all();
    }
    s->th.join();
    server_free(s);
}
