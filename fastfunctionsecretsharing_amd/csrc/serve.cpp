// Continuous-batching request server (see include/FastFss/hip/serve.h).
//
// One scheduler thread owns all device work: it admits queued requests in
// arrival order under the in-flight and HBM-budget limits, launches each on
// the least recently used of its streams, records a completion event and
// retires requests whose events have fired.  Client threads only touch the
// queue and the request records under the mutex.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

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
    hipEvent_t          ev     = nullptr;
};

struct Server
{
    std::mutex                                            mu;
    std::condition_variable                               cvWork, cvDone;
    std::deque<ServeItem *>                               queue;
    std::vector<ServeItem *>                              running;
    std::unordered_map<uint64_t, std::unique_ptr<ServeItem>> items;
    std::vector<hipStream_t>                              streams;
    size_t                                                budget, inflightBytes = 0;
    int                                                   maxInflight, launchMode, device = 0, nextStream = 0;
    uint64_t                                              nextId = 1;
    bool                                                  stop   = false;
    FastFssServeStats                                     stats{};
    Clock::time_point                                     busySince;
    std::thread                                           th;
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
    if (!elem_size_ok(es)) return FAST_FSS_INVALID_ELEMENT_SIZE_ERROR;
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
        case FAST_FSS_SERVE_MIC_EVAL:
            return launch_mic_eval_rt(r.elementSize, grid, st, r.sharedOut, r.maskedX, r.key, r.sharedZ, r.seed,
                                      r.partyId, r.leftEndpoints, r.rightEndpoints, r.intervalNum, n, m,
                                      r.elementNum, nullptr, ls.block);
        default:
            return launch_spline_eval_rt(r.elementSize, grid, st, r.sharedOut, r.maskedX, r.key, r.sharedZ,
                                         r.sharedBeta, r.seed, r.partyId, r.leftEndpoints, r.rightEndpoints,
                                         r.intervalNum, (int)r.degree, n, m, r.elementNum, nullptr, ls.block);
    }
}

static void scheduler(Server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lock(s->mu);
    while (true)
    {
        // retire finished requests
        for (size_t i = 0; i < s->running.size();)
        {
            ServeItem  *it = s->running[i];
            hipError_t q  = hipEventQuery(it->ev);
            if (q == hipErrorNotReady)
            {
                i++;
                continue;
            }
            it->tDone  = Clock::now();
            it->status = (q == hipSuccess) ? 0 : FAST_FSS_RUNTIME_ERROR;
            (void)hipEventDestroy(it->ev);
            it->ev = nullptr;
            s->inflightBytes -= it->bytes;
            if (it->status == 0)
            {
                s->stats.completed++;
                s->stats.elementsDone += it->req.elementNum;
            }
            else
                s->stats.failed++;
            s->running[i] = s->running.back();
            s->running.pop_back();
            if (s->running.empty())
                s->stats.busySeconds += std::chrono::duration<double>(it->tDone - s->busySince).count();
            s->cvDone.notify_all();
        }
        // admit in arrival order
        while (!s->queue.empty() && (int)s->running.size() < s->maxInflight)
        {
            ServeItem *it = s->queue.front();
            if (!s->running.empty() && s->inflightBytes + it->bytes > s->budget) break;
            s->queue.pop_front();
            hipStream_t st = s->streams[s->nextStream];
            s->nextStream  = (s->nextStream + 1) % (int)s->streams.size();
            int rc         = launch(s, it, st);
            if (rc == FAST_FSS_SUCCESS && hipEventCreateWithFlags(&it->ev, hipEventDisableTiming) != hipSuccess)
                rc = FAST_FSS_RUNTIME_ERROR;
            if (rc == FAST_FSS_SUCCESS && hipEventRecord(it->ev, st) != hipSuccess) rc = FAST_FSS_RUNTIME_ERROR;
            if (rc != FAST_FSS_SUCCESS)
            {
                if (it->ev) (void)hipEventDestroy(it->ev);
                it->ev     = nullptr;
                it->status = rc;
                it->tDone  = Clock::now();
                s->stats.failed++;
                s->cvDone.notify_all();
                continue;
            }
            if (s->running.empty()) s->busySince = Clock::now();
            it->status = 2;
            s->running.push_back(it);
            s->inflightBytes += it->bytes;
            if (s->running.size() > s->stats.maxInflightSeen) s->stats.maxInflightSeen = s->running.size();
        }
        if (s->stop && s->queue.empty() && s->running.empty()) break;
        if (s->running.empty())
            s->cvWork.wait(lock, [&] { return s->stop || !s->queue.empty(); });
        else
        {
            // poll completions at a fine grain while work is in flight
            lock.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lock.lock();
        }
    }
}

} // namespace fss

using namespace fss;

extern "C" FAST_FSS_API int FastFss_hip_serverCreate(void **server, size_t hbmBudgetBytes, int maxInflight,
                                                     int numStreams, int launchMode)
{
    if (!server || maxInflight < 1 || numStreams < 1 || numStreams > 64 || (launchMode != 0 && launchMode != 1))
        return FAST_FSS_RUNTIME_ERROR;
    Server *s      = new Server();
    s->budget      = hbmBudgetBytes;
    s->maxInflight = maxInflight;
    s->launchMode  = launchMode;
    if (hipGetDevice(&s->device) != hipSuccess)
    {
        delete s;
        return FAST_FSS_RUNTIME_ERROR;
    }
    for (int i = 0; i < numStreams; i++)
    {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        {
            for (auto x : s->streams) (void)hipStreamDestroy(x);
            delete s;
            return FAST_FSS_RUNTIME_ERROR;
        }
        s->streams.push_back(st);
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
    s->cvDone.wait(lock, [&] { return s->queue.empty() && s->running.empty(); });
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
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    delete s;
}
