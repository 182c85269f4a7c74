/* Continuous-batching request server on MI355X (no counterpart in the
 * reference, whose C API is one synchronous-launch call per evaluation).
 *
 * Requests (DCF, DPF, DCF-MIC or DCF-spline evaluations whose keys and inputs are
 * already resident in device memory) are submitted from any host thread.  A
 * scheduler thread admits queued requests in arrival order while
 *   - the number of requests in flight is below `maxInflight`, and
 *   - the bytes they touch (key + inputs + outputs) fit the HBM budget
 *     (a request larger than the budget is admitted alone),
 * launches the admitted requests on `numStreams` HIP streams (each launch on
 * the stream with the fewest launches in flight) and retires them when their
 * completion events fire.  The scheduler launches without holding the lock
 * that Submit, Query and Wait take.
 *
 * DCF and DPF requests admitted together (two or more in one scheduling pass)
 * share one persistent launch of the ragged-batch kernel
 * (include/FastFss/hip/batch.h), which pulls tiles of all of them from one work
 * queue.  MIC and spline requests, and a DCF or DPF request admitted alone, get
 * a launch each.  Work of concurrently admitted launches overlaps on the
 * device: with launchMode 1 every single-request launch uses one workgroup per
 * 1024 elements, so the hardware interleaves the workgroups of all launches in
 * flight across the CUs.
 *
 * The runtime gives each process GPU_MAX_HW_QUEUES hardware queues (4 unless
 * set); streams beyond that share queues and add no concurrency.
 *
 * Each request's output is bit-identical to the per-call entry point with
 * the same arguments (FastFss_cuda_dcfEval, FastFss_cuda_dpfEval,
 * FastFss_cuda_dcfMICEval, FastFss_cuda_dcfSplineEval).  Latency = completion observed by the
 * scheduler minus submission.
 */
#ifndef FAST_FSS_HIP_SERVE_H
#define FAST_FSS_HIP_SERVE_H

#include <FastFss/api.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum FastFssServeOp
{
    FAST_FSS_SERVE_DCF_EVAL    = 0, /* groupSize 1 */
    FAST_FSS_SERVE_MIC_EVAL    = 1,
    FAST_FSS_SERVE_SPLINE_EVAL = 2, /* degree <= 7 */
    FAST_FSS_SERVE_DPF_EVAL    = 3, /* groupSize 1 */
};

typedef struct FastFssServeRequest
{
    int         op;
    int         partyId;
    size_t      bitWidthIn;
    size_t      bitWidthOut;
    size_t      elementSize;
    size_t      elementNum;
    void       *sharedOut;
    const void *maskedX;
    const void *key;
    size_t      keyDataSize;
    const void *seed;
    const void *sharedZ;        /* MIC: z shares; spline: e shares */
    const void *sharedBeta;     /* spline only */
    const void *leftEndpoints;  /* MIC / spline */
    const void *rightEndpoints; /* MIC / spline */
    size_t      intervalNum;    /* MIC / spline */
    size_t      degree;         /* spline */
} FastFssServeRequest;

typedef struct FastFssServeStats
{
    uint64_t submitted;
    uint64_t completed;
    uint64_t failed;
    uint64_t maxInflightSeen;
    uint64_t elementsDone;
    double   busySeconds;   /* wall time with at least one request in flight */
    uint64_t launches;      /* kernel launches made (a fused launch counts once) */
    uint64_t fusedRequests; /* requests served by fused DCF/DPF batch launches */
} FastFssServeStats;

/* launchMode: 0 = the per-call entry points' shape (one 1024-lane workgroup
 * per CU once the request fills every CU, smaller workgroups spread over all
 * CUs below that), 1 = one 1024-lane workgroup per 1024 elements (interleaves
 * across requests). */
FAST_FSS_API int FastFss_hip_serverCreate(void **server, size_t hbmBudgetBytes, int maxInflight, int numStreams,
                                          int launchMode);

/* Validate and enqueue; *ticket identifies the request.  Non-blocking. */
FAST_FSS_API int FastFss_hip_serverSubmit(void *server, const FastFssServeRequest *req, uint64_t *ticket);

/* 1 if done, 0 if pending or running, negative FastFssError if it failed. */
FAST_FSS_API int FastFss_hip_serverQuery(void *server, uint64_t ticket);

/* Block until the request is done; returns its status and, if latencySeconds
 * is not NULL, its submit-to-completion latency. */
FAST_FSS_API int FastFss_hip_serverWait(void *server, uint64_t ticket, double *latencySeconds);

/* Forget a finished (done or failed) request: its ticket becomes unknown to
 * Query and Wait.  A request's record otherwise stays until
 * FastFss_hip_serverDestroy, so a long-running caller releases each ticket
 * once it has the result.  Returns FAST_FSS_RUNTIME_ERROR for an unknown
 * ticket or one still queued or running. */
FAST_FSS_API int FastFss_hip_serverRelease(void *server, uint64_t ticket);

/* Block until every submitted request is done. */
FAST_FSS_API int FastFss_hip_serverDrain(void *server);

FAST_FSS_API int FastFss_hip_serverStats(void *server, FastFssServeStats *stats);

/* Drain, stop the scheduler thread and release the server. */
FAST_FSS_API void FastFss_hip_serverDestroy(void *server);

#ifdef __cplusplus
}
#endif

#endif
