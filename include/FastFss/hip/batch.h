/* Ragged batched evaluation on MI355X: many DCF / DPF key groups of mixed
 * widths in one persistent launch.
 *
 * These entry points have no counterpart in the reference (whose API takes
 * one homogeneous group per call, include/FastFss/cuda/dcf.h:167-183,
 * include/FastFss/cuda/dpf.h:30-46).  Each group's shares are bit-identical
 * to a FastFss_cuda_dcfEval / FastFss_cuda_dpfEval call with the same
 * arguments (groupSize 1, no cache); the batch only changes how the work is
 * scheduled: one persistent workgroup per CU pulls tiles of elements from a
 * device work queue that spans all groups.
 *
 * A plan holds the device-side group table and queue counter; launching a
 * plan performs no allocation and no memset, so the launch can be captured
 * into a hipGraph (FastFss_hip_batchPlanCapture) and replayed.
 */
#ifndef FAST_FSS_HIP_BATCH_H
#define FAST_FSS_HIP_BATCH_H

#include <FastFss/api.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum FastFssBatchOp
{
    FAST_FSS_BATCH_DCF_EVAL = 0, /* same shares as FastFss_cuda_dcfEval, groupSize 1 */
    FAST_FSS_BATCH_DPF_EVAL = 1, /* same shares as FastFss_cuda_dpfEval, groupSize 1 */
};

typedef struct FastFssBatchGroup
{
    int         op;          /* FastFssBatchOp */
    int         partyId;     /* 0 or 1 */
    size_t      bitWidthIn;  /* 1 .. 8*elementSize */
    size_t      bitWidthOut; /* 1 .. 8*elementSize */
    size_t      elementSize; /* 1, 2, 4 or 8 */
    size_t      elementNum;
    void       *sharedOut;   /* device, elementNum * elementSize bytes */
    const void *maskedX;     /* device, elementNum * elementSize bytes */
    const void *key;         /* device, FastFss_{dcf,dpf}GetKeyDataSize(..., groupSize 1) bytes */
    const void *seed;        /* device, 16 * elementNum bytes */
    size_t      keyDataSize; /* validated like the per-call API */
    uint32_t    tileElems;   /* elements per queue tile (multiple of 64); 0 = tuned default for the width */
} FastFssBatchGroup;

/* Validate `groups` (host array) and build a plan on the current device.
 * Returns 0 or the first FastFssError found (with *plan untouched). */
FAST_FSS_API int FastFss_hip_batchPlanCreate(void **plan, const FastFssBatchGroup *groups, size_t groupNum);

/* Enqueue the plan's persistent launch on *streamPtr (NULL: per-thread default stream). */
FAST_FSS_API int FastFss_hip_batchPlanLaunch(void *plan, void *streamPtr);

/* Record the plan's launch into a hipGraph once (stream capture on *streamPtr). */
FAST_FSS_API int FastFss_hip_batchPlanCapture(void *plan, void *streamPtr);

/* Replay the captured graph on *streamPtr. */
FAST_FSS_API int FastFss_hip_batchPlanReplay(void *plan, void *streamPtr);

/* Total queue tiles of the plan (for reporting). */
FAST_FSS_API size_t FastFss_hip_batchPlanTileCount(const void *plan);

/* Release the plan (synchronises the device first). */
FAST_FSS_API void FastFss_hip_batchPlanDestroy(void *plan);

/* One-shot convenience: create, launch and (stream-ordered) release. */
FAST_FSS_API int FastFss_hip_batchEval(const FastFssBatchGroup *groups, size_t groupNum, void *streamPtr);

/* Tile size (elements per queue tile) used when a group asks for 0:
 * per (op, elementSize) entry of the tuned table; the autotuner writes it. */
FAST_FSS_API int FastFss_hip_batchSetDefaultTile(int op, size_t elementSize, uint32_t tileElems);
FAST_FSS_API uint32_t FastFss_hip_batchGetDefaultTile(int op, size_t elementSize);

/* Workgroups per CU for the persistent launch (1 by default; the AES tables
 * take 128 KB of the CU's 160 KB LDS, so a second resident workgroup cannot
 * fit and values above 1 only queue extra workgroups behind the first). */
FAST_FSS_API int FastFss_hip_batchSetBlocksPerCu(int blocks);

/* Groupsize-1 DCF eval kernel variant used by FastFss_cuda_dcfEval: -1 =
 * the per-element-size table, else the default (variant 8 up to 64 elements
 * per CU, variant 7 up to 256, variant 2 up to 2^20 elements, variant 6 from
 * there); 0 = per-level key loads; 1..4 = burst loads of CH levels with E
 * elements per lane {(4,1), (8,1), (4,2), (2,2)}, static grid split; 5 =
 * 8-level bursts on the work queue; 6 = variant 5 with each element's value
 * corrections loaded when its walk starts (elements of at most 32 bits; 64-bit
 * elements run variant 5); 7 = two lanes per element, one AES block each per
 * level; 8 = eight lanes per element, one column of one AES block each.
 * Every variant gives the same shares. */
FAST_FSS_API int FastFss_hip_setEvalVariant(int variant);
FAST_FSS_API int FastFss_hip_getEvalVariantCount(void);
FAST_FSS_API int FastFss_hip_setEvalVariantForElementSize(size_t elementSize, int variant);

#ifdef __cplusplus
}
#endif

#endif
