/* Device-resident key store on MI355X.
 *
 * No counterpart in the reference, whose keys live in caller buffers only
 * (its FastFss_*KeyZip / KeyUnzip wire-format functions are stubs,
 * src/dcf.cpp:10-34, kept as such here).  A store is one device allocation
 * of a fixed capacity, made once, from which the library hands out
 * 256-byte-aligned slots for keys of any primitive: a serving process puts
 * its key sets into HBM once and evaluates from them for as long as it runs,
 * without a hipMalloc/hipFree per key set.
 *
 *  - Put copies a key (host or device memory) into a new slot on a stream;
 *    Reserve hands out a slot without copying, for a keyGen to write into.
 *    Get returns the slot's device pointer and size; that pointer is an
 *    ordinary key buffer for every FastFss_cuda_* entry point and the server.
 *  - Release is stream-ordered: the slot is reused only after the work
 *    enqueued on the given stream before the Release call has completed (an
 *    event recorded there).  Work that reads the key on other streams must be
 *    ordered before that stream by the caller.
 *  - Slots are placed first-fit and adjacent free space is merged; a Put or
 *    Reserve that finds no room waits for pending releases before it fails.
 *  - Every call is thread-safe.  Handles are never reused; 0 is no handle.
 *
 * Errors: FAST_FSS_SUCCESS, FAST_FSS_RUNTIME_ERROR for a null store or
 * pointer, an unknown handle, a zero size or one above the store's capacity,
 * no room, or a failed HIP call.
 */
#ifndef FAST_FSS_HIP_KEY_STORE_H
#define FAST_FSS_HIP_KEY_STORE_H

#include <FastFss/api.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct FastFssKeyStore FastFssKeyStore;

/* One device allocation of capacityBytes (rounded up to 256) on the current
 * device.  No counterpart in the reference. */
FAST_FSS_API int FastFss_hip_keyStoreCreate(FastFssKeyStore **store, size_t capacityBytes);

/* Waits for pending releases and frees the store's memory.  Every slot's
 * pointer becomes invalid. */
FAST_FSS_API int FastFss_hip_keyStoreDestroy(FastFssKeyStore *store);

/* A new slot holding a copy of keyDataSize bytes at `key` (host or device
 * memory), copied on *hipStreamPtr (NULL: the per-thread default stream);
 * the copy is ordered before later work on that stream, and the caller keeps
 * `key` unchanged until it has run. */
FAST_FSS_API int FastFss_hip_keyStorePut(FastFssKeyStore *store, const void *key, size_t keyDataSize,
                                         uint64_t *handle, void *hipStreamPtr);

/* A new slot of keyDataSize bytes, uninitialised, and its device pointer. */
FAST_FSS_API int FastFss_hip_keyStoreReserve(FastFssKeyStore *store, size_t keyDataSize, uint64_t *handle,
                                             void **keyPtr);

/* Device pointer and size of a slot. */
FAST_FSS_API int FastFss_hip_keyStoreGet(FastFssKeyStore *store, uint64_t handle, void **keyPtr,
                                         size_t *keyDataSize);

/* Returns the slot to the store once the work enqueued on *hipStreamPtr so
 * far has completed (NULL: the per-thread default stream). */
FAST_FSS_API int FastFss_hip_keyStoreRelease(FastFssKeyStore *store, uint64_t handle, void *hipStreamPtr);

/* Bytes in live slots (rounded to 256), bytes free now (pending releases
 * not counted), live slots, and the largest free extent. */
FAST_FSS_API int FastFss_hip_keyStoreStats(FastFssKeyStore *store, size_t *bytesUsed, size_t *bytesFree,
                                           size_t *keyCount, size_t *largestFree);

#ifdef __cplusplus
}
#endif

#endif
