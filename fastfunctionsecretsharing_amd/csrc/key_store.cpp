// Device-resident key store: one device allocation carved into 256-byte
// aligned slots (first fit, adjacent free extents merged), with
// stream-ordered release through events.  See include/FastFss/hip/key_store.h.
#include "FastFss/hip/key_store.h"

#include <hip/hip_runtime.h>

#include <deque>
#include <iterator>
#include <map>
#include <mutex>
#include <unordered_map>

#include "fss_host.h"

namespace {

constexpr size_t kAlign = 256;

size_t round_up(size_t v)
{
    return (v + kAlign - 1) / kAlign * kAlign;
}

struct Slot
{
    size_t offset, bytes, keyBytes;
};

struct PendingRelease
{
    hipEvent_t ev;
    size_t     offset, bytes;
};

} // namespace

struct FastFssKeyStore
{
    uint8_t                           *base     = nullptr;
    size_t                             capacity = 0;
    int                                device   = 0;
    std::mutex                         mu;
    std::map<size_t, size_t>           freeList; // offset -> bytes, disjoint, never adjacent
    std::unordered_map<uint64_t, Slot> slots;
    std::deque<PendingRelease>         pending; // in release order
    std::deque<hipEvent_t>             spareEvents;
    uint64_t                           nextHandle = 1;
    size_t                             used       = 0;

    // both under mu
    void give_back(size_t offset, size_t bytes)
    {
        auto next = freeList.lower_bound(offset);
        if (next != freeList.begin())
        {
            auto prev = std::prev(next);
            if (prev->first + prev->second == offset)
            {
                offset = prev->first;
                bytes += prev->second;
                freeList.erase(prev);
            }
        }
        if (next != freeList.end() && offset + bytes == next->first)
        {
            bytes += next->second;
            freeList.erase(next);
        }
        freeList[offset] = bytes;
    }

    // Moves completed releases to the free list; with `wait`, waits for the
    // oldest pending one first.  Returns false when a HIP call failed.
    bool collect(bool wait)
    {
        while (!pending.empty())
        {
            PendingRelease &p = pending.front();
            hipError_t      q = wait ? hipEventSynchronize(p.ev) : hipEventQuery(p.ev);
            if (q == hipErrorNotReady) return true;
            if (q != hipSuccess) return false;
            give_back(p.offset, p.bytes);
            spareEvents.push_back(p.ev);
            pending.pop_front();
            wait = false;
        }
        return true;
    }

    // offset of a free extent of `bytes`, or SIZE_MAX
    size_t take(size_t bytes)
    {
        for (int attempt = 0;; attempt++)
        {
            if (!collect(attempt > 0)) return SIZE_MAX;
            for (auto it = freeList.begin(); it != freeList.end(); ++it)
            {
                if (it->second < bytes) continue;
                const size_t off = it->first, rest = it->second - bytes;
                freeList.erase(it);
                if (rest) freeList[off + bytes] = rest;
                return off;
            }
            if (pending.empty()) return SIZE_MAX;
        }
    }

    int add_slot(size_t keyBytes, uint64_t *handle, size_t *offset)
    {
        // checked before rounding up, which would wrap for sizes near SIZE_MAX
        if (keyBytes > capacity) return FAST_FSS_RUNTIME_ERROR;
        const size_t bytes = round_up(keyBytes);
        const size_t off   = take(bytes);
        if (off == SIZE_MAX) return FAST_FSS_RUNTIME_ERROR;
        const uint64_t h = nextHandle++;
        slots[h]         = Slot{off, bytes, keyBytes};
        used += bytes;
        *handle = h;
        *offset = off;
        return FAST_FSS_SUCCESS;
    }
};

extern "C" FAST_FSS_API int FastFss_hip_keyStoreCreate(FastFssKeyStore **store, size_t capacityBytes)
{
    if (!store) return FAST_FSS_RUNTIME_ERROR;
    *store = nullptr;
    if (capacityBytes == 0 || capacityBytes > SIZE_MAX - kAlign) return FAST_FSS_RUNTIME_ERROR;
    auto  *s   = new FastFssKeyStore();
    size_t cap = round_up(capacityBytes);
    if (hipGetDevice(&s->device) != hipSuccess || hipMalloc((void **)&s->base, cap) != hipSuccess)
    {
        delete s;
        return FAST_FSS_RUNTIME_ERROR;
    }
    s->capacity    = cap;
    s->freeList[0] = cap;
    *store         = s;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_hip_keyStoreDestroy(FastFssKeyStore *store)
{
    if (!store) return FAST_FSS_RUNTIME_ERROR;
    int rc = FAST_FSS_SUCCESS;
    {
        std::lock_guard<std::mutex> g(store->mu);
        for (PendingRelease &p : store->pending)
        {
            if (hipEventSynchronize(p.ev) != hipSuccess) rc = FAST_FSS_RUNTIME_ERROR;
            (void)hipEventDestroy(p.ev);
        }
        for (hipEvent_t e : store->spareEvents) (void)hipEventDestroy(e);
        // the slots' last users are the caller's streams; finish them before
        // the memory goes back
        int prev = 0;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(store->device);
        if (hipDeviceSynchronize() != hipSuccess) rc = FAST_FSS_RUNTIME_ERROR;
        if (hipFree(store->base) != hipSuccess) rc = FAST_FSS_RUNTIME_ERROR;
        (void)hipSetDevice(prev);
    }
    delete store;
    return rc;
}

extern "C" FAST_FSS_API int FastFss_hip_keyStorePut(FastFssKeyStore *store, const void *key, size_t keyDataSize,
                                                    uint64_t *handle, void *hipStreamPtr)
{
    if (!store || !key || !handle || keyDataSize == 0) return FAST_FSS_RUNTIME_ERROR;
    size_t off = 0;
    {
        std::lock_guard<std::mutex> g(store->mu);
        const int                   rc = store->add_slot(keyDataSize, handle, &off);
        if (rc) return rc;
    }
    if (hipMemcpyAsync(store->base + off, key, keyDataSize, hipMemcpyDefault, fss::resolve_stream(hipStreamPtr)) !=
        hipSuccess)
    {
        std::lock_guard<std::mutex> g(store->mu);
        const Slot                  sl = store->slots[*handle];
        store->slots.erase(*handle);
        store->used -= sl.bytes;
        store->give_back(sl.offset, sl.bytes);
        *handle = 0;
        return FAST_FSS_RUNTIME_ERROR;
    }
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_hip_keyStoreReserve(FastFssKeyStore *store, size_t keyDataSize, uint64_t *handle,
                                                        void **keyPtr)
{
    if (!store || !handle || !keyPtr || keyDataSize == 0) return FAST_FSS_RUNTIME_ERROR;
    std::lock_guard<std::mutex> g(store->mu);
    size_t                      off = 0;
    const int                   rc  = store->add_slot(keyDataSize, handle, &off);
    if (rc) return rc;
    *keyPtr = store->base + off;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_hip_keyStoreGet(FastFssKeyStore *store, uint64_t handle, void **keyPtr,
                                                    size_t *keyDataSize)
{
    if (!store || !keyPtr || !keyDataSize) return FAST_FSS_RUNTIME_ERROR;
    std::lock_guard<std::mutex> g(store->mu);
    auto                        it = store->slots.find(handle);
    if (it == store->slots.end()) return FAST_FSS_RUNTIME_ERROR;
    *keyPtr      = store->base + it->second.offset;
    *keyDataSize = it->second.keyBytes;
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_hip_keyStoreRelease(FastFssKeyStore *store, uint64_t handle, void *hipStreamPtr)
{
    if (!store) return FAST_FSS_RUNTIME_ERROR;
    std::lock_guard<std::mutex> g(store->mu);
    auto                        it = store->slots.find(handle);
    if (it == store->slots.end()) return FAST_FSS_RUNTIME_ERROR;
    hipEvent_t ev = nullptr;
    if (!store->spareEvents.empty())
    {
        ev = store->spareEvents.back();
        store->spareEvents.pop_back();
    }
    else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
        return FAST_FSS_RUNTIME_ERROR;
    if (hipEventRecord(ev, fss::resolve_stream(hipStreamPtr)) != hipSuccess)
    {
        store->spareEvents.push_back(ev);
        return FAST_FSS_RUNTIME_ERROR;
    }
    const Slot sl = it->second;
    store->slots.erase(it);
    store->used -= sl.bytes;
    store->pending.push_back(PendingRelease{ev, sl.offset, sl.bytes});
    return FAST_FSS_SUCCESS;
}

extern "C" FAST_FSS_API int FastFss_hip_keyStoreStats(FastFssKeyStore *store, size_t *bytesUsed, size_t *bytesFree,
                                                      size_t *keyCount, size_t *largestFree)
{
    if (!store) return FAST_FSS_RUNTIME_ERROR;
    std::lock_guard<std::mutex> g(store->mu);
    if (!store->collect(false)) return FAST_FSS_RUNTIME_ERROR;
    size_t freeBytes = 0, largest = 0;
    for (const auto &kv : store->freeList)
    {
        freeBytes += kv.second;
        if (kv.second > largest) largest = kv.second;
    }
    if (bytesUsed) *bytesUsed = store->used;
    if (bytesFree) *bytesFree = freeBytes;
    if (keyCount) *keyCount = store->slots.size();
    if (largestFree) *largestFree = largest;
    return FAST_FSS_SUCCESS;
}
