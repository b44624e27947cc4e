// ffr_pool.hip - the context's cache of scene resources: device blocks
// (scene, multigrid and lattice arenas), pinned host blocks (the per-scene
// mirrors of the step record and solver scalars), streams and events.
//
// A serving loop creates a scene for every request and destroys it when the
// request ends, while the other scenes' steps run on their own streams.
// hipFree and hipHostFree wait for the whole device before they return, so
// each request's end would drain every stream; handing the blocks back to the
// context instead keeps the other streams busy, and the next scene of about
// the same size takes the block without a hipMalloc. The caller of
// dev_give/pinned_give/stream_give has already waited for the scene's own
// stream (and the gang stream of a scene in a gang), so no launch still uses
// a block when it is cached.
#include "ffr_scene.hpp"

#include <algorithm>
#include <iterator>

namespace {
constexpr size_t kMaxCachedBlocks = 64;       // device blocks kept at most
constexpr size_t kMaxCachedSmall = 256;       // pinned blocks, streams, events kept at most
constexpr size_t kSlackFloor = size_t(1) << 20;  // a cached block may exceed the request by
                                                 // max(1 MiB, a quarter of it)
}  // namespace

void *ffr_ctx::dev_take(size_t bytes, size_t *got)
{
    std::lock_guard<std::mutex> hold(pool.mu);
    auto it = pool.dev.lower_bound(bytes);
    if (it != pool.dev.end() && it->first <= bytes + std::max(bytes / 4, kSlackFloor) &&
        in_use + it->first <= budget) {
        void *p = it->second;
        *got = it->first;
        pool.dev_bytes -= it->first;
        pool.dev.erase(it);
        in_use += *got;
        return p;
    }
    // a new block: make room beside the cached ones, largest first
    while (!pool.dev.empty() && in_use + pool.dev_bytes + bytes > budget) {
        auto last = std::prev(pool.dev.end());
        hipFree(last->second);
        pool.dev_bytes -= last->first;
        pool.dev.erase(last);
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess && !pool.dev.empty()) {  // the device itself is short: drop the cache
        (void)hipGetLastError();
        for (auto &kv : pool.dev) hipFree(kv.second);
        pool.dev.clear();
        pool.dev_bytes = 0;
        e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    *got = bytes;
    in_use += bytes;
    return p;
}

void ffr_ctx::dev_give(void *p, size_t bytes)
{
    if (!p) return;
    std::lock_guard<std::mutex> hold(pool.mu);
    in_use -= bytes;
    if (pool.dev.size() >= kMaxCachedBlocks) {
        hipFree(p);
        return;
    }
    pool.dev.emplace(bytes, p);
    pool.dev_bytes += bytes;
}

void *ffr_ctx::pinned_take()
{
    {
        std::lock_guard<std::mutex> hold(pool.mu);
        if (!pool.pinned.empty()) {
            void *p = pool.pinned.back();
            pool.pinned.pop_back();
            return p;
        }
    }
    void *p = nullptr;
    if (hipHostMalloc(&p, ffr::kPinnedBlock) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

void ffr_ctx::pinned_give(void *p)
{
    if (!p) return;
    std::lock_guard<std::mutex> hold(pool.mu);
    if (pool.pinned.size() >= kMaxCachedSmall) {
        hipHostFree(p);
        return;
    }
    pool.pinned.push_back(p);
}

hipStream_t ffr_ctx::stream_take()
{
    {
        std::lock_guard<std::mutex> hold(pool.mu);
        if (!pool.streams.empty()) {
            hipStream_t st = pool.streams.back();
            pool.streams.pop_back();
            return st;
        }
    }
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return st;
}

void ffr_ctx::stream_give(hipStream_t st)
{
    if (!st) return;
    std::lock_guard<std::mutex> hold(pool.mu);
    if (pool.streams.size() >= kMaxCachedSmall) {
        hipStreamDestroy(st);
        return;
    }
    pool.streams.push_back(st);
}

hipEvent_t ffr_ctx::event_take()
{
    {
        std::lock_guard<std::mutex> hold(pool.mu);
        if (!pool.events.empty()) {
            hipEvent_t ev = pool.events.back();
            pool.events.pop_back();
            return ev;
        }
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return ev;
}

void ffr_ctx::event_give(hipEvent_t ev)
{
    if (!ev) return;
    std::lock_guard<std::mutex> hold(pool.mu);
    if (pool.events.size() >= kMaxCachedSmall) {
        hipEventDestroy(ev);
        return;
    }
    pool.events.push_back(ev);
}

size_t ffr_ctx::pool_trim()
{
    std::lock_guard<std::mutex> hold(pool.mu);
    const size_t freed = pool.dev_bytes;
    for (auto &kv : pool.dev) hipFree(kv.second);
    pool.dev.clear();
    pool.dev_bytes = 0;
    for (void *p : pool.pinned) hipHostFree(p);
    pool.pinned.clear();
    for (hipStream_t st : pool.streams) hipStreamDestroy(st);
    pool.streams.clear();
    for (hipEvent_t ev : pool.events) hipEventDestroy(ev);
    pool.events.clear();
    return freed;
}

extern "C" {

size_t ffr_ctx_trim(ffr_ctx *c)
{
    if (!c) return 0;
    hipSetDevice(c->device);
    return c->pool_trim();
}

size_t ffr_ctx_bytes_cached(const ffr_ctx *c)
{
    if (!c) return 0;
    std::lock_guard<std::mutex> hold(const_cast<ffr_ctx *>(c)->pool.mu);
    return c->pool.dev_bytes;
}

}  // extern "C"
