// ffr_gang.hpp - one launch for the same kernel of several scenes.
//
// Small scenes are bound by the ~3-4 us a dependent launch costs, not by
// memory: a scene of 2^17 cells issues ~500 launches per step. A gang runs
// the launches of several scenes merged: each merged launch is one grid made
// of every member's own grid (the same block count and the same per-block
// work as its plain launch, so reductions and results are bit-identical),
// and each block finds its scene and that scene's arguments in a table in
// device memory.
//
// Table layout (GangHdr, then one argument record per scene):
//   off[s]   first block of scene s (a multiple of 8, so that block b runs on
//            the XCD it would run on in the scene's own launch); off[s] for
//            s >= n is the grid size
//   G[s]     blocks of scene s's own launch (blocks between off[s] + G[s]
//            and off[s + 1] are padding and return at once)
//   records  GangArgs<P...> of the kernel's parameters, `stride` bytes apart
//
// Kernels take part by having their body in a struct with a static
// run(const VB &, params...) (the plain __global__ calls it with vb_here());
// FFR_GANG launches such a kernel, or records it when a recorder is active.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

#include "ffr_device.hpp"

namespace ffr {

constexpr int kGangMax = 32;  // scenes in one merged launch

struct GangHdr {
    uint32_t n, stride;
    uint32_t off[kGangMax + 1];
    uint32_t G[kGangMax];
};
constexpr uint32_t kGangArgsOff = (sizeof(GangHdr) + 15u) & ~15u;

// ---- argument records: a tuple with the same layout on host and device ----
template <size_t I, typename T>
struct GLeaf {
    T v;
};
template <typename S, typename... P>
struct GangArgsImpl;
template <size_t... I, typename... P>
struct GangArgsImpl<std::index_sequence<I...>, P...> : GLeaf<I, P>... {
    __host__ __device__ explicit GangArgsImpl(P... p) : GLeaf<I, P>{p}... {}
};
template <typename... P>
using GangArgs = GangArgsImpl<std::index_sequence_for<P...>, P...>;

template <size_t I, typename T>
__device__ __forceinline__ const T &gget(const GLeaf<I, T> &l)
{
    return l.v;
}
template <typename B, typename A, size_t... I>
__device__ __forceinline__ void gang_call(const VB &vb, const A &a, std::index_sequence<I...>)
{
    B::run(vb, gget<I>(a)...);
}

// block -> (scene, its VB); false for a padding block (uniform per block)
__device__ __forceinline__ bool gang_locate(const GangHdr *h, VB &vb, uint32_t &s)
{
    const uint32_t b = blockIdx.x;
    uint32_t k = 0;
#pragma unroll
    for (int i = 1; i <= kGangMax; i++) k += h->off[i] <= b ? 1u : 0u;
    const uint32_t local = b - h->off[k], G = h->G[k];
    if (local >= G) return false;
    vb = VB{xcd_remap(local, G), G, local};
    s = k;
    return true;
}

template <typename B, typename... P>
__device__ __forceinline__ void gang_body(const GangHdr *h)
{
    VB vb;
    uint32_t s;
    if (!gang_locate(h, vb, s)) return;
    using A = GangArgs<P...>;
    const A &a = *reinterpret_cast<const A *>(reinterpret_cast<const char *>(h) + kGangArgsOff +
                                              (size_t)s * h->stride);
    gang_call<B>(vb, a, std::index_sequence_for<P...>{});
}

// the launch shapes of the plain kernels: 256 threads, 256 threads at 8
// waves per SIMD, 1024 threads (one-workgroup coarse solves)
enum { GB_PLAIN = 0, GB_RES = 1, GB_WIDE = 2 };

template <typename B, typename... P>
__global__ void __launch_bounds__(kBlock) k_gang(const GangHdr *h)
{
    gang_body<B, P...>(h);
}
template <typename B, typename... P>
__global__ void __launch_bounds__(kBlock) FFR_RESIDENT k_gang_res(const GangHdr *h)
{
    gang_body<B, P...>(h);
}
template <typename B, typename... P>
__global__ void __launch_bounds__(1024) k_gang_wide(const GangHdr *h)
{
    gang_body<B, P...>(h);
}

template <typename B, int BD, typename... P>
const void *gang_kernel()
{
    if constexpr (BD == GB_RES) return reinterpret_cast<const void *>(&k_gang_res<B, P...>);
    else if constexpr (BD == GB_WIDE) return reinterpret_cast<const void *>(&k_gang_wide<B, P...>);
    else return reinterpret_cast<const void *>(&k_gang<B, P...>);
}

// ---- recording -------------------------------------------------------------
// One recorded operation of a scene: a kernel launch (fn = its gang kernel)
// or, with fn == nullptr, a copy of device scalars to pinned host memory.
struct GangOp {
    const void *fn = nullptr;
    uint32_t grid = 0, block = 0, lds = 0, asize = 0;
    size_t aoff = 0;  // argument record in the recorder's arena
    void *hdst = nullptr;
    const void *dsrc = nullptr;
    uint32_t bytes = 0;
};

struct Recorder {
    std::vector<GangOp> ops;
    std::vector<uint8_t> arena;
    bool unsupported = false;  // a launch that cannot join a gang was reached
    void clear()
    {
        ops.clear();
        arena.clear();
        unsupported = false;
    }
    void push(const void *fn, uint32_t grid, uint32_t block, uint32_t lds, const void *args,
              uint32_t asize)
    {
        GangOp op;
        op.fn = fn;
        op.grid = grid;
        op.block = block;
        op.lds = lds;
        op.asize = asize;
        op.aoff = (arena.size() + 15u) & ~(size_t)15u;
        arena.resize(op.aoff + asize);
        std::memcpy(arena.data() + op.aoff, args, asize);
        ops.push_back(op);
    }
    void push_copy(void *hdst, const void *dsrc, uint32_t bytes)
    {
        GangOp op;
        op.hdst = hdst;
        op.dsrc = dsrc;
        op.bytes = bytes;
        ops.push_back(op);
    }
    // repeats the operations [b, e) (same argument records) `times` times
    void repeat(size_t b, size_t e, int times)
    {
        if (times <= 0) return;
        ops.reserve(ops.size() + (e - b) * (size_t)times);
        for (int t = 0; t < times; t++)
            for (size_t i = b; i < e; i++) ops.push_back(ops[i]);
    }
};

template <typename B, int BD, typename... P, typename... A>
void gang_record(void (*)(P...), dim3 grid, dim3 block, size_t shm, A &&...a)
{
    static_assert(sizeof...(P) == sizeof...(A), "argument count");
    const GangArgs<P...> args(static_cast<P>(a)...);
    t_rec->push(gang_kernel<B, BD, P...>(), grid.x, block.x, (uint32_t)shm, &args,
                (uint32_t)sizeof(args));
}

template <typename T>
struct unparen;
template <typename T>
struct unparen<void(T)> {
    using type = T;
};

}  // namespace ffr

// FFR_GANG((Body<...>), GB_*, kernel, grid, block, shm, stream, args...):
// the kernel's plain launch, or its record while a gang records.
#define FFR_GANG(BODY, BD, kernel, grid, block, shm, stream, ...)                               \
    do {                                                                                        \
        if (ffr::t_rec)                                                                         \
            ffr::gang_record<typename ffr::unparen<void BODY>::type, BD>(kernel, grid, block, shm, \
                                                                          __VA_ARGS__);         \
        else                                                                                    \
            FFR_LAUNCH(kernel, grid, block, shm, stream, __VA_ARGS__);                          \
    } while (0)
