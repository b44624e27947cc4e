// ffr_gang.hip - merged launches of several scenes' step tasks (ffr_gang.hpp).
//
// A round: every member scene advances its own step task (ffr_task.hip) with
// a recorder active, exactly as it would on its own stream, and its kernel
// launches and its copy of device scalars to pinned host memory are recorded
// instead of issued. The recorded sequences are then merged greedily: the
// next merged launch takes the head operation of the member with the most
// operations left, together with every other member whose head is the same
// kernel; each member's operations keep their order, and members are
// independent, so any such merge computes what the members' own launches
// compute. All tables of a round go to the device in one copy, the merged
// launches follow on the gang's stream, the scalar copies become one gather
// kernel, and one event ends the round.
#include "ffr_gang.hpp"
#include "ffr_scene.hpp"

#include <algorithm>
#include <string>
#include <unordered_map>

namespace ffr {

thread_local Recorder *t_rec = nullptr;

void rec_unsupported()
{
    if (t_rec) t_rec->unsupported = true;
}

// ---- the gather of device scalars into pinned host memory ------------------
struct GatherEnt {
    const uint32_t *src;
    uint32_t *dst;
    uint32_t words, pad;
};
struct GatherHdr {
    uint32_t n, pad[3];
};

// block b copies entry b
__global__ void __launch_bounds__(kBlock) k_gang_gather(const GatherHdr *h)
{
    const GatherEnt *e = reinterpret_cast<const GatherEnt *>(h + 1) + blockIdx.x;
    const uint32_t *src = e->src;
    uint32_t *dst = e->dst;
    for (uint32_t i = threadIdx.x; i < e->words; i += kBlock) dst[i] = src[i];
}

namespace {

struct Launch {
    const void *fn;  // nullptr: the gather
    uint32_t grid, block, lds;
    size_t toff;
};

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

Gang::~Gang() { release(); }

void Gang::release()
{
    if (stream) hipStreamSynchronize(stream);
    if (ev) hipEventDestroy(ev);
    if (stream) hipStreamDestroy(stream);
    if (d_tab) hipFree(d_tab);
    if (h_stage) hipHostFree(h_stage);
    ev = nullptr;
    stream = nullptr;
    d_tab = nullptr;
    h_stage = nullptr;
    cap = 0;
}

int Gang::init(ffr_ctx *c)
{
    ctx = c;
    HIP_CHECK_CTX(c, hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), FFR_E_DEVICE);
    HIP_CHECK_CTX(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming), FFR_E_DEVICE);
    return FFR_OK;
}

int Gang::reserve(size_t bytes)
{
    if (bytes <= cap) return FFR_OK;
    size_t nc = std::max<size_t>(cap ? cap * 2 : (size_t)1 << 20, bytes);
    // the previous round has completed (the caller waited for its event)
    if (d_tab) hipFree(d_tab);
    if (h_stage) hipHostFree(h_stage);
    d_tab = nullptr;
    h_stage = nullptr;
    cap = 0;
    HIP_CHECK_CTX(ctx, hipMalloc((void **)&d_tab, nc), FFR_E_NOMEM);
    HIP_CHECK_CTX(ctx, hipHostMalloc((void **)&h_stage, nc), FFR_E_NOMEM);
    cap = nc;
    return FFR_OK;
}

// Merges the members' recorded operations into launches, uploads their
// tables and issues them on the gang's stream, then records the event.
int Gang::launch_round(const std::vector<Recorder *> &recs)
{
    const int n = (int)recs.size();
    std::vector<size_t> head(n, 0);
    std::vector<Launch> launches;
    std::vector<uint8_t> &stage = stage_buf;
    stage.clear();
    std::unordered_map<std::string, size_t> seen;  // identical tables are uploaded once
    std::string key;
    std::vector<int> batch;
    batch.reserve(n);

    auto remaining = [&](int i) { return recs[i]->ops.size() - head[i]; };
    for (;;) {
        int best = -1;
        for (int i = 0; i < n; i++)
            if (remaining(i) > 0 && (best < 0 || remaining(i) > remaining(best))) best = i;
        if (best < 0) break;
        const GangOp &ref = recs[best]->ops[head[best]];
        batch.clear();
        for (int i = 0; i < n && (int)batch.size() < kGangMax; i++) {
            if (remaining(i) == 0) continue;
            const GangOp &op = recs[i]->ops[head[i]];
            if (op.fn == ref.fn && op.block == ref.block) batch.push_back(i);
        }
        key.assign(reinterpret_cast<const char *>(&ref.fn), sizeof(ref.fn));
        for (int i : batch) {
            const GangOp &op = recs[i]->ops[head[i]];
            const uintptr_t id[2] = {(uintptr_t)recs[i], ref.fn ? (uintptr_t)op.aoff : (uintptr_t)op.dsrc};
            key.append(reinterpret_cast<const char *>(id), sizeof(id));
        }
        size_t toff;
        uint32_t grid = 0, lds = 0;
        if (ref.fn) {
            const uint32_t stride = (uint32_t)align_up(ref.asize, 16);
            GangHdr h{};
            h.n = (uint32_t)batch.size();
            h.stride = stride;
            uint32_t off = 0;
            for (size_t j = 0; j < batch.size(); j++) {
                const GangOp &op = recs[batch[j]]->ops[head[batch[j]]];
                h.off[j] = off;
                h.G[j] = op.grid;
                off += (uint32_t)align_up(op.grid, kXcds);
                lds = std::max(lds, op.lds);
            }
            for (int j = (int)batch.size(); j <= kGangMax; j++) h.off[j] = off;
            grid = off;
            auto it = seen.find(key);
            if (it != seen.end()) {
                toff = it->second;
            } else {
                toff = align_up(stage.size(), 256);
                stage.resize(toff + kGangArgsOff + (size_t)stride * batch.size(), 0);
                std::memcpy(stage.data() + toff, &h, sizeof(h));
                for (size_t j = 0; j < batch.size(); j++) {
                    const Recorder *r = recs[batch[j]];
                    const GangOp &op = r->ops[head[batch[j]]];
                    std::memcpy(stage.data() + toff + kGangArgsOff + j * stride, r->arena.data() + op.aoff,
                                op.asize);
                }
                seen.emplace(key, toff);
            }
        } else {
            toff = align_up(stage.size(), 256);
            stage.resize(toff + sizeof(GatherHdr) + sizeof(GatherEnt) * batch.size(), 0);
            GatherHdr gh{};
            gh.n = (uint32_t)batch.size();
            std::memcpy(stage.data() + toff, &gh, sizeof(gh));
            for (size_t j = 0; j < batch.size(); j++) {
                const GangOp &op = recs[batch[j]]->ops[head[batch[j]]];
                if (op.bytes % 4 != 0) {
                    ctx->fail("gang: a scalar copy of %u bytes is not whole words", op.bytes);
                    return FFR_E_INVALID;
                }
                GatherEnt e{(const uint32_t *)op.dsrc, (uint32_t *)op.hdst, op.bytes / 4, 0};
                std::memcpy(stage.data() + toff + sizeof(GatherHdr) + j * sizeof(GatherEnt), &e, sizeof(e));
            }
            grid = (uint32_t)batch.size();
        }
        launches.push_back(Launch{ref.fn, grid, ref.block, lds, toff});
        for (int i : batch) head[i]++;
    }
    if (launches.empty()) return FFR_OK;

    int rc = reserve(stage.size());
    if (rc != FFR_OK) return rc;
    std::memcpy(h_stage, stage.data(), stage.size());
    HIP_CHECK_CTX(ctx, hipSetDevice(ctx->device), FFR_E_DEVICE);
    HIP_CHECK_CTX(ctx, hipMemcpyAsync(d_tab, h_stage, stage.size(), hipMemcpyHostToDevice, stream),
                  FFR_E_DEVICE);
    for (const Launch &l : launches) {
        const void *tab = d_tab + l.toff;
        void *args[] = {(void *)&tab};
        if (l.fn) {
            HIP_CHECK_CTX(ctx, hipLaunchKernel(l.fn, dim3(l.grid), dim3(l.block), args, l.lds, stream),
                          FFR_E_DEVICE);
        } else {
            const GatherHdr *gh = (const GatherHdr *)tab;
            hipLaunchKernelGGL(k_gang_gather, dim3(l.grid), dim3(kBlock), 0, stream, gh);
            HIP_CHECK_CTX(ctx, hipGetLastError(), FFR_E_DEVICE);
        }
    }
    rounds++;
    merged_launches += launches.size();
    size_t longest = 0;
    for (const Recorder *r : recs) {
        recorded_ops += r->ops.size();
        longest = std::max(longest, r->ops.size());
    }
    longest_ops += longest;
    HIP_CHECK_CTX(ctx, hipEventRecord(ev, stream), FFR_E_DEVICE);
    return FFR_OK;
}

}  // namespace ffr
