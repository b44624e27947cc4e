// ffr_snapshot.hip - snapshot egress: a scene's fields copied to the host
// while later steps run (the captures of SolverBridge.run_animated,
// solver_bridge.py:172-281, which copy the fields between steps).
//
// ffr_snapshot_begin enqueues, on the scene's stream, a device-to-device copy
// of u, v, w, T (in the state's precision) and p into a slot of the scene;
// steps enqueued afterwards may overwrite the live fields at once. A second
// stream then copies the slot to pinned host memory, overlapping the next
// steps' kernels. ffr_snapshot_end waits for that copy, writes FP64 fields in
// the reference layout and frees the slot. Nothing here runs inside a graph.
#include "ffr_scene.hpp"

#include <algorithm>

using namespace ffr;

namespace ffr {
// frees a scene's snapshot slots (with its stream idle)
void snapshot_release(ffr_scene *s)
{
    if (s->copy_stream) hipStreamSynchronize(s->copy_stream);
    for (auto &q : s->snaps) {
        if (q.copied) hipEventDestroy(q.copied);
        if (q.landed) hipEventDestroy(q.landed);
        if (q.dev) {
            hipFree(q.dev);
            s->ctx->in_use -= s->snap_bytes;
        }
        if (q.host) hipHostFree(q.host);
        q = ffr_scene::Snap{};
    }
    if (s->copy_stream) hipStreamDestroy(s->copy_stream);
    s->copy_stream = nullptr;
}
}  // namespace ffr

namespace {
// field q (u, v, w, T, p): element count, bytes per element, offset in a slot
struct SnapLayout {
    uint64_t cnt[5];
    size_t esz[5], off[5], total;
};
SnapLayout snap_layout(const ffr_scene *s)
{
    SnapLayout L{};
    const size_t r = s->fp64 ? 8 : 4;
    const uint64_t cnt[5] = {s->nu, s->nv, s->nw, s->nc, s->nc};
    const size_t esz[5] = {r, r, r, r, 8};
    size_t off = 0;
    for (int q = 0; q < 5; q++) {
        L.cnt[q] = cnt[q];
        L.esz[q] = esz[q];
        L.off[q] = off;
        off += (cnt[q] * esz[q] + 255) & ~(size_t)255;
    }
    L.total = off;
    return L;
}
}  // namespace

extern "C" {

int ffr_snapshot_begin(ffr_scene *s, int *slot)
{
    if (!s || !slot) return FFR_E_INVALID;
    ffr_ctx *c = s->ctx;
    if (task_busy(s)) {
        c->fail("ffr_snapshot_begin: the scene has a step in flight");
        return FFR_E_STATE;
    }
    int k = -1;
    for (int i = 0; i < FFR_SNAPSHOT_SLOTS; i++)
        if (!s->snaps[i].busy) {
            k = i;
            break;
        }
    if (k < 0) {
        c->fail("ffr_snapshot_begin: %d snapshots are pending; end one first", FFR_SNAPSHOT_SLOTS);
        return FFR_E_STATE;
    }
    HIP_CHECK_CTX(c, hipSetDevice(c->device), FFR_E_DEVICE);
    const SnapLayout L = snap_layout(s);
    ffr_scene::Snap &q = s->snaps[k];
    if (!s->copy_stream)
        HIP_CHECK_CTX(c, hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking), FFR_E_DEVICE);
    if (!q.dev) {  // slots are allocated on first use and count against the budget
        if (c->in_use + L.total > c->budget) {
            c->fail("ffr_snapshot_begin: a slot needs %zu bytes, the context budget has %zu free",
                    L.total, c->budget - c->in_use);
            return FFR_E_NOMEM;
        }
        HIP_CHECK_CTX(c, hipMalloc(&q.dev, L.total), FFR_E_NOMEM);
        c->in_use += L.total;
        s->snap_bytes = L.total;
        HIP_CHECK_CTX(c, hipHostMalloc(&q.host, L.total), FFR_E_NOMEM);
        HIP_CHECK_CTX(c, hipEventCreateWithFlags(&q.copied, hipEventDisableTiming), FFR_E_DEVICE);
        HIP_CHECK_CTX(c, hipEventCreateWithFlags(&q.landed, hipEventDisableTiming), FFR_E_DEVICE);
    }
    const void *src[5] = {s->u, s->v, s->w, s->T, s->p};
    for (int f = 0; f < 5; f++)
        HIP_CHECK_CTX(c, hipMemcpyAsync((char *)q.dev + L.off[f], src[f], L.cnt[f] * L.esz[f],
                                        hipMemcpyDeviceToDevice, s->stream), FFR_E_DEVICE);
    HIP_CHECK_CTX(c, hipEventRecord(q.copied, s->stream), FFR_E_DEVICE);
    HIP_CHECK_CTX(c, hipStreamWaitEvent(s->copy_stream, q.copied, 0), FFR_E_DEVICE);
    HIP_CHECK_CTX(c, hipMemcpyAsync(q.host, q.dev, L.total, hipMemcpyDeviceToHost, s->copy_stream),
                  FFR_E_DEVICE);
    HIP_CHECK_CTX(c, hipEventRecord(q.landed, s->copy_stream), FFR_E_DEVICE);
    q.busy = true;
    q.step = s->step;
    *slot = k;
    return FFR_OK;
}

int ffr_snapshot_end(ffr_scene *s, int slot, int *step, double *u, double *v, double *w,
                     double *p, double *T)
{
    if (!s || slot < 0 || slot >= FFR_SNAPSHOT_SLOTS) return FFR_E_INVALID;
    ffr_ctx *c = s->ctx;
    ffr_scene::Snap &q = s->snaps[slot];
    if (!q.busy) {
        c->fail("ffr_snapshot_end: slot %d holds no snapshot", slot);
        return FFR_E_STATE;
    }
    HIP_CHECK_CTX(c, hipSetDevice(c->device), FFR_E_DEVICE);
    HIP_CHECK_CTX(c, hipEventSynchronize(q.landed), FFR_E_DEVICE);
    const SnapLayout L = snap_layout(s);
    double *dst[5] = {u, v, w, T, p};
    for (int f = 0; f < 5; f++) {
        if (!dst[f]) continue;
        const char *h = (const char *)q.host + L.off[f];
        if (L.esz[f] == 8) {
            std::copy((const double *)h, (const double *)h + L.cnt[f], dst[f]);
        } else {
            const float *x = (const float *)h;
            for (uint64_t i = 0; i < L.cnt[f]; i++) dst[f][i] = (double)x[i];
        }
    }
    if (step) *step = q.step;
    q.busy = false;
    return FFR_OK;
}

}  // extern "C"
