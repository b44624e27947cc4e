// ffr_snapshot.hip - snapshot egress: a scene's fields copied to the host
// while later steps run (the captures of SolverBridge.run_animated,
// solver_bridge.py:172-281, which copy the fields between steps).
//
// ffr_snapshot_begin enqueues, on the scene's stream, a copy of the fields
// into a slot of the scene: u, v, w, T (in the state's precision) and p
// (FFR_SNAP_FIELDS), or the quantities of a FieldSnapshot computed from them
// in FP64 (FFR_SNAP_CELLS: cell-centred velocity, speed, T); steps enqueued
// afterwards may overwrite the live fields at once. A second stream then
// copies the slot to pinned host memory, overlapping the next steps' kernels.
// ffr_snapshot_end waits for that copy, writes FP64 arrays and frees the
// slot. Nothing here runs inside a graph.
#include "ffr_scene.hpp"

#include <algorithm>
#include <cstring>

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
// Cell-centred velocity, speed and T of one cell in FP64, with the
// operations of the reference's capture (solver_bridge.py:224-236: 0.5 *
// (a + b) per axis, sqrt(uc**2 + vc**2 + wc**2)) in that order and without
// fused multiply-adds, so the values are numpy's bit for bit. vel is
// (nx, ny, nz, 3) interleaved, as np.stack(..., axis=-1) lays it out.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_snap_cells(Geo g, const R *u, const R *v, const R *w,
                                                        const R *T, double *vel, double *speed,
                                                        double *Tc)
{
#pragma clang fp contract(off)
    const uint32_t nc = g.nc;
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
        int i, j, k;
        const uint32_t ii = fdiv(c, g.f_nynz), rem = c - ii * (uint32_t)(g.ny * g.nz);
        const uint32_t jj = fdiv(rem, g.f_nz);
        i = ii;
        j = jj;
        k = rem - jj * g.nz;
        const double uc = 0.5 * ((double)u[uidx(g, i, j, k)] + (double)u[uidx(g, i + 1, j, k)]);
        const double vc = 0.5 * ((double)v[vidx(g, i, j, k)] + (double)v[vidx(g, i, j + 1, k)]);
        const double wc = 0.5 * ((double)w[widx(g, i, j, k)] + (double)w[widx(g, i, j, k + 1)]);
        vel[3ull * c] = uc;
        vel[3ull * c + 1] = vc;
        vel[3ull * c + 2] = wc;
        const double uu = uc * uc, vv = vc * vc, ww = wc * wc;
        speed[c] = sqrt((uu + vv) + ww);
        Tc[c] = (double)T[c];
    }
}

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
// cell kind: vel (3 nc), speed (nc), T (nc), FP64
size_t cells_bytes(const ffr_scene *s) { return 5ull * 8ull * s->nc; }
}  // namespace

extern "C" {

int ffr_snapshot_begin(ffr_scene *s, int kind, int *slot)
{
    if (!s || !slot || (kind != FFR_SNAP_FIELDS && kind != FFR_SNAP_CELLS)) return FFR_E_INVALID;
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
    const size_t need = kind == FFR_SNAP_CELLS ? cells_bytes(s) : L.total;
    ffr_scene::Snap &q = s->snaps[k];
    if (!s->copy_stream)
        HIP_CHECK_CTX(c, hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking), FFR_E_DEVICE);
    if (!q.dev) {  // slots are allocated on first use (for either kind) and count against the budget
        const size_t bytes = std::max(L.total, cells_bytes(s));
        if (c->in_use + bytes > c->budget) {
            c->fail("ffr_snapshot_begin: a slot needs %zu bytes, the context budget has %zu free",
                    bytes, c->budget - c->in_use);
            return FFR_E_NOMEM;
        }
        // taken into locals and given to the slot only when every part exists
        void *dev = nullptr, *host = nullptr;
        hipEvent_t copied = nullptr, landed = nullptr;
        hipError_t e = hipMalloc(&dev, bytes);
        if (e == hipSuccess) e = hipHostMalloc(&host, bytes);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&copied, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&landed, hipEventDisableTiming);
        if (e != hipSuccess) {
            if (landed) hipEventDestroy(landed);
            if (copied) hipEventDestroy(copied);
            if (host) hipHostFree(host);
            if (dev) hipFree(dev);
            c->fail("ffr_snapshot_begin: allocating a %zu-byte slot failed: %s", bytes,
                    hipGetErrorString(e));
            return FFR_E_NOMEM;
        }
        q.dev = dev;
        q.host = host;
        q.copied = copied;
        q.landed = landed;
        c->in_use += bytes;
        s->snap_bytes = bytes;
    }
    if (kind == FFR_SNAP_CELLS) {
        double *vel = (double *)q.dev, *speed = vel + 3 * s->nc, *Tc = speed + s->nc;
        const dim3 grid(grid_for(s->nc)), block(kBlock);
        if (s->fp64)
            hipLaunchKernelGGL(k_snap_cells<double>, grid, block, 0, s->stream, s->geo,
                               (const double *)s->u, (const double *)s->v, (const double *)s->w,
                               (const double *)s->T, vel, speed, Tc);
        else
            hipLaunchKernelGGL(k_snap_cells<float>, grid, block, 0, s->stream, s->geo,
                               (const float *)s->u, (const float *)s->v, (const float *)s->w,
                               (const float *)s->T, vel, speed, Tc);
        HIP_CHECK_CTX(c, hipGetLastError(), FFR_E_DEVICE);
    } else {
        const void *src[5] = {s->u, s->v, s->w, s->T, s->p};
        for (int f = 0; f < 5; f++)
            HIP_CHECK_CTX(c, hipMemcpyAsync((char *)q.dev + L.off[f], src[f], L.cnt[f] * L.esz[f],
                                            hipMemcpyDeviceToDevice, s->stream), FFR_E_DEVICE);
    }
    HIP_CHECK_CTX(c, hipEventRecord(q.copied, s->stream), FFR_E_DEVICE);
    HIP_CHECK_CTX(c, hipStreamWaitEvent(s->copy_stream, q.copied, 0), FFR_E_DEVICE);
    HIP_CHECK_CTX(c, hipMemcpyAsync(q.host, q.dev, need, hipMemcpyDeviceToHost, s->copy_stream),
                  FFR_E_DEVICE);
    HIP_CHECK_CTX(c, hipEventRecord(q.landed, s->copy_stream), FFR_E_DEVICE);
    q.busy = true;
    q.kind = kind;
    q.step = s->step;
    *slot = k;
    return FFR_OK;
}

int ffr_snapshot_end(ffr_scene *s, int slot, int *step, double *a, double *b, double *c3,
                     double *d, double *e)
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
    if (q.kind == FFR_SNAP_CELLS) {  // a: vel (3 nc), b: speed, c3: T
        const double *h = (const double *)q.host;
        double *dst[3] = {a, b, c3};
        const size_t off[3] = {0, 3 * s->nc, 4 * s->nc}, cnt[3] = {3 * s->nc, s->nc, s->nc};
        for (int f = 0; f < 3; f++)
            if (dst[f]) std::memcpy(dst[f], h + off[f], cnt[f] * sizeof(double));
        if (step) *step = q.step;
        q.busy = false;
        return FFR_OK;
    }
    const SnapLayout L = snap_layout(s);
    double *dst[5] = {a, b, c3, e, d};  // u, v, w, T, p
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
