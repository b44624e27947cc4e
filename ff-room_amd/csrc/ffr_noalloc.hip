// ffr_noalloc.hip - ffr_poisson_solve_noalloc: the reference CG pressure
// solve (PoissonSolver.cpp:93-158) of a scene, the same kernels and the same
// result as ffr_poisson_solve in the CG form, with every device buffer the
// solve needs beyond the scene itself (the right-hand side staged on the
// device, the CG vectors r, d and A d, the reduction partials and the
// iteration scalars) in one fixed array of this library instead of the
// scene's arena. It allocates nothing: no hipMalloc, no hipFree, no pinned
// memory (the scalars are read through the scene's existing pinned mirror).
//
// The array `scratch` is a file-scope __device__ array: each GPU the library
// runs on holds its own copy, at its own address, until the process ends.
// Host code never uses the array's name as an address: the first call on a
// device, with that device current, takes the device's address with
// hipGetSymbolAddress and keeps it per device; kernels get that pointer.
//
// Every call on a GPU shares that GPU's array and finds the previous call's
// contents there. Nothing needs to start at zero: k_cg_init writes x, r and
// d in full, k_cg_start writes every scalar, each reduction's producer writes
// every partial its consumer reads, and the SpMV writes A d in full. Calls
// are serialised per device by a lock and return only when their stream has
// drained, so calls from one process never overlap; two calls that did
// overlap on one GPU would race on the array.
#include "ffr_scene.hpp"

#include <mutex>

using namespace ffr;

namespace {

constexpr uint64_t kMaxCells = FFR_NOALLOC_MAX_CELLS;
constexpr size_t kStateBytes = 256;  // CgState, padded to the array's alignment
constexpr size_t kPartialBytes = 3 * (size_t)kMaxBlocksCap * sizeof(double);  // SLOT_INIT..UPDATE
constexpr size_t kVecBytes = kMaxCells * sizeof(double);
constexpr size_t kScratchBytes = kStateBytes + kPartialBytes + 4 * kVecBytes;  // b, r, d, A d
static_assert(sizeof(CgState) <= kStateBytes, "CgState outgrew its slot");
static_assert(SLOT_INIT < 3 && SLOT_SPMV < 3 && SLOT_UPDATE < 3, "CG slots outside the partials");
static_assert(kPartialBytes % 256 == 0, "vectors keep 256-byte alignment");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock[kMaxDevices];
void *g_scratch[kMaxDevices] = {};

}  // namespace

extern "C" {

int ffr_poisson_solve_noalloc(ffr_scene *s, const double *rhs, double *p, ffr_solve_info *info)
{
    if (!s || !rhs) return FFR_E_INVALID;
    ffr_ctx *c = s->ctx;
    if (c->device < 0 || c->device >= kMaxDevices) {
        c->fail("ffr_poisson_solve_noalloc: device %d beyond the %d this library tracks", c->device,
                kMaxDevices);
        return FFR_E_UNSUPPORTED;
    }
    if (s->nc > kMaxCells) {
        c->fail("ffr_poisson_solve_noalloc: %llu cells, above the fixed scratch's %llu "
                "(FFR_NOALLOC_MAX_CELLS); use ffr_poisson_solve",
                (unsigned long long)s->nc, (unsigned long long)kMaxCells);
        return FFR_E_INVALID;
    }
    if (resolve_method(s->prm) != FFR_POISSON_CG) {
        c->fail("ffr_poisson_solve_noalloc: the scene's pressure solver is not the CG form");
        return FFR_E_UNSUPPORTED;
    }
    if (task_busy(s)) {
        c->fail("ffr_poisson_solve_noalloc: the scene has a step in flight");
        return FFR_E_STATE;
    }
    std::lock_guard<std::mutex> hold(g_lock[c->device]);
    HIP_CHECK_CTX(c, hipSetDevice(c->device), FFR_E_DEVICE);
    if (!g_scratch[c->device]) {  // once per device, with that device current
        void *ptr = nullptr;
        HIP_CHECK_CTX(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), FFR_E_DEVICE);
        g_scratch[c->device] = ptr;
    }
    unsigned char *base = static_cast<unsigned char *>(g_scratch[c->device]);
    CgState *cg = reinterpret_cast<CgState *>(base);
    RedScratch red{reinterpret_cast<double *>(base + kStateBytes), nullptr};
    double *b = reinterpret_cast<double *>(base + kStateBytes + kPartialBytes);
    double *r = b + kMaxCells, *d = r + kMaxCells, *Ad = d + kMaxCells;

    hipStream_t st = s->stream;
    const Geo &g = s->geo;
    const int max_iter = s->prm.poisson_max_iter;
    HIP_CHECK_CTX(c, hipMemcpyAsync(b, rhs, s->nc * sizeof(double), hipMemcpyHostToDevice, st),
                  FFR_E_DEVICE);
    launch_cg_init(st, g, b, s->p, r, d, cg, red, max_iter, s->prm.poisson_tol);
    HIP_CHECK_CTX(c, hipGetLastError(), FFR_E_DEVICE);
    // iterations in chunks; kernels past the device-side stopping test return at once
    int enq = 0, chunk = 16;
    for (;;) {
        const int n = std::max(0, std::min(chunk, max_iter - enq));
        for (int m = 0; m < n; m++) {
            launch_cg_spmv(st, g, d, Ad, cg, red);
            launch_cg_update(st, g.nc, s->p, r, d, Ad, cg, red);
            launch_cg_dir(st, g.nc, d, r, cg, red);
        }
        HIP_CHECK_CTX(c, hipGetLastError(), FFR_E_DEVICE);
        enq += n;
        HIP_CHECK_CTX(c, hipMemcpyAsync(s->h_cg, cg, sizeof(CgState), hipMemcpyDeviceToHost, st),
                      FFR_E_DEVICE);
        HIP_CHECK_CTX(c, hipStreamSynchronize(st), FFR_E_DEVICE);
        if (s->h_cg->done || enq >= max_iter) break;
        chunk = std::max(8, enq / 4);
    }
    if (p)
        HIP_CHECK_CTX(c, hipMemcpyAsync(p, s->p, s->nc * sizeof(double), hipMemcpyDeviceToHost, st),
                      FFR_E_DEVICE);
    HIP_CHECK_CTX(c, hipStreamSynchronize(st), FFR_E_DEVICE);
    if (info) {
        info->residual = s->h_cg->res;
        info->iterations = s->h_cg->iters;
        info->outer_iterations = 0;
    }
    return FFR_OK;
}

}  // extern "C"
