// ffr_capi.hip - the C ABI (include/ffroom.h): contexts, device-resident
// scenes and the orchestration of one projection step on a HIP stream.
//
// One scene = one device arena (a single hipMalloc) holding the geometry
// (flags, per-axis coefficient arrays, boundary lists) and the state in the
// precision the parameters ask for. A step is a fixed sequence of kernel
// launches on the scene's stream; the only host waits are the pressure
// solver's convergence polls and the final 64-byte diagnostics copy that
// FluidSolver::step's StepResult needs (FluidSolver.cpp:549).
#include "ffr_scene.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

namespace ffr {
unsigned g_max_blocks = 2048;
unsigned g_resident_blocks = 2048;
thread_local ProfEvents g_prof_events;
}

using namespace ffr;

// ---- errors ----------------------------------------------------------------
void ffr_ctx::fail(const char *fmt, ...)
{
    std::lock_guard<std::mutex> hold(err_mu);
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, sizeof err, fmt, ap);
    va_end(ap);
}

static thread_local char g_noctx_err[256];

extern "C" {

int ffr_abi_version(void) { return FFR_ABI_VERSION; }

const char *ffr_status_string(int st)
{
    switch (st) {
    case FFR_OK: return "ok";
    case FFR_E_INVALID: return "invalid argument";
    case FFR_E_DEVICE: return "device error";
    case FFR_E_NOMEM: return "out of device memory or HBM budget";
    case FFR_E_STATE: return "invalid state";
    case FFR_E_NUMERIC: return "non-finite values";
    case FFR_E_UNSUPPORTED: return "unsupported option";
    case FFR_E_NOT_FOUND: return "unknown ticket";
    default: return "unknown status";
    }
}

void ffr_params_default(ffr_params *p)
{
    // FluidSolver.hpp:11-38, PoissonSolver.hpp:14-18
    std::memset(p, 0, sizeof *p);
    p->rho = 1.2;
    p->nu = 1.5e-5;
    p->dt = 0.01;
    p->max_steps = 500;
    p->convergence_tol = 1.0;
    p->poisson_max_iter = 1000;
    p->poisson_tol = 1e-6;
    p->advection = FFR_UPWIND1;
    p->T_initial = 20.0;
    p->T_target = 1e30;
    p->g_accel = 9.81;
    p->beta = 1.0 / 293.0;
    p->k_thermal = 0.026;
    p->cp = 1005.0;
    p->precision = FFR_PREC_FP64;
    p->poisson_method = FFR_POISSON_AUTO;
    p->mg_levels = 0;
    p->mg_bf16 = 1;
    p->mg_pre_sweeps = 2;
    p->mg_level_sweeps = 2;
    p->mg_coarse_sweeps = 8;
    p->inner_tol = 1e-4;
    p->inner_max_iter = 60;
    p->graphs = 1;
    p->mg_tail_kb = 0;
    p->warm_start = 3;
    p->mg_tail_cells = 0;
    p->mg_tail_groups = 0;
}

int ffr_ctx_create(int device, size_t hbm_budget, ffr_ctx **out)
{
    if (!out) return FFR_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        snprintf(g_noctx_err, sizeof g_noctx_err, "no HIP device available");
        return FFR_E_DEVICE;
    }
    if (device < 0 || device >= n) {
        snprintf(g_noctx_err, sizeof g_noctx_err, "device %d out of range (%d devices)", device, n);
        return FFR_E_INVALID;
    }
    auto *c = new (std::nothrow) ffr_ctx();
    if (!c) return FFR_E_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return FFR_E_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return FFR_E_DEVICE;
    }
    c->prop = prop;
    // blocks of kBlock threads resident at once at 8 waves per SIMD (4 SIMDs
    // per CU): the grid cap of the level-0 gather kernels
    ffr::g_resident_blocks = (unsigned)std::max(1, prop.multiProcessorCount) * 8u * 4u /
                             (unsigned)(kBlock / kWave);
    c->budget = hbm_budget ? hbm_budget : (size_t)prop.totalGlobalMem * 9 / 10;
    *out = c;
    return FFR_OK;
}

void ffr_ctx_destroy(ffr_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->device);
    c->serve_shutdown();
    // scenes still alive go with their context
    const std::set<ffr_scene *> left = c->scenes;
    for (ffr_scene *s : left) ffr_scene_destroy(s);
    c->pool_trim();
    c->prof_clear();
    delete c;
}

const char *ffr_last_error(const ffr_ctx *c) { return c ? c->err : g_noctx_err; }

size_t ffr_ctx_bytes_in_use(const ffr_ctx *c) { return c ? c->in_use : 0; }

int ffr_device_query(ffr_ctx *c, ffr_device_info *o)
{
    if (!c || !o) return FFR_E_INVALID;
    std::memset(o, 0, sizeof *o);
    std::snprintf(o->name, sizeof o->name, "%s", c->prop.name);
    std::snprintf(o->arch, sizeof o->arch, "%s", c->prop.gcnArchName);
    o->compute_units = c->prop.multiProcessorCount;
    o->xcds = kXcds;
    o->hbm_bytes = c->prop.totalGlobalMem;
    o->l2_bytes = c->prop.l2CacheSize;
    o->lds_bytes_per_cu = c->prop.maxSharedMemoryPerMultiProcessor;
    o->clock_khz = c->prop.clockRate;
    return FFR_OK;
}

}  // extern "C"

// ---- profiling (HIP events around launches on the scene's stream) -----------
void ffr_ctx::prof_begin(hipStream_t s, const char *name, double bytes)
{
    if (!prof_on) return;
    PendingEvent pe;
    if (!ev_pool.empty()) {
        pe.a = ev_pool.back(); ev_pool.pop_back();
    } else {
        hipEventCreate(&pe.a);
    }
    if (!ev_pool.empty()) {
        pe.b = ev_pool.back(); ev_pool.pop_back();
    } else {
        hipEventCreate(&pe.b);
    }
    pe.name = name;
    pe.bytes = bytes;
    pending.push_back(pe);
    open_event = true;
    // the next launch on this thread records a/b at its dispatch begin/end
    ffr::g_prof_events.start = pe.a;
    ffr::g_prof_events.stop = pe.b;
}

void ffr_ctx::prof_end(hipStream_t s)
{
    if (!prof_on || !open_event) return;
    if (ffr::g_prof_events.start) {  // the scope launched nothing: record an empty interval
        ffr::g_prof_events = ffr::ProfEvents{};
        hipEventRecord(pending.back().a, s);
        hipEventRecord(pending.back().b, s);
    }
    open_event = false;
    if (pending.size() > 4096) prof_resolve();
}

void ffr_ctx::prof_resolve()
{
    for (auto &pe : pending) {
        hipEventSynchronize(pe.b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, pe.a, pe.b);
        auto &e = prof_map[pe.name];
        e.launches++;
        e.total_ms += ms;
        e.bytes += pe.bytes;  // launches of one name may move different bytes
        ev_pool.push_back(pe.a);
        ev_pool.push_back(pe.b);
    }
    pending.clear();
}

void ffr_ctx::prof_clear()
{
    prof_resolve();
    for (auto ev : ev_pool) hipEventDestroy(ev);
    ev_pool.clear();
    prof_map.clear();
}

extern "C" {
int ffr_prof_enable(ffr_ctx *c, int on)
{
    if (!c) return FFR_E_INVALID;
    if (!on) c->prof_resolve();
    c->prof_on = on != 0;
    return FFR_OK;
}
int ffr_prof_reset(ffr_ctx *c)
{
    if (!c) return FFR_E_INVALID;
    c->prof_resolve();
    c->prof_map.clear();
    return FFR_OK;
}
int ffr_prof_read(ffr_ctx *c, ffr_prof_entry *out, int cap, int *n_out)
{
    if (!c || !n_out) return FFR_E_INVALID;
    c->prof_resolve();
    int n = 0;
    for (auto &kv : c->prof_map) {
        if (out && n < cap) {
            std::memset(&out[n], 0, sizeof out[n]);
            std::snprintf(out[n].name, sizeof out[n].name, "%s", kv.first.c_str());
            out[n].launches = kv.second.launches;
            out[n].total_ms = kv.second.total_ms;
            out[n].bytes_per_launch =
                kv.second.launches ? kv.second.bytes / (double)kv.second.launches : 0.0;
        }
        n++;
    }
    *n_out = n;
    return FFR_OK;
}

int ffr_set_max_blocks(int max_blocks)
{
    if (max_blocks < 256 || max_blocks > (int)ffr::kMaxBlocksCap) return FFR_E_INVALID;
    ffr::g_max_blocks = (unsigned)max_blocks;
    return FFR_OK;
}

int ffr_get_max_blocks(void) { return (int)ffr::g_max_blocks; }
}  // extern "C"
