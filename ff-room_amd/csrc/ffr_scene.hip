// ffr_scene.hip - scene creation, host transfers and the projection step.
#include "ffr_scene.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <new>
#include <string>
#include <vector>

using namespace ffr;

namespace {

bool valid_desc(const ffr_scene_desc *d, char *err, size_t errn)
{
    if (!d) { snprintf(err, errn, "null scene descriptor"); return false; }
    if (d->nx < 2 || d->ny < 2 || d->nz < 2) {
        snprintf(err, errn, "grid dimensions must be >= 2 (got %d x %d x %d)", d->nx, d->ny, d->nz);
        return false;
    }
    const uint64_t faces = (uint64_t)(d->nx + 1) * d->ny * d->nz +
                           (uint64_t)d->nx * (d->ny + 1) * d->nz +
                           (uint64_t)d->nx * d->ny * (d->nz + 1);
    if (faces >= (1ull << 31)) { snprintf(err, errn, "grid too large for 32-bit indexing"); return false; }
    // the kernels read fields through buffer resources with 32-bit byte offsets
    const uint64_t most = std::max({(uint64_t)(d->nx + 1) * d->ny * d->nz,
                                    (uint64_t)d->nx * (d->ny + 1) * d->nz,
                                    (uint64_t)d->nx * d->ny * (d->nz + 1)});
    if (most * sizeof(double) >= (1ull << 31)) {
        snprintf(err, errn, "grid too large: a face array of %llu elements exceeds 2 GiB in FP64",
                 (unsigned long long)most);
        return false;
    }
    if (d->n_fans > kMaxFans) { snprintf(err, errn, "at most %d fans per scene", kMaxFans); return false; }
    if (!d->dxv || !d->dyv || !d->dzv) {
        snprintf(err, errn, "spacings are required");
        return false;
    }
    if (!d->cell_type) {  // flags from marks, built on the device
        if (d->n_marks < 0 || d->n_marks > FFR_MAX_MARKS || (d->n_marks && !d->marks)) {
            snprintf(err, errn, "cell_type or 0..%d marks are required", FFR_MAX_MARKS);
            return false;
        }
        for (int q = 0; q < d->n_marks; q++)
            if (d->marks[q].type < 0 || d->marks[q].type > 3) {
                snprintf(err, errn, "mark %d: type %d out of range", q, d->marks[q].type);
                return false;
            }
    }
    const int n[3] = {d->nx, d->ny, d->nz};
    const double *h[3] = {d->dxv, d->dyv, d->dzv};
    for (int a = 0; a < 3; a++)
        for (int i = 0; i < n[a]; i++)
            if (!(h[a][i] > 0.0) || !std::isfinite(h[a][i])) {
                snprintf(err, errn, "spacing %d of axis %d is not a positive finite number", i, a);
                return false;
            }
    if ((d->n_fans && !d->fans) || (d->n_outflows && !d->outflows) ||
        (d->n_openings && !d->openings) || d->n_fans < 0 || d->n_outflows < 0 || d->n_openings < 0) {
        snprintf(err, errn, "boundary list missing or negative count");
        return false;
    }
    auto check_patch = [&](int axis, int side, int a0, int a1, int b0, int b1, const char *what) {
        if (axis < 0 || axis > 2 || (side != 0 && side != 1)) {
            snprintf(err, errn, "%s: axis must be 0..2 and side 0/1", what);
            return false;
        }
        const int ta = axis == 0 ? 1 : 0, tb = axis == 2 ? 1 : 2;
        if (a0 < 0 || b0 < 0 || a1 >= n[ta] || b1 >= n[tb]) {
            // the reference indexes these ranges unchecked (BoundaryManager.cpp:24-44, 180-220)
            snprintf(err, errn, "%s patch [%d,%d]x[%d,%d] outside the wall", what, a0, a1, b0, b1);
            return false;
        }
        return true;
    };
    for (int q = 0; q < d->n_outflows; q++) {
        const ffr_outflow &o = d->outflows[q];
        if (!check_patch(o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max, "outflow")) return false;
    }
    for (int q = 0; q < d->n_openings; q++) {
        const ffr_opening &o = d->openings[q];
        if (!check_patch(o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max, "opening")) return false;
    }
    if (d->cell_type) {
        const long nc = (long)d->nx * d->ny * d->nz;
        for (long c = 0; c < nc; c++)
            if (d->cell_type[c] > 3) { snprintf(err, errn, "cell_type[%ld] out of range", c); return false; }
    }
    return true;
}

bool valid_params(const ffr_params *p, char *err, size_t errn)
{
    if (!p) { snprintf(err, errn, "null params"); return false; }
    if (!(p->dt > 0.0) || !(p->rho > 0.0)) { snprintf(err, errn, "dt and rho must be > 0"); return false; }
    if (p->advection < 0 || p->advection > 2) { snprintf(err, errn, "unknown advection scheme"); return false; }
    if (p->precision != FFR_PREC_FP64 && p->precision != FFR_PREC_MIXED) {
        snprintf(err, errn, "unknown precision");
        return false;
    }
    if (p->poisson_method < 0 || p->poisson_method > 2) {
        snprintf(err, errn, "unknown poisson_method");
        return false;
    }
    return true;
}

struct ArenaPtrs {
    uint8_t *flags;
    uint16_t *nbm;
    uint16_t *fcode;
    double *ax64;
    float *ax32;
    FanDev *fans;
    OpeningDev *opens;
    uint8_t *wall;
    char *u, *v, *w, *us, *vs, *ws, *T, *Tn;
    double *p, *b, *r, *d, *Ad, *scratch, *p_hist, *p_hist2;
    StepRec *rec;
    CgState *cg;
    double *partials;
    unsigned int *tickets;
};

// Layout of a scene's arena (sizing pass with base == nullptr).
size_t layout(char *base, const ffr_scene_desc *d, const ffr_params *p, ArenaPtrs *out)
{
    size_t off = 0;
    auto take = [&](size_t bytes) -> char * {
        char *ptr = base ? base + off : nullptr;
        off = (off + bytes + 255) & ~(size_t)255;
        return ptr;
    };
    const uint64_t nx = d->nx, ny = d->ny, nz = d->nz;
    const uint64_t nu = (nx + 1) * ny * nz, nv = nx * (ny + 1) * nz, nw = nx * ny * (nz + 1);
    const uint64_t nc = nx * ny * nz, na = nx + ny + nz;
    const size_t s = (p->precision == FFR_PREC_MIXED) ? sizeof(float) : sizeof(double);
    ArenaPtrs a{};
    a.flags = (uint8_t *)take(nc);
    a.nbm = (uint16_t *)take(2 * nc);
    a.fcode = (uint16_t *)take(2 * (nu + nv + nw));
    a.ax64 = (double *)take(sizeof(double) * 8 * na);
    a.ax32 = (float *)take(sizeof(float) * 10 * na);
    a.fans = (FanDev *)take(sizeof(FanDev) * std::max(1, d->n_fans));
    a.opens = (OpeningDev *)take(sizeof(OpeningDev) * std::max(1, d->n_openings));
    a.wall = (uint8_t *)take(2 * (ny * nz + nx * nz + nx * ny));
    a.u = take(nu * s); a.v = take(nv * s); a.w = take(nw * s);
    a.us = take(nu * s); a.vs = take(nv * s); a.ws = take(nw * s);
    a.T = take(nc * s); a.Tn = take(nc * s);
    a.p = (double *)take(8 * nc);
    a.b = (double *)take(8 * nc);
    a.r = (double *)take(8 * nc);
    a.d = (double *)take(8 * nc);
    a.Ad = (double *)take(8 * nc);
    a.scratch = (p->precision == FFR_PREC_MIXED) ? (double *)take(8 * std::max({nu, nv, nw})) : nullptr;
    a.p_hist = (p->precision == FFR_PREC_MIXED) ? (double *)take(8 * nc) : nullptr;
    a.p_hist2 = (p->precision == FFR_PREC_MIXED) ? (double *)take(8 * nc) : nullptr;
    a.rec = (StepRec *)take(sizeof(StepRec));
    a.cg = (CgState *)take(sizeof(CgState));
    a.partials = (double *)take(sizeof(double) * kMaxBlocksCap * kSlots);
    a.tickets = (unsigned int *)take(sizeof(unsigned int) * 16);
    if (out) *out = a;
    return off;
}

}  // namespace

namespace ffr {
int resolve_method(const ffr_params &p)
{
    if (p.poisson_method == FFR_POISSON_CG) return FFR_POISSON_CG;
    if (p.poisson_method == FFR_POISSON_MGPCG) return FFR_POISSON_MGPCG;
    // FFR_POISSON_AUTO: the reference solver for the reference precision
    if (p.use_multigrid || p.precision == FFR_PREC_MIXED) return FFR_POISSON_MGPCG;
    return FFR_POISSON_CG;
}

StepParams make_step_params(const ffr_params &p)
{
    StepParams sp{};
    sp.dt = p.dt;
    sp.nu = p.nu;
    sp.dt_rho = p.dt / p.rho;                   // FluidSolver.cpp:333
    sp.rhs_factor = -(p.rho / p.dt);            // :315
    sp.buoy_coeff = p.g_accel * p.beta * p.dt;  // :377
    sp.T_ref = p.T_initial;                     // :378
    sp.alpha_T = p.k_thermal / (p.rho * p.cp);  // :390
    sp.scheme = p.advection;
    sp.buoyancy = p.buoyancy;
    return sp;
}

double bits_to_double(unsigned long long b)
{
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

size_t scene_bytes(const ffr_scene_desc *d, const ffr_params *p, bool with_mg)
{
    size_t b = layout(nullptr, d, p, nullptr);
    if (with_mg) {
        b += mg_bytes(d->nx, d->ny, d->nz, d->dxv, d->dyv, d->dzv, *p);
    }
    return b;
}
bool scene_uses_mg(const ffr_params &p) { return resolve_method(p) == FFR_POISSON_MGPCG; }
}  // namespace ffr

extern "C" {

size_t ffr_scene_footprint(const ffr_scene_desc *d, const ffr_params *p)
{
    char err[256];
    if (!valid_desc(d, err, sizeof err) || !valid_params(p, err, sizeof err)) return 0;
    return scene_bytes(d, p, scene_uses_mg(*p));
}

int ffr_scene_create(ffr_ctx *ctx, const ffr_scene_desc *d, const ffr_params *prm, ffr_scene **out)
{
    if (!ctx || !out) return FFR_E_INVALID;
    *out = nullptr;
    if (!valid_desc(d, ctx->err, sizeof ctx->err)) return FFR_E_INVALID;
    if (!valid_params(prm, ctx->err, sizeof ctx->err)) return FFR_E_INVALID;
    HIP_CHECK_CTX(ctx, hipSetDevice(ctx->device), FFR_E_DEVICE);

    auto *s = new (std::nothrow) ffr_scene();
    if (!s) return FFR_E_NOMEM;
    s->ctx = ctx;
    s->prm = *prm;
    s->nx = d->nx; s->ny = d->ny; s->nz = d->nz;
    s->nu = (uint64_t)(d->nx + 1) * d->ny * d->nz;
    s->nv = (uint64_t)d->nx * (d->ny + 1) * d->nz;
    s->nw = (uint64_t)d->nx * d->ny * (d->nz + 1);
    s->nc = (uint64_t)d->nx * d->ny * d->nz;
    s->fp64 = prm->precision == FFR_PREC_FP64;
    s->dxv.assign(d->dxv, d->dxv + d->nx);
    s->dyv.assign(d->dyv, d->dyv + d->ny);
    s->dzv.assign(d->dzv, d->dzv + d->nz);
    if (d->cell_type) s->flags.assign(d->cell_type, d->cell_type + s->nc);
    s->fans.assign(d->fans, d->fans + d->n_fans);
    s->outflows.assign(d->outflows, d->outflows + d->n_outflows);
    s->openings.assign(d->openings, d->openings + d->n_openings);
    // a stream per scene: the serving loop overlaps the steps of many scenes
    // (stream, event, pinned and device blocks come from the context's cache)
    s->stream = ctx->stream_take();
    s->task.ev = ctx->event_take();
    s->h_block = ctx->pinned_take();
    if (!s->stream || !s->task.ev || !s->h_block) {
        ctx->fail("creating the scene's stream, event or pinned mirror failed");
        ffr_scene_destroy(s);
        return FFR_E_DEVICE;
    }
    s->own_stream = true;
    {
        char *h = static_cast<char *>(s->h_block);
        s->h_rec = reinterpret_cast<StepRec *>(h);
        s->h_cg = reinterpret_cast<CgState *>(h + pinned_slot(sizeof(StepRec)));
        s->h_mgs = reinterpret_cast<MgState *>(h + pinned_slot(sizeof(StepRec)) + pinned_slot(sizeof(CgState)));
    }

    const size_t bytes = layout(nullptr, d, prm, nullptr);
    if (ctx->in_use + bytes > ctx->budget) {
        ctx->fail("scene needs %zu bytes, context budget has %zu free", bytes, ctx->budget - ctx->in_use);
        ffr_scene_destroy(s);
        return FFR_E_NOMEM;
    }
    s->arena = ctx->dev_take(bytes, &s->arena_bytes);
    if (!s->arena) {
        ctx->fail("hipMalloc of %zu bytes failed", bytes);
        ffr_scene_destroy(s);
        return FFR_E_NOMEM;
    }
    ArenaPtrs a;
    layout((char *)s->arena, d, prm, &a);
    if (!d->cell_type) {  // the flags are marked on the device; the host keeps a copy
        MarkList ml{};
        ml.n = d->n_marks;
        for (int q = 0; q < d->n_marks; q++) {
            const ffr_mark &m = d->marks[q];
            const int v[7] = {m.type, m.i_min, m.i_max, m.j_min, m.j_max, m.k_min, m.k_max};
            for (int e = 0; e < 7; e++) ml.m[q][e] = v[e];
        }
        s->flags.resize(s->nc);
        launch_mark_cells(s->stream, d->nx, d->ny, d->nz, ml, a.flags);
        if (hipMemcpyAsync(s->flags.data(), a.flags, s->nc, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
            hipStreamSynchronize(s->stream) != hipSuccess) {
            ctx->fail("marking the cell flags failed");
            ffr_scene_destroy(s);
            return FFR_E_DEVICE;
        }
    }

    // ---- geometry, computed on the host exactly as the reference does -----
    const int nx = d->nx, ny = d->ny, nz = d->nz, na = nx + ny + nz;
    std::vector<double> ax64(8 * (size_t)na);
    std::vector<float> ax32(10 * (size_t)na);
    // per-axis blocks: [h | i2 | c | 1/h | 1/c] then lap (sl | sr) per axis
    double *h64[3], *i2_64[3], *c64[3], *ih64[3], *ic64[3];
    float *h32[3], *ih32[3], *i2_32[3], *c32[3], *ic32[3];
    const int n[3] = {nx, ny, nz};
    const std::vector<double> *hv[3] = {&s->dxv, &s->dyv, &s->dzv};
    {
        size_t o = 0, o32 = 0;
        for (int q = 0; q < 3; q++) { h64[q] = &ax64[o]; o += n[q]; }
        for (int q = 0; q < 3; q++) { i2_64[q] = &ax64[o]; o += n[q]; }
        for (int q = 0; q < 3; q++) { c64[q] = &ax64[o]; o += n[q]; }
        for (int q = 0; q < 3; q++) { ih64[q] = &ax64[o]; o += n[q]; }
        for (int q = 0; q < 3; q++) { ic64[q] = &ax64[o]; o += n[q]; }
        for (int q = 0; q < 3; q++) { h32[q] = &ax32[o32]; o32 += n[q]; }
        for (int q = 0; q < 3; q++) { ih32[q] = &ax32[o32]; o32 += n[q]; }
        for (int q = 0; q < 3; q++) { i2_32[q] = &ax32[o32]; o32 += n[q]; }
        for (int q = 0; q < 3; q++) { c32[q] = &ax32[o32]; o32 += n[q]; }
        for (int q = 0; q < 3; q++) { ic32[q] = &ax32[o32]; o32 += n[q]; }
        // lap blocks follow at o (2*na doubles) and o32 (2*na floats)
        double *lap = &ax64[o];
        float *lapf = &ax32[o32];
        size_t lo = 0;
        for (int q = 0; q < 3; q++) {
            const std::vector<double> &hh = *hv[q];
            for (int i = 0; i < n[q]; i++) {
                const double hi = hh[i];
                h64[q][i] = hi;
                i2_64[q][i] = 1.0 / (hi * hi);                           // FluidSolver.cpp:101-103
                c64[q][i] = i > 0 ? 0.5 * (hh[i - 1] + hh[i]) : hi;      // :343, 355, 367
                ih64[q][i] = 1.0 / hi;           // reciprocals: the FP32 form multiplies
                ic64[q][i] = 1.0 / c64[q][i];
                h32[q][i] = (float)hi;
                ih32[q][i] = (float)(1.0 / hi);
                i2_32[q][i] = (float)(1.0 / (hi * hi));
                c32[q][i] = (float)c64[q][i];
                ic32[q][i] = (float)ic64[q][i];
                const double hl = (i > 0) ? 0.5 * (hh[i - 1] + hh[i]) : hh[i];         // PoissonSolver.cpp:34-49
                const double hr = (i < n[q] - 1) ? 0.5 * (hh[i] + hh[i + 1]) : hh[i];
                lap[lo + i] = 1.0 / (hl * hh[i]);
                lap[lo + n[q] + i] = 1.0 / (hr * hh[i]);
                lapf[lo + i] = (float)lap[lo + i];
                lapf[lo + n[q] + i] = (float)lap[lo + n[q] + i];
            }
            lo += 2 * n[q];
        }
    }
    std::vector<FanDev> fans(std::max(1, d->n_fans));
    for (int q = 0; q < d->n_fans; q++) {
        const ffr_fan &f = d->fans[q];
        FanDev &g = fans[q];
        // boxes clamped as add_fan / apply_inflow do (BoundaryManager.cpp:18-20, 113-115)
        g.i0 = std::max(0, f.i_min); g.i1 = std::min(nx - 1, f.i_max);
        g.j0 = std::max(0, f.j_min); g.j1 = std::min(ny - 1, f.j_max);
        g.k0 = std::max(0, f.k_min); g.k1 = std::min(nz - 1, f.k_max);
        for (int a2 = 0; a2 < 3; a2++) { g.vel[a2] = f.vel[a2]; g.velf[a2] = (float)f.vel[a2]; }
    }
    std::vector<OpeningDev> opens(std::max(1, d->n_openings));
    int open_threads = 0;
    for (int q = 0; q < d->n_openings; q++) {
        const ffr_opening &o = d->openings[q];
        OpeningDev &g = opens[q];
        g.axis = o.axis; g.side = o.side; g.a0 = o.a_min; g.a1 = o.a_max; g.b0 = o.b_min; g.b1 = o.b_max;
        g.T_out = o.T_outside;
        g.offset = open_threads;
        open_threads += std::max(0, o.a_max - o.a_min + 1) * std::max(0, o.b_max - o.b_min + 1);
    }
    const size_t nwall = 2 * ((size_t)ny * nz + (size_t)nx * nz + (size_t)nx * ny);
    std::vector<uint8_t> wall(nwall, 0);
    for (auto &o : s->outflows) {
        for (int aa = o.a_min; aa <= o.a_max; aa++)
            for (int bb = o.b_min; bb <= o.b_max; bb++) {
                size_t idx;
                if (o.axis == 0) idx = ((size_t)o.side * ny + aa) * nz + bb;
                else if (o.axis == 1) idx = 2 * (size_t)ny * nz + ((size_t)o.side * nx + aa) * nz + bb;
                else idx = 2 * (size_t)ny * nz + 2 * (size_t)nx * nz + ((size_t)o.side * nx + aa) * ny + bb;
                wall[idx] = 1;
            }
    }
    s->n_fluid = 0;
    for (uint64_t c = 0; c < s->nc; c++) s->n_fluid += (s->flags[c] == FLUID);
    s->opening_touches_fluid = false;
    for (auto &o : s->openings)
        for (int aa = o.a_min; aa <= o.a_max; aa++)
            for (int bb = o.b_min; bb <= o.b_max; bb++) {
                long c;
                if (o.axis == 0) c = ((long)(o.side ? nx - 1 : 0) * ny + aa) * nz + bb;
                else if (o.axis == 1) c = ((long)aa * ny + (o.side ? ny - 1 : 0)) * nz + bb;
                else c = ((long)aa * ny + bb) * nz + (o.side ? nz - 1 : 0);
                if (s->flags[c] == FLUID) s->opening_touches_fluid = true;
            }

    hipStream_t st = s->stream;
    bool ok = true;
    if (d->cell_type)
        ok &= hipMemcpyAsync(a.flags, s->flags.data(), s->nc, hipMemcpyHostToDevice, st) == hipSuccess;
    // neighbour codes of the pressure operator, from the flags on the device
    launch_nb_codes(st, nx, ny, nz, a.flags, a.nbm);
    ok &= hipGetLastError() == hipSuccess;
    ok &= hipMemcpyAsync(a.ax64, ax64.data(), ax64.size() * 8, hipMemcpyHostToDevice, st) == hipSuccess;
    ok &= hipMemcpyAsync(a.ax32, ax32.data(), ax32.size() * 4, hipMemcpyHostToDevice, st) == hipSuccess;
    ok &= hipMemcpyAsync(a.fans, fans.data(), fans.size() * sizeof(FanDev), hipMemcpyHostToDevice, st) == hipSuccess;
    ok &= hipMemcpyAsync(a.opens, opens.data(), opens.size() * sizeof(OpeningDev), hipMemcpyHostToDevice, st) == hipSuccess;
    ok &= hipMemcpyAsync(a.wall, wall.data(), nwall, hipMemcpyHostToDevice, st) == hipSuccess;
    // zero the state, counters and diagnostics (creation only, never captured)
    const size_t state_bytes = (size_t)((char *)a.scratch ? (char *)a.scratch : (char *)a.rec) - (size_t)a.u;
    ok &= hipMemsetAsync(a.u, 0, state_bytes, st) == hipSuccess;
    ok &= hipMemsetAsync(a.rec, 0, (size_t)((char *)(a.tickets + 16) - (char *)a.rec), st) == hipSuccess;
    ok &= hipStreamSynchronize(st) == hipSuccess;
    if (!ok) {
        ctx->fail("scene upload failed");
        ffr_scene_destroy(s);
        return FFR_E_DEVICE;
    }

    Geo &g = s->geo;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.nu = (uint32_t)s->nu; g.nv = (uint32_t)s->nv; g.nw = (uint32_t)s->nw; g.nc = (uint32_t)s->nc;
    g.f_nz = make_fastdiv(nz);
    g.f_nz1 = make_fastdiv(nz + 1);
    g.f_nynz = make_fastdiv(ny * nz);
    g.f_ny1nz = make_fastdiv((ny + 1) * nz);
    g.f_nynz1 = make_fastdiv(ny * (nz + 1));
    g.flags = a.flags;
    g.nbm = a.nbm;
    {
        auto uni = [](const std::vector<double> &h) {
            for (double x : h)
                if (x != h[0]) return false;  // exactly: the uniform kernels read h[0] for every cell
            return true;
        };
        g.uniform = uni(s->dxv) && uni(s->dyv) && uni(s->dzv);
        // 1/(h*h) is what PoissonSolver.cpp:34-49 gives every face of a uniform axis
        g.sxd = g.uniform ? 1.0 / (s->dxv[0] * s->dxv[0]) : 0.0;
        g.syd = g.uniform ? 1.0 / (s->dyv[0] * s->dyv[0]) : 0.0;
        g.szd = g.uniform ? 1.0 / (s->dzv[0] * s->dzv[0]) : 0.0;
        g.sx = (float)g.sxd;
        g.sy = (float)g.syd;
        g.sz = (float)g.szd;
    }
    {
        const double *b64 = a.ax64;
        g.hx = b64; g.hy = b64 + nx; g.hz = b64 + nx + ny; b64 += na;
        g.i2x = b64; g.i2y = b64 + nx; g.i2z = b64 + nx + ny; b64 += na;
        g.cx = b64; g.cy = b64 + nx; g.cz = b64 + nx + ny; b64 += na;
        g.ihx = b64; g.ihy = b64 + nx; g.ihz = b64 + nx + ny; b64 += na;
        g.icx = b64; g.icy = b64 + nx; g.icz = b64 + nx + ny; b64 += na;
        g.lap = b64;
        const float *b32 = a.ax32;
        g.hxf = b32; g.hyf = b32 + nx; g.hzf = b32 + nx + ny; b32 += na;
        g.ihxf = b32; g.ihyf = b32 + nx; g.ihzf = b32 + nx + ny; b32 += na;
        g.i2xf = b32; g.i2yf = b32 + nx; g.i2zf = b32 + nx + ny; b32 += na;
        g.cxf = b32; g.cyf = b32 + nx; g.czf = b32 + nx + ny; b32 += na;
        g.icxf = b32; g.icyf = b32 + nx; g.iczf = b32 + nx + ny; b32 += na;
        g.lapf = b32;
    }
    g.fans = a.fans;
    g.n_fans = d->n_fans;
    g.wall_of = a.wall;
    g.openings = a.opens;
    g.n_openings = d->n_openings;
    g.open_threads = open_threads;
    g.fcode = a.fcode;
    launch_face_codes(st, g, a.fcode);
    if (hipStreamSynchronize(st) != hipSuccess) {
        ctx->fail("face code construction failed");
        ffr_scene_destroy(s);
        return FFR_E_DEVICE;
    }

    s->u = a.u; s->v = a.v; s->w = a.w; s->us = a.us; s->vs = a.vs; s->ws = a.ws;
    s->T = a.T; s->Tn = a.Tn;
    s->p = a.p; s->b = a.b; s->r = a.r; s->d = a.d; s->Ad = a.Ad;
    s->scratch64 = a.scratch;
    s->p_hist = a.p_hist;
    s->p_hist2 = a.p_hist2;
    s->rec = a.rec; s->cg = a.cg;
    s->red.partials = a.partials;
    s->red.tickets = a.tickets;

    if (scene_uses_mg(s->prm)) {
        int rc = mg_create(s);
        if (rc != FFR_OK) {
            ffr_scene_destroy(s);
            return rc;
        }
    }
    ctx->scenes.insert(s);
    *out = s;
    return FFR_OK;
}

void ffr_scene_destroy(ffr_scene *s)
{
    if (!s) return;
    hipSetDevice(s->ctx->device);
    ffr_ctx *c = s->ctx;
    c->scenes.erase(s);
    serve_forget(c, s);  // waits for a gang stream that may still read the scene
    // the blocks go back to the context's cache only once the stream is idle;
    // a stream that reports an error is not trusted with another scene
    const bool idle = !s->stream || hipStreamSynchronize(s->stream) == hipSuccess;
    if (!idle) hipDeviceSynchronize();
    snapshot_release(s);
    lbm_release(s);
    mg_destroy(s);
    c->dev_give(s->arena, s->arena_bytes);
    c->event_give(s->task.ev);
    c->pinned_give(s->h_block);
    if (s->own_stream) {
        if (idle) c->stream_give(s->stream);
        else hipStreamDestroy(s->stream);
    }
    delete s;
}

int ffr_scene_cell_type(const ffr_scene *s, uint8_t *cell_type)
{
    if (!s || !cell_type) return FFR_E_INVALID;
    std::copy(s->flags.begin(), s->flags.end(), cell_type);
    return FFR_OK;
}

int ffr_scene_set_params(ffr_scene *s, const ffr_params *p)
{
    if (!s) return FFR_E_INVALID;
    if (!valid_params(p, s->ctx->err, sizeof s->ctx->err)) return FFR_E_INVALID;
    if (p->precision != s->prm.precision) {
        s->ctx->fail("precision is fixed at scene creation");
        return FFR_E_STATE;
    }
    if (task_busy(s)) {
        s->ctx->fail("the scene has a step in flight");
        return FFR_E_STATE;
    }
    const bool need_mg = scene_uses_mg(*p);
    // the hierarchy (levels, storage, sweep counts, coarsest inverse, tail
    // layout and captured graphs) is built from these fields at creation
    const ffr_params &o = s->prm;
    const bool mg_changed = o.mg_levels != p->mg_levels || o.mg_bf16 != p->mg_bf16 ||
                            o.mg_pre_sweeps != p->mg_pre_sweeps ||
                            o.mg_level_sweeps != p->mg_level_sweeps ||
                            o.mg_coarse_sweeps != p->mg_coarse_sweeps ||
                            o.mg_tail_kb != p->mg_tail_kb || o.mg_tail_cells != p->mg_tail_cells ||
                            o.mg_tail_groups != p->mg_tail_groups;
    s->prm = *p;
    s->pred = SolvePred{};  // iteration counts of other settings do not carry over
    s->p_hist_n = 0;
    if (s->mg.levels && mg_changed) mg_destroy(s);
    if (need_mg && !s->mg.levels) return mg_create(s);
    return FFR_OK;
}

int ffr_scene_step_count(const ffr_scene *s) { return s ? s->step : -1; }

int ffr_scene_set_step_count(ffr_scene *s, int step)
{
    if (!s || step < 0) return FFR_E_INVALID;
    s->step = step;
    return FFR_OK;
}

int ffr_scene_upload(ffr_scene *s, const double *u, const double *v, const double *w,
                     const double *p, const double *T)
{
    if (!s) return FFR_E_INVALID;
    hipSetDevice(s->ctx->device);
    hipStream_t st = s->stream;
    const double *src[5] = {u, v, w, p, T};
    void *dst[5] = {s->u, s->v, s->w, s->p, s->T};
    // an uploaded p takes the place of the last step's pressure; the older
    // ones stay (a host mirror pushed back unchanged then steps exactly as
    // if it had not been)
    const uint64_t cnt[5] = {s->nu, s->nv, s->nw, s->nc, s->nc};
    for (int q = 0; q < 5; q++) {
        if (!src[q]) continue;
        if (s->fp64 || q == 3) {  // p is FP64 in both precisions
            HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(dst[q], src[q], cnt[q] * 8, hipMemcpyHostToDevice, st), FFR_E_DEVICE);
        } else {
            HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(s->scratch64, src[q], cnt[q] * 8, hipMemcpyHostToDevice, st), FFR_E_DEVICE);
            launch_convert<float>(st, s->scratch64, (float *)dst[q], cnt[q]);
        }
    }
    HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(st), FFR_E_DEVICE);
    return FFR_OK;
}

int ffr_scene_download(ffr_scene *s, double *u, double *v, double *w, double *p, double *T)
{
    if (!s) return FFR_E_INVALID;
    hipSetDevice(s->ctx->device);
    hipStream_t st = s->stream;
    double *dst[5] = {u, v, w, p, T};
    void *src[5] = {s->u, s->v, s->w, s->p, s->T};
    const uint64_t cnt[5] = {s->nu, s->nv, s->nw, s->nc, s->nc};
    for (int q = 0; q < 5; q++) {
        if (!dst[q]) continue;
        if (s->fp64 || q == 3) {
            HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(dst[q], src[q], cnt[q] * 8, hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
        } else {
            launch_convert_back<float>(st, (const float *)src[q], s->scratch64, cnt[q]);
            HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(dst[q], s->scratch64, cnt[q] * 8, hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
            // the scratch buffer is reused by the next field: wait per field
            HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(st), FFR_E_DEVICE);
        }
    }
    HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(st), FFR_E_DEVICE);
    return FFR_OK;
}

}  // extern "C"

extern "C" {

int ffr_step(ffr_scene *s, ffr_step_result *out)
{
    if (!s) return FFR_E_INVALID;
    ffr_step_result tmp;
    return scene_step(s, out ? out : &tmp);
}

int ffr_run(ffr_scene *s, int max_steps, ffr_step_callback cb, void *user, ffr_step_result *last)
{
    // FluidSolver::run (FluidSolver.cpp:552-560)
    if (!s) return FFR_E_INVALID;
    ffr_step_result r{};
    for (int k = 0; k < max_steps; k++) {
        int rc = scene_step(s, &r);
        if (rc != FFR_OK) return rc;
        if (cb) cb(&r, user);
        if (r.converged) break;
    }
    if (last) *last = r;
    return FFR_OK;
}

int ffr_divergence_max(ffr_scene *s, double *out)
{
    if (!s || !out) return FFR_E_INVALID;
    hipSetDevice(s->ctx->device);
    hipStream_t st = s->stream;
    HIP_CHECK_CTX(s->ctx, hipMemsetAsync(&s->rec->divmax_bits, 0, sizeof(unsigned long long), st), FFR_E_DEVICE);
    if (s->fp64) launch_divmax<double>(st, s->geo, (double *)s->u, (double *)s->v, (double *)s->w, s->rec);
    else launch_divmax<float>(st, s->geo, (float *)s->u, (float *)s->v, (float *)s->w, s->rec);
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(s->h_rec, s->rec, sizeof(StepRec), hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(st), FFR_E_DEVICE);
    *out = bits_to_double(s->h_rec->divmax_bits);
    return FFR_OK;
}

int ffr_T_mean(ffr_scene *s, double *out)
{
    if (!s || !out) return FFR_E_INVALID;
    hipSetDevice(s->ctx->device);
    hipStream_t st = s->stream;
    const unsigned G = s->fp64 ? launch_tsum<double>(st, s->geo, (double *)s->T, s->rec, s->red)
                               : launch_tsum<float>(st, s->geo, (float *)s->T, s->rec, s->red);
    launch_step_finalize(st, s->rec, s->red, G);
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(s->h_rec, s->rec, sizeof(StepRec), hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(st), FFR_E_DEVICE);
    *out = s->n_fluid > 0 ? s->h_rec->tsum / (double)s->n_fluid : s->prm.T_initial;  // :458-468
    return FFR_OK;
}

int ffr_apply_laplacian(ffr_scene *s, const double *x, double *Ax)
{
    if (!s || !x || !Ax) return FFR_E_INVALID;
    hipSetDevice(s->ctx->device);
    hipStream_t st = s->stream;
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(s->d, x, s->nc * 8, hipMemcpyHostToDevice, st), FFR_E_DEVICE);
    launch_laplacian(st, s->geo, s->d, s->Ad);
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(Ax, s->Ad, s->nc * 8, hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(st), FFR_E_DEVICE);
    return FFR_OK;
}

// The scene's u, v, w copied into its u*, v*, w* buffers: the single-phase
// entry points below take the current velocity as the predictor's output.
static int state_to_star(ffr_scene *s)
{
    const size_t sz = s->fp64 ? 8 : 4;
    hipStream_t st = s->stream;
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(s->us, s->u, s->nu * sz, hipMemcpyDeviceToDevice, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(s->vs, s->v, s->nv * sz, hipMemcpyDeviceToDevice, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(s->ws, s->w, s->nw * sz, hipMemcpyDeviceToDevice, st), FFR_E_DEVICE);
    return FFR_OK;
}

int ffr_build_rhs(ffr_scene *s, double *rhs)
{
    if (!s || !rhs) return FFR_E_INVALID;
    hipSetDevice(s->ctx->device);
    if (task_busy(s)) {
        s->ctx->fail("ffr_build_rhs: the scene has a step in flight");
        return FFR_E_STATE;
    }
    hipStream_t st = s->stream;
    const StepParams sp = make_step_params(s->prm);
    int rc = state_to_star(s);
    if (rc != FFR_OK) return rc;
    if (s->fp64) launch_rhs<double>(st, s->geo, sp, (double *)s->us, (double *)s->vs, (double *)s->ws, s->b);
    else launch_rhs<float>(st, s->geo, sp, (float *)s->us, (float *)s->vs, (float *)s->ws, s->b);
    HIP_CHECK_CTX(s->ctx, hipGetLastError(), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(rhs, s->b, s->nc * 8, hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(st), FFR_E_DEVICE);
    return FFR_OK;
}

int ffr_apply_correction(ffr_scene *s, const double *p, double *velocity_change)
{
    if (!s || !p) return FFR_E_INVALID;
    hipSetDevice(s->ctx->device);
    if (task_busy(s)) {
        s->ctx->fail("ffr_apply_correction: the scene has a step in flight");
        return FFR_E_STATE;
    }
    hipStream_t st = s->stream;
    const StepParams sp = make_step_params(s->prm);
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(s->p, p, s->nc * 8, hipMemcpyHostToDevice, st), FFR_E_DEVICE);
    s->p_hist_n = 0;
    HIP_CHECK_CTX(s->ctx, hipMemsetAsync(&s->rec->vmax_bits, 0, sizeof(unsigned long long), st), FFR_E_DEVICE);
    int rc = state_to_star(s);
    if (rc != FFR_OK) return rc;
    if (s->fp64)
        launch_correct<double>(st, s->geo, sp, (double *)s->us, (double *)s->vs, (double *)s->ws, s->p,
                               (double *)s->u, (double *)s->v, (double *)s->w, s->rec);
    else
        launch_correct<float>(st, s->geo, sp, (float *)s->us, (float *)s->vs, (float *)s->ws, s->p,
                              (float *)s->u, (float *)s->v, (float *)s->w, s->rec);
    HIP_CHECK_CTX(s->ctx, hipGetLastError(), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(s->h_rec, s->rec, sizeof(StepRec), hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(st), FFR_E_DEVICE);
    if (velocity_change) *velocity_change = bits_to_double(s->h_rec->vmax_bits);
    return FFR_OK;
}

int ffr_prof_kernel(ffr_scene *s, const char *name, int reps, double *mean_ms, double *bytes)
{
    if (!s || !name || reps <= 0) return FFR_E_INVALID;
    hipSetDevice(s->ctx->device);
    hipStream_t st = s->stream;
    const Geo &g = s->geo;
    const double nc = (double)s->nc, faces = (double)(s->nu + s->nv + s->nw);
    const double sz = s->fp64 ? 8.0 : 4.0;
    const StepParams sp = make_step_params(s->prm);
    std::string k(name);
    std::function<void()> launch;
    double b = 0.0;
    const bool mg = s->mg.levels > 0;
    if (k == "predict") {
        // u, v, w and their face codes read, u*, v*, w* written, T read with buoyancy
        b = faces * (2 * sz + 2) + (s->prm.buoyancy ? nc * sz : 0);
        if (s->fp64)
            launch = [&] { launch_predict<double>(st, g, sp, (double *)s->u, (double *)s->v, (double *)s->w, (double *)s->T, (double *)s->us, (double *)s->vs, (double *)s->ws, s->rec); };
        else
            launch = [&] { launch_predict<float>(st, g, sp, (float *)s->u, (float *)s->v, (float *)s->w, (float *)s->T, (float *)s->us, (float *)s->vs, (float *)s->ws, s->rec); };
    } else if (k == "correct") {  // u = u* - dt/rho grad p from the last step's u* and p
        // u*, v*, w*, u^n and the face codes read, u written, p read
        b = faces * (3 * sz + 2) + nc * 8;
        if (s->fp64)
            launch = [&] { launch_correct<double>(st, g, sp, (double *)s->us, (double *)s->vs, (double *)s->ws, s->p, (double *)s->u, (double *)s->v, (double *)s->w, s->rec); };
        else
            launch = [&] { launch_correct<float>(st, g, sp, (float *)s->us, (float *)s->vs, (float *)s->ws, s->p, (float *)s->u, (float *)s->v, (float *)s->w, s->rec); };
    } else if (s->lbm.on && (k == "lbm_relax" || k == "lbm_stream" || k == "lbm_faces")) {
        // the lattice's kernels on its current state (algorithmic bytes:
        // populations read and written, flags, cell velocities and faces)
        LbmState &L = s->lbm;
        if (k == "lbm_relax") {  // on the spare buffer: the state stays as it is
            b = nc * (19 * 16 + 1);
            launch = [&] { lbm_launch_relax(st, g, L.f2, L.dt / L.dx_ref, 1.0 / L.tau, true, s->rec); };
        } else if (k == "lbm_stream" && L.fused) {  // the step's streaming pass (writes the spare buffer)
            b = nc * (19 * 16 + 4 + 32);
            launch = [&] {
                lbm_launch_stream_collide(st, g, L.mask, L.f, L.f2, L.cu, s->p, L.dx_ref / L.dt,
                                          L.dt / L.dx_ref, 1.0 / L.tau, s->rec);
            };
        } else if (k == "lbm_stream") {  // populations, mask, cell velocities and p
            b = nc * (19 * 16 + 4 + 32);
            launch = [&] { lbm_launch_stream(st, g, L.mask, L.f, L.f2, L.cu, s->p, L.dx_ref / L.dt); };
        } else {
            b = faces * 16 + nc * 24;
            launch = [&] { lbm_launch_faces(st, g, L.cu, (double *)s->u, (double *)s->v, (double *)s->w, s->rec); };
        }
    } else if (k == "temperature" && s->prm.thermal) {  // T -> Tn (the state keeps T)
        b = nc * 2 * sz + faces * sz + nc;
        if (s->fp64)
            launch = [&] { launch_temperature<double>(st, g, sp, (double *)s->u, (double *)s->v, (double *)s->w, (double *)s->T, (double *)s->Tn, s->rec, s->red); };
        else
            launch = [&] { launch_temperature<float>(st, g, sp, (float *)s->u, (float *)s->v, (float *)s->w, (float *)s->T, (float *)s->Tn, s->rec, s->red); };
    } else if (k == "rhs") {
        b = faces * sz + nc + nc * 8;
        if (s->fp64) launch = [&] { launch_rhs<double>(st, g, sp, (double *)s->us, (double *)s->vs, (double *)s->ws, s->b); };
        else launch = [&] { launch_rhs<float>(st, g, sp, (float *)s->us, (float *)s->vs, (float *)s->ws, s->b); };
    } else if (k == "cg_spmv" || k == "cg_update" || k == "cg_dir") {
        // keep the CG state live for the timed launches
        HIP_CHECK_CTX(s->ctx, hipMemsetAsync(&s->cg->done, 0, sizeof(int), st), FFR_E_DEVICE);
        if (k == "cg_spmv") { b = nc * 17; launch = [&] { launch_cg_spmv(st, g, s->d, s->Ad, s->cg, s->red); }; }
        else if (k == "cg_update") { b = nc * 48; launch = [&] { launch_cg_update(st, g.nc, s->p, s->r, s->d, s->Ad, s->cg, s->red); }; }
        else { b = nc * 24; launch = [&] { launch_cg_dir(st, g.nc, s->d, s->r, s->cg, s->red); }; }
    } else if (k == "copy") {  // plain FP64 copy b -> r (counter calibration)
        b = nc * 16;
        launch = [&] { launch_copy64(st, s->b, s->r, s->nc); };
    } else if (mg && (k == "mg_smooth_l0" || k == "mg_resid0" || k == "pcg_dirq" ||
                      k == "pcg_update" || k == "dc_residual")) {
        HIP_CHECK_CTX(s->ctx, hipMemsetAsync(&s->mgs->solve_done, 0, 2 * sizeof(int), st), FFR_E_DEVICE);
        b = mg_prof_launch(s, k.c_str(), &launch);
    } else {
        s->ctx->fail("ffr_prof_kernel: unknown kernel '%s' for this scene", name);
        return FFR_E_INVALID;
    }
    // each launch timed by its own dispatch begin/end events
    std::vector<hipEvent_t> ev(2 * (size_t)reps);
    for (auto &e : ev) HIP_CHECK_CTX(s->ctx, hipEventCreate(&e), FFR_E_DEVICE);
    launch();  // warm
    for (int r = 0; r < reps; r++) {
        g_prof_events.start = ev[2 * r];
        g_prof_events.stop = ev[2 * r + 1];
        launch();
        g_prof_events = ProfEvents{};
    }
    HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(st), FFR_E_DEVICE);
    double total = 0.0;
    for (int r = 0; r < reps; r++) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]);
        total += ms;
    }
    for (auto &e : ev) hipEventDestroy(e);
    if (mean_ms) *mean_ms = total / reps;
    if (bytes) *bytes = b;
    return FFR_OK;
}

int ffr_poisson_solve(ffr_scene *s, const double *rhs, double *p, ffr_solve_info *info)
{
    if (!s || !rhs) return FFR_E_INVALID;
    hipSetDevice(s->ctx->device);
    hipStream_t st = s->stream;
    if (task_busy(s)) {
        s->ctx->fail("ffr_poisson_solve: the scene has a step in flight");
        return FFR_E_STATE;
    }
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(s->b, rhs, s->nc * 8, hipMemcpyHostToDevice, st), FFR_E_DEVICE);
    int rc = task_begin(s, true);
    bool done = false;
    while (rc == FFR_OK && !done) rc = task_advance(s, true, &done);
    if (rc != FFR_OK) return rc;
    if (p) HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(p, s->p, s->nc * 8, hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(st), FFR_E_DEVICE);
    if (info) {
        info->residual = s->task.res;
        info->iterations = s->task.iters;
        info->outer_iterations = s->task.outer;
    }
    return FFR_OK;
}

}  // extern "C"
