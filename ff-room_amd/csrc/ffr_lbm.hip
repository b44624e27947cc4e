// ffr_lbm.hip - the LBMSolver of the reference (LBMSolver.hpp:17-39) on a
// scene: ffr_lbm_attach builds the lattice (rest equilibrium, then the
// fans' inflow, LBMSolver.cpp:39-75), ffr_lbm_step advances it one step and
// writes u, v, w, p as the reference writes its Grid (:403-444).
#include "ffr_lbm.hpp"
#include "ffr_scene.hpp"

#include <cmath>

using namespace ffr;

namespace ffr {
namespace {
// the reference copies every outflow patch (openings included), then every
// opening again (LBMSolver.cpp:196-282); f(patch) for each, in that order
template <typename F>
void each_patch(const ffr_scene *s, F f)
{
    auto one = [&](int axis, int side, int a0, int a1, int b0, int b1) {
        const int n_ax = axis == 0 ? s->nx : (axis == 1 ? s->ny : s->nz);
        const int src = side ? n_ax - 2 : 1;
        if (src < 0 || src >= n_ax || a1 < a0 || b1 < b0) return;
        f(LbmPatch{axis, side, a0, a1, b0, b1});
    };
    for (const auto &o : s->outflows) one(o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max);
    for (const auto &o : s->openings) one(o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max);
}

// true when some patch copies into an OUTFLOW cell from an INFLOW cell (the
// fused form's patch copy would then copy an equilibrium, not a collision)
bool patch_reads_inflow(const ffr_scene *s)
{
    bool hit = false;
    each_patch(s, [&](const LbmPatch &pt) {
        const int n_ax = pt.axis == 0 ? s->nx : (pt.axis == 1 ? s->ny : s->nz);
        const int wall = pt.side ? n_ax - 1 : 0, src = pt.side ? n_ax - 2 : 1;
        for (int a = pt.a0; a <= pt.a1; a++)
            for (int b = pt.b0; b <= pt.b1; b++) {
                int cw[3];
                if (pt.axis == 0) { cw[0] = wall; cw[1] = a; cw[2] = b; }
                else if (pt.axis == 1) { cw[0] = a; cw[1] = wall; cw[2] = b; }
                else { cw[0] = a; cw[1] = b; cw[2] = wall; }
                if (cw[0] < 0 || cw[0] >= s->nx || cw[1] < 0 || cw[1] >= s->ny || cw[2] < 0 ||
                    cw[2] >= s->nz)
                    continue;
                int cs[3] = {cw[0], cw[1], cw[2]};
                cs[pt.axis] = src;
                const size_t co = ((size_t)cw[0] * s->ny + cw[1]) * s->nz + cw[2];
                const size_t ci = ((size_t)cs[0] * s->ny + cs[1]) * s->nz + cs[2];
                if (s->flags[co] == OUTFLOW && s->flags[ci] == INFLOW) hit = true;
            }
    });
    return hit;
}
}  // namespace

void lbm_release(ffr_scene *s)
{
    s->ctx->dev_give(s->lbm.arena, s->lbm.bytes);
    s->lbm = LbmState{};
}
}  // namespace ffr

extern "C" {

int ffr_lbm_attach(ffr_scene *s, double tau, double dt, double convergence_tol, double dx_ref)
{
    if (!s) return FFR_E_INVALID;
    ffr_ctx *c = s->ctx;
    if (!s->fp64) {
        c->fail("ffr_lbm_attach: the lattice runs on FP64 scenes (precision FFR_PREC_FP64)");
        return FFR_E_UNSUPPORTED;
    }
    if (!(tau > 0.0) || !(dt > 0.0) || !(dx_ref > 0.0) || !std::isfinite(tau) || !std::isfinite(dt)) {
        c->fail("ffr_lbm_attach: tau, dt and dx_ref must be positive");
        return FFR_E_INVALID;
    }
    if (task_busy(s)) {
        c->fail("ffr_lbm_attach: the scene has a step in flight");
        return FFR_E_STATE;
    }
    HIP_CHECK_CTX(c, hipSetDevice(c->device), FFR_E_DEVICE);
    HIP_CHECK_CTX(c, hipStreamSynchronize(s->stream), FFR_E_DEVICE);
    lbm_release(s);
    const size_t nc = s->nc, bytes = (2 * 19 * nc + 3 * nc) * sizeof(double) + nc * sizeof(uint32_t);
    if (c->in_use + bytes > c->budget) {
        c->fail("ffr_lbm_attach: the lattice needs %zu bytes, the context budget has %zu free", bytes,
                c->budget - c->in_use);
        return FFR_E_NOMEM;
    }
    LbmState &L = s->lbm;
    L.arena = c->dev_take(bytes, &L.bytes);
    if (!L.arena) {
        c->fail("ffr_lbm_attach: hipMalloc of %zu bytes failed", bytes);
        return FFR_E_NOMEM;
    }
    L.f = (double *)L.arena;
    L.f2 = L.f + 19 * nc;
    L.cu = L.f2 + 19 * nc;
    L.mask = (uint32_t *)(L.cu + 3 * nc);
    L.tau = tau;
    L.dt = dt;
    L.tol = convergence_tol;
    L.dx_ref = dx_ref;
    L.steps = 0;
    L.on = true;
    L.fused = !patch_reads_inflow(s);
    lbm_launch_mask(s->stream, s->geo, L.mask);
    lbm_launch_rest(s->stream, (uint32_t)nc, L.f);
    lbm_launch_relax(s->stream, s->geo, L.f, dt / dx_ref, 1.0 / tau, false, s->rec);
    // fused form: the first step's inflow and collision now (each later
    // step's come at the end of the step before)
    if (L.fused) lbm_launch_relax(s->stream, s->geo, L.f, dt / dx_ref, 1.0 / tau, true, s->rec);
    HIP_CHECK_CTX(c, hipStreamSynchronize(s->stream), FFR_E_DEVICE);
    return FFR_OK;
}

int ffr_lbm_step(ffr_scene *s, ffr_step_result *out)
{
    if (!s) return FFR_E_INVALID;
    ffr_ctx *c = s->ctx;
    LbmState &L = s->lbm;
    if (!L.on) {
        c->fail("ffr_lbm_step: no lattice on this scene (ffr_lbm_attach)");
        return FFR_E_STATE;
    }
    if (task_busy(s)) {
        c->fail("ffr_lbm_step: the scene has a step in flight");
        return FFR_E_STATE;
    }
    HIP_CHECK_CTX(c, hipSetDevice(c->device), FFR_E_DEVICE);
    hipStream_t st = s->stream;
    const Geo &g = s->geo;
    const double n = (double)g.nc;
    const double l2s = L.dx_ref / L.dt, s2l = L.dt / L.dx_ref, inv_tau = 1.0 / L.tau;
    if (L.fused) {
        {
            ProfScope ps(c, st, "lbm_stream", n * (19 * 16 + 4 + 32));
            lbm_launch_stream_collide(st, g, L.mask, L.f, L.f2, L.cu, s->p, l2s, s2l, inv_tau, s->rec);
        }
        std::swap(L.f, L.f2);
        each_patch(s, [&](const LbmPatch &pt) { lbm_launch_patch_post(st, g, L.f, pt, L.cu, s->p); });
    } else {
        {
            ProfScope ps(c, st, "lbm_relax", n * 19 * 16 + n);
            lbm_launch_relax(st, g, L.f, s2l, inv_tau, true, s->rec);
        }
        {
            ProfScope ps(c, st, "lbm_stream", n * (19 * 16 + 4 + 32));
            lbm_launch_stream(st, g, L.mask, L.f, L.f2, L.cu, s->p, l2s);
        }
        std::swap(L.f, L.f2);
        each_patch(s, [&](const LbmPatch &pt) { lbm_launch_patch(st, g, L.f, pt, L.cu, s->p, l2s); });
    }
    {
        const double faces = (double)(s->nu + s->nv + s->nw);
        ProfScope ps(c, st, "lbm_faces", faces * 16 + n * 24);
        lbm_launch_faces(st, g, L.cu, (double *)s->u, (double *)s->v, (double *)s->w, s->rec);
    }
    launch_divmax<double>(st, g, (const double *)s->u, (const double *)s->v, (const double *)s->w,
                          s->rec);
    HIP_CHECK_CTX(c, hipMemcpyAsync(s->h_rec, s->rec, sizeof(StepRec), hipMemcpyDeviceToHost, st),
                  FFR_E_DEVICE);
    HIP_CHECK_CTX(c, hipStreamSynchronize(st), FFR_E_DEVICE);
    L.steps++;
    const double vel_change = bits_to_double(s->h_rec->vmax_bits) / L.dt;
    if (out) {
        *out = ffr_step_result{};
        out->step = L.steps;
        out->velocity_change = vel_change;
        out->pressure_residual = 0.0;
        out->divergence_max = bits_to_double(s->h_rec->divmax_bits);
        out->converged = vel_change < L.tol ? 1 : 0;
        out->T_mean = 0.0;
    }
    return FFR_OK;
}

}  // extern "C"
