// ffr_task.hip - the projection step (FluidSolver::step, FluidSolver.cpp:486-550)
// and the pressure solve as resumable tasks.
//
// A task enqueues the step's kernels on the scene's stream as far as it can
// without looking at device results, copies the solver's device scalars to
// pinned host memory, records an event and returns. task_advance() resumes it
// once the device has passed that event: the pressure solve either finished
// (the rest of the step is enqueued) or needs more iterations (enqueued, next
// event). How far a solve is enqueued ahead comes from the iteration counts
// of the scene's previous solve; kernels enqueued past the point where the
// device-side stopping test ended the solve return at once. A blocking
// caller (ffr_step) waits on each event; the serving loop polls them, so the
// steps of many scenes overlap on the device.
#include "ffr_gang.hpp"
#include "ffr_scene.hpp"

#include <algorithm>

namespace ffr {

// Copies `bytes` of device state to pinned host memory and records the
// task's event after it.
int wait_point(ffr_scene *s, void *host, const void *dev, size_t bytes)
{
    if (t_rec) {  // a gang member: its round's gather copies, its event ends the round
        t_rec->push_copy(host, dev, (uint32_t)bytes);
        return FFR_OK;
    }
    hipStream_t st = s->stream;
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipEventRecord(s->task.ev, st), FFR_E_DEVICE);
    return FFR_OK;
}

// ---- reference CG (PoissonSolver.cpp:93-158), b -> p -------------------------
static int cg_enqueue(ffr_scene *s, int chunk)
{
    StepTask &t = s->task;
    hipStream_t st = s->stream;
    const Geo &g = s->geo;
    const double nc = (double)s->nc;
    const int nrun = std::max(0, std::min(chunk, s->prm.poisson_max_iter - t.cg_enq));
    for (int m = 0; m < nrun; m++) {
        { ProfScope ps(s->ctx, st, "cg_spmv", nc * 17); launch_cg_spmv(st, g, s->d, s->Ad, s->cg, s->red); }
        { ProfScope ps(s->ctx, st, "cg_update", nc * 48); launch_cg_update(st, g.nc, s->p, s->r, s->d, s->Ad, s->cg, s->red); }
        { ProfScope ps(s->ctx, st, "cg_dir", nc * 24); launch_cg_dir(st, g.nc, s->d, s->r, s->cg, s->red); }
    }
    t.cg_enq += nrun;
    return wait_point(s, s->h_cg, s->cg, sizeof(CgState));
}

static int cg_begin(ffr_scene *s)
{
    const double nc = (double)s->nc;
    {
        ProfScope ps(s->ctx, s->stream, "cg_init", nc * (8 + 1 + 8 * 4));
        launch_cg_init(s->stream, s->geo, s->b, s->p, s->r, s->d, s->cg, s->red,
                       s->prm.poisson_max_iter, s->prm.poisson_tol);
    }
    s->task.cg_enq = 0;
    // the previous solve's count plus a margin: usually one host wait per solve
    return cg_enqueue(s, s->pred.cg_iters > 0 ? s->pred.cg_iters + 2 : 16);
}

static int cg_continue(ffr_scene *s, bool *finished)
{
    StepTask &t = s->task;
    const CgState &h = *s->h_cg;
    if (!h.done && t.cg_enq < s->prm.poisson_max_iter) {
        *finished = false;
        return cg_enqueue(s, std::max(8, t.cg_enq / 4));
    }
    t.iters = h.iters;
    t.res = h.res;
    t.outer = 0;
    s->pred.cg_iters = h.iters;
    *finished = true;
    return FFR_OK;
}

// ---- the solve ---------------------------------------------------------------
static bool use_cg(const ffr_scene *s) { return resolve_method(s->prm) == FFR_POISSON_CG; }

static int solve_begin(ffr_scene *s)
{
    if (use_cg(s)) {
        if (s->prm.use_multigrid) {
            s->ctx->fail("CG with use_multigrid (pinned outflow pressure) is not provided; use MG-PCG");
            return FFR_E_UNSUPPORTED;
        }
        return cg_begin(s);
    }
    return mg_begin(s);
}

static int solve_continue(ffr_scene *s, bool *finished)
{
    return use_cg(s) ? cg_continue(s, finished) : mg_continue(s, finished);
}

// ---- the step around the solve -----------------------------------------------
template <typename R>
static void enqueue_predict(ffr_scene *s)
{
    ffr_ctx *ctx = s->ctx;
    hipStream_t st = s->stream;
    const ffr_params &P = s->prm;
    const StepParams sp = make_step_params(P);
    const double faces = (double)(s->nu + s->nv + s->nw), nc = (double)s->nc, sz = sizeof(R);
    R *us = (R *)s->us, *vs = (R *)s->vs, *ws = (R *)s->ws;
    {   // K1 predictor + buoyancy + pre-projection masks
        ProfScope ps(ctx, st, "predict", faces * (2 * sz + 2) + (P.buoyancy ? nc * sz : 0));
        launch_predict<R>(st, s->geo, sp, (R *)s->u, (R *)s->v, (R *)s->w, (R *)s->T, us, vs, ws,
                          s->rec);
    }
    {   // K4 right-hand side
        ProfScope ps(ctx, st, "rhs", faces * sz + nc + nc * 8);
        launch_rhs<R>(st, s->geo, sp, us, vs, ws, s->b);
    }
}

template <typename R>
static int enqueue_finish(ffr_scene *s)
{
    ffr_ctx *ctx = s->ctx;
    hipStream_t st = s->stream;
    const Geo &g = s->geo;
    const ffr_params &P = s->prm;
    const StepParams sp = make_step_params(P);
    R *u = (R *)s->u, *v = (R *)s->v, *w = (R *)s->w;
    const double faces = (double)(s->nu + s->nv + s->nw), nc = (double)s->nc, sz = sizeof(R);
    {   // K6 projection + masks + outflow + velocity change
        ProfScope ps(ctx, st, "correct", faces * (3 * sz + 2) + nc * 8);
        launch_correct<R>(st, g, sp, (R *)s->us, (R *)s->vs, (R *)s->ws, s->p, u, v, w, s->rec);
    }
    if (P.thermal) {
        R *T = (R *)s->T, *Tn = (R *)s->Tn;
        if (g.n_openings) {
            ProfScope ps(ctx, st, "opening_T", 0);
            launch_opening_T<R>(st, g, u, v, w, T);
        }
        unsigned G = 0;  // blocks that wrote the T-sum partials
        {
            ProfScope ps(ctx, st, "temperature", nc * 2 * sz + faces * sz + nc);
            G = launch_temperature<R>(st, g, sp, u, v, w, T, Tn, s->rec, s->red);
        }
        if (g.n_openings) {
            ProfScope ps(ctx, st, "opening_T", 0);
            launch_opening_T<R>(st, g, u, v, w, Tn);
        }
        std::swap(s->T, s->Tn);
        if (s->opening_touches_fluid) {
            ProfScope ps(ctx, st, "tsum", nc * sz + nc);
            G = launch_tsum<R>(st, g, (R *)s->T, s->rec, s->red);
        }
        launch_step_finalize(st, s->rec, s->red, G);
    } else {
        ProfScope ps(ctx, st, "divmax", faces * sz + nc);
        launch_divmax<R>(st, g, u, v, w, s->rec);
    }
    return wait_point(s, s->h_rec, s->rec, sizeof(StepRec));
}

// StepResult from the device record (FluidSolver.cpp:527-547)
static void build_result(ffr_scene *s)
{
    const StepRec &r = *s->h_rec;
    const ffr_params &P = s->prm;
    StepTask &t = s->task;
    double tmean = P.T_initial;
    if (P.thermal) tmean = s->n_fluid > 0 ? r.tsum / (double)s->n_fluid : P.T_initial;
    const double vel_change = bits_to_double(r.vmax_bits) / P.dt;
    bool conv;
    if (P.thermal && P.T_target < 1e29)
        conv = (P.T_target >= P.T_initial) ? (tmean >= P.T_target) : (tmean <= P.T_target);
    else
        conv = vel_change < P.convergence_tol;
    ffr_step_result &o = t.result;
    o.step = s->step;
    o.velocity_change = vel_change;
    o.pressure_residual = t.res;
    o.divergence_max = bits_to_double(r.divmax_bits);
    o.converged = conv ? 1 : 0;
    o.T_mean = tmean;
    o.poisson_iters = t.iters;
    o.outer_iters = t.outer;
    o.nonfinite = r.nonfinite;
}

bool task_busy(const ffr_scene *s) { return s->task.phase != StepTask::IDLE; }

// After a failed enqueue: waits for the scene's stream so that the scene
// stays usable; a stream that cannot be drained is reported as a device error.
static int idle_after_error(ffr_scene *s, int rc)
{
    const hipError_t e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        s->ctx->fail("hipStreamSynchronize after status %d failed: %s", rc, hipGetErrorString(e));
        return FFR_E_DEVICE;
    }
    return rc;
}

int task_begin(ffr_scene *s, bool solve_only)
{
    if (task_busy(s)) {
        s->ctx->fail("the scene has a step in flight");
        return FFR_E_STATE;
    }
    HIP_CHECK_CTX(s->ctx, hipSetDevice(s->ctx->device), FFR_E_DEVICE);
    StepTask &t = s->task;
    if (!t.ev) HIP_CHECK_CTX(s->ctx, hipEventCreateWithFlags(&t.ev, hipEventDisableTiming), FFR_E_DEVICE);
    t.solve_only = solve_only;
    t.iters = t.outer = 0;
    t.res = 0.0;
    if (!solve_only) {
        if (s->fp64) enqueue_predict<double>(s);
        else enqueue_predict<float>(s);
    }
    const int rc = solve_begin(s);
    if (rc != FFR_OK) return idle_after_error(s, rc);
    t.phase = StepTask::SOLVE;
    return FFR_OK;
}

int task_resume(ffr_scene *s, bool *done)
{
    StepTask &t = s->task;
    *done = false;
    if (t.phase == StepTask::SOLVE) {
        bool finished = false;
        int rc = solve_continue(s, &finished);
        if (rc == FFR_OK && finished) {
            if (t.solve_only) rc = wait_point(s, s->h_rec, s->rec, sizeof(StepRec));
            else rc = s->fp64 ? enqueue_finish<double>(s) : enqueue_finish<float>(s);
            t.phase = StepTask::RESULT;
        }
        if (rc != FFR_OK) {
            t.phase = StepTask::IDLE;
            if (!t_rec) rc = idle_after_error(s, rc);
        }
        return rc;
    }
    // RESULT: the step's record is on the host
    if (!t.solve_only) {
        s->step++;
        build_result(s);
    }
    t.phase = StepTask::IDLE;
    *done = true;
    return FFR_OK;
}

int task_advance(ffr_scene *s, bool block, bool *done)
{
    StepTask &t = s->task;
    *done = false;
    if (t.phase == StepTask::IDLE) {
        s->ctx->fail("no step in flight");
        return FFR_E_STATE;
    }
    HIP_CHECK_CTX(s->ctx, hipSetDevice(s->ctx->device), FFR_E_DEVICE);
    for (;;) {
        if (block) {
            HIP_CHECK_CTX(s->ctx, hipEventSynchronize(t.ev), FFR_E_DEVICE);
        } else {
            const hipError_t q = hipEventQuery(t.ev);
            if (q == hipErrorNotReady) return FFR_OK;
            if (q != hipSuccess) {
                s->ctx->fail("hipEventQuery failed: %s", hipGetErrorString(q));
                t.phase = StepTask::IDLE;
                return FFR_E_DEVICE;
            }
        }
        const int rc = task_resume(s, done);
        if (rc != FFR_OK || *done) return rc;
        // wait for (or poll) the event just recorded
    }
}

int task_prepare_gang(ffr_scene *s, bool *eligible)
{
    // mixed precision with the multigrid-preconditioned solve: every kernel
    // of its step has a gang form; the FP64 form and plain CG run alone
    *eligible = false;
    if (s->fp64 || use_cg(s) || s->ctx->prof_on) return FFR_OK;
    if (!s->mg.levels) {
        const int rc = mg_create(s);
        if (rc != FFR_OK) return rc;
    }
    *eligible = true;
    return FFR_OK;
}

int scene_step(ffr_scene *s, ffr_step_result *out)
{
    int rc = task_begin(s, false);
    bool done = false;
    while (rc == FFR_OK && !done) rc = task_advance(s, true, &done);
    if (rc != FFR_OK) return rc;
    if (out) *out = s->task.result;
    return FFR_OK;
}

}  // namespace ffr
