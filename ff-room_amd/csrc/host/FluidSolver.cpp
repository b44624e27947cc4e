// FluidSolver.cpp - the projection step on the GPU behind the reference
// interface (cpp/src/core/FluidSolver.cpp:18-32, 486-560).
#include "FluidSolver.hpp"

#include <algorithm>
#include <stdexcept>

#include "Device.hpp"

namespace ffroom {

namespace {
ffr_params to_ffr(const FluidSolverParams &p)
{
    ffr_params o;
    ffr_params_default(&o);
    o.rho = p.rho; o.nu = p.nu; o.dt = p.dt; o.max_steps = p.max_steps;
    o.convergence_tol = p.convergence_tol;
    o.poisson_max_iter = p.poisson.max_iter;
    o.poisson_tol = p.poisson.tol;
    o.advection = static_cast<int>(p.advection);
    o.use_openmp = p.use_openmp ? 1 : 0;
    o.use_multigrid = p.use_multigrid ? 1 : 0;
    o.thermal = p.thermal ? 1 : 0;
    o.T_initial = p.T_initial; o.T_target = p.T_target;
    o.buoyancy = p.buoyancy ? 1 : 0;
    o.g_accel = p.g_accel; o.beta = p.beta; o.k_thermal = p.k_thermal; o.cp = p.cp;
    o.precision = p.precision == Precision::MIXED ? FFR_PREC_MIXED : FFR_PREC_FP64;
    o.poisson_method = static_cast<int>(p.poisson_method);
    o.mg_levels = p.mg_levels;
    o.mg_bf16 = p.mg_bf16 ? 1 : 0;
    o.inner_tol = p.inner_tol;
    o.inner_max_iter = p.inner_max_iter;
    return o;
}
}  // namespace

FluidSolver::FluidSolver(Grid &grid, BoundaryManager &bm, FluidSolverParams params)
    : grid_(grid), bm_(bm), params_(params)
{
    // ctor semantics of the reference (FluidSolver.cpp:27-31): T := T_initial,
    // then the opening temperatures when thermal.
    if (grid_.attached) grid_.attached->sync_to_host();
    std::fill(grid_.T.begin(), grid_.T.end(), params_.T_initial);
    if (params_.thermal) bm_.apply_opening_temperature(grid_);
    // optional fan inlet temperatures: the fan's INFLOW cells carry T_inlet
    for (auto &ft : bm_.fan_temperatures()) {
        const FanBC &f = bm_.fans()[ft.first];
        for (int i = std::max(0, f.i_min); i <= std::min(grid_.Nx - 1, f.i_max); i++)
            for (int j = std::max(0, f.j_min); j <= std::min(grid_.Ny - 1, f.j_max); j++)
                for (int k = std::max(0, f.k_min); k <= std::min(grid_.Nz - 1, f.k_max); k++) {
                    const int c = grid_.c_idx(i, j, k);
                    if (grid_.cell_type[c] == CellType::INFLOW) grid_.T[c] = ft.second;
                }
    }
    grid_.attached = this;
    rebuild_device();
}

FluidSolver::~FluidSolver()
{
    if (grid_.attached == this) {
        try {
            sync_to_host();
        } catch (...) {
        }
        grid_.attached = nullptr;
    }
}

void FluidSolver::rebuild_device()
{
    if (dev_ && host_stale_) sync_to_host();
    dev_.reset(new DeviceScene(grid_, &bm_, to_ffr(params_)));
    ffr_check(ffr_scene_upload(dev_->get(), grid_.u.data(), grid_.v.data(), grid_.w.data(),
                               grid_.p.data(), grid_.T.data()),
              "ffr_scene_upload");
    ffr_check(ffr_scene_set_step_count(dev_->get(), step_), "ffr_scene_set_step_count");
    geom_seen_ = grid_.geometry_version + bm_.version();
    host_stale_ = false;
    host_dirty_ = false;
}

void FluidSolver::sync_to_host()
{
    if (!host_stale_ || !dev_) return;
    ffr_check(ffr_scene_download(dev_->get(), grid_.u.data(), grid_.v.data(), grid_.w.data(),
                                 grid_.p.data(), grid_.T.data()),
              "ffr_scene_download");
    host_stale_ = false;
}

void FluidSolver::mark_host_modified() { host_dirty_ = true; }

int FluidSolver::snapshot_begin(int kind)
{
    push_if_dirty();
    int slot = -1;
    ffr_check(ffr_snapshot_begin(dev_->get(), kind, &slot), "ffr_snapshot_begin");
    snap_kind_[slot & 3] = kind;
    return slot;
}

int FluidSolver::snapshot_end(int slot, double *a, double *b, double *c, double *d, double *e)
{
    int step = 0;
    ffr_check(ffr_snapshot_end(dev_->get(), slot, &step, a, b, c, d, e), "ffr_snapshot_end");
    return step;
}

void FluidSolver::push_if_dirty()
{
    if (grid_.geometry_version + bm_.version() != geom_seen_) {
        rebuild_device();  // flags or boundary lists changed since the last step
        return;
    }
    if (!host_dirty_) return;
    ffr_check(ffr_scene_upload(dev_->get(), grid_.u.data(), grid_.v.data(), grid_.w.data(),
                               grid_.p.data(), grid_.T.data()),
              "ffr_scene_upload");
    host_dirty_ = false;
}

StepResult FluidSolver::step()
{
    push_if_dirty();
    ffr_step_result r{};
    ffr_check(ffr_step(dev_->get(), &r), "ffr_step");
    host_stale_ = true;
    step_ = r.step;
    if (params_.sync_host_every_step || grid_.live_views > 0) sync_to_host();
    // a held view may be written before the next step: upload it then
    if (grid_.live_views > 0) host_dirty_ = true;
    StepResult out;
    out.step = r.step;
    out.velocity_change = r.velocity_change;
    out.pressure_residual = r.pressure_residual;
    out.divergence_max = r.divergence_max;
    out.converged = r.converged != 0;
    out.T_mean = r.T_mean;
    out.poisson_iterations = r.poisson_iters;
    out.nonfinite = r.nonfinite != 0;
    return out;
}

StepResult FluidSolver::run(std::function<void(const StepResult &)> callback)
{   // FluidSolver.cpp:552-560
    StepResult result{};
    for (int s = 0; s < params_.max_steps; s++) {
        result = step();
        if (callback) callback(result);
        if (result.converged) break;
    }
    sync_to_host();
    return result;
}

double FluidSolver::compute_divergence_max() const
{
    auto *self = const_cast<FluidSolver *>(this);
    self->push_if_dirty();
    double d = 0.0;
    ffr_check(ffr_divergence_max(dev_->get(), &d), "ffr_divergence_max");
    return d;
}

double FluidSolver::compute_T_mean() const
{
    auto *self = const_cast<FluidSolver *>(this);
    self->push_if_dirty();
    double t = 0.0;
    ffr_check(ffr_T_mean(dev_->get(), &t), "ffr_T_mean");
    return t;
}

}  // namespace ffroom
