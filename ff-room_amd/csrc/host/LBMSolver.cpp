// LBMSolver.cpp - host side of the GPU lattice-Boltzmann solver
// (cpp/src/core/LBMSolver.cpp:36-60, 403-456).
#include "LBMSolver.hpp"

#include <stdexcept>

#include "Device.hpp"

namespace ffroom {

LBMSolver::LBMSolver(Grid &grid, BoundaryManager &bm, double tau, int max_steps,
                     double convergence_tol, double dt)
    : grid_(grid), bm_(bm), tau_(tau), convergence_tol_(convergence_tol), dt_(dt),
      max_steps_(max_steps)
{
    if (grid_.attached) grid_.attached->sync_to_host();
    ffr_params p;
    ffr_params_default(&p);
    p.dt = dt;
    p.precision = FFR_PREC_FP64;
    p.poisson_method = FFR_POISSON_CG;  // the lattice needs no pressure solver
    dev_.reset(new DeviceScene(grid_, &bm_, p));
    ffr_check(ffr_scene_upload(dev_->get(), grid_.u.data(), grid_.v.data(), grid_.w.data(),
                               grid_.p.data(), grid_.T.data()),
              "ffr_scene_upload");
    // lattice units: Grid::dx is the reference length (LBMSolver.cpp:43-48)
    ffr_check(ffr_lbm_attach(dev_->get(), tau, dt, convergence_tol, grid_.dx), "ffr_lbm_attach");
    geom_seen_ = grid_.geometry_version + bm_.version();
    grid_.attached = this;
}

LBMSolver::~LBMSolver()
{
    if (grid_.attached == this) {
        try {
            sync_to_host();
        } catch (...) {
        }
        grid_.attached = nullptr;
    }
}

void LBMSolver::sync_to_host()
{
    if (!host_stale_ || !dev_) return;
    ffr_check(ffr_scene_download(dev_->get(), grid_.u.data(), grid_.v.data(), grid_.w.data(),
                                 grid_.p.data(), nullptr),
              "ffr_scene_download");
    host_stale_ = false;
}

void LBMSolver::mark_host_modified() { host_dirty_ = true; }

void LBMSolver::push_if_dirty()
{
    if (grid_.geometry_version + bm_.version() != geom_seen_)
        // the lattice holds populations per cell; new flags or patches after
        // construction would need a new lattice
        throw std::runtime_error("LBMSolver: the geometry changed after construction; "
                                 "create a new solver");
    if (!host_dirty_) return;
    ffr_check(ffr_scene_upload(dev_->get(), grid_.u.data(), grid_.v.data(), grid_.w.data(),
                               grid_.p.data(), nullptr),
              "ffr_scene_upload");
    host_dirty_ = false;
}

StepResult LBMSolver::step()
{
    push_if_dirty();
    ffr_step_result r{};
    ffr_check(ffr_lbm_step(dev_->get(), &r), "ffr_lbm_step");
    host_stale_ = true;
    step_ = r.step;
    if (grid_.live_views > 0) {  // held views stay current; their edits go up next step
        sync_to_host();
        host_dirty_ = true;
    }
    StepResult out;
    out.step = r.step;
    out.velocity_change = r.velocity_change;
    out.pressure_residual = r.pressure_residual;
    out.divergence_max = r.divergence_max;
    out.converged = r.converged != 0;
    out.T_mean = r.T_mean;
    return out;
}

StepResult LBMSolver::run(std::function<void(const StepResult &)> callback)
{   // LBMSolver.cpp:446-454
    StepResult result{};
    for (int s = 0; s < max_steps_; s++) {
        result = step();
        if (callback) callback(result);
        if (result.converged) break;
    }
    sync_to_host();
    return result;
}

double LBMSolver::compute_divergence_max() const
{
    double d = 0.0;
    ffr_check(ffr_divergence_max(dev_->get(), &d), "ffr_divergence_max");
    return d;
}

}  // namespace ffroom
