// PoissonSolver.cpp - GPU pressure solves behind the reference interface
// (cpp/src/core/PoissonSolver.hpp:20-98).
#include "PoissonSolver.hpp"

#include <stdexcept>

#include "Device.hpp"

namespace ffroom {

PoissonSolver::PoissonSolver(PoissonSolverParams params) : params_(params) {}
PoissonSolver::~PoissonSolver() = default;

MultigridPoissonSolver::MultigridPoissonSolver(PoissonSolverParams params) : PoissonSolver(params)
{
    multigrid_ = true;
}

double PoissonSolver::solve(Grid &grid, const std::vector<double> &rhs)
{
    const size_t n = (size_t)grid.Nx * grid.Ny * grid.Nz;
    if (rhs.size() != n) throw std::invalid_argument("rhs size must be Nx*Ny*Nz");
    ffr_params prm;
    ffr_params_default(&prm);
    prm.poisson_max_iter = params_.max_iter;
    prm.poisson_tol = params_.tol;
    prm.use_multigrid = multigrid_ ? 1 : 0;
    prm.poisson_method = multigrid_ ? FFR_POISSON_MGPCG : FFR_POISSON_CG;
    // the scene (flags + spacings) is kept while the Grid's geometry is unchanged
    if (!cache_ || cache_->geometry_version() != grid.geometry_version)
        cache_.reset(new DeviceScene(grid, nullptr, prm));
    else
        ffr_check(ffr_scene_set_params(cache_->get(), &prm), "ffr_scene_set_params");
    grid.sync_to_host();
    ffr_solve_info info{};
    ffr_check(ffr_poisson_solve(cache_->get(), rhs.data(), grid.p.data(), &info), "ffr_poisson_solve");
    grid.mark_host_modified();
    last_iter_ = info.iterations;
    last_res_ = info.residual;
    return info.residual;
}

}  // namespace ffroom
