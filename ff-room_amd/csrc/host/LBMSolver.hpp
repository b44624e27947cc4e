// LBMSolver.hpp - the reference's D3Q19 lattice-Boltzmann solver behind its
// own interface (cpp/src/core/LBMSolver.hpp:17-39), run on the GPU through
// ffr_lbm_attach / ffr_lbm_step. The Grid's u, v, w, p mirror the device
// (synchronised as FluidSolver does it); T is left alone, as the reference's
// isothermal solver leaves it.
#pragma once
#include <functional>
#include <memory>

#include "BoundaryManager.hpp"
#include "FluidSolver.hpp"  // StepResult
#include "Grid.hpp"

namespace ffroom {

class DeviceScene;

class LBMSolver : public DeviceMirror {
public:
    // tau: BGK relaxation time; dt: physical time step [s] (lattice velocity
    // = physical velocity * dt / Grid::dx); max_steps and convergence_tol as
    // in FluidSolver
    LBMSolver(Grid &grid, BoundaryManager &bm, double tau, int max_steps, double convergence_tol,
              double dt = 0.01);
    ~LBMSolver();
    LBMSolver(const LBMSolver &) = delete;
    LBMSolver &operator=(const LBMSolver &) = delete;

    StepResult step();
    StepResult run(std::function<void(const StepResult &)> callback = nullptr);
    double compute_divergence_max() const;
    double compute_T_mean() const { return 0.0; }  // isothermal
    int current_step() const { return step_; }

    void sync_to_host() override;
    void mark_host_modified() override;

private:
    Grid &grid_;
    BoundaryManager &bm_;
    double tau_, convergence_tol_, dt_;
    int max_steps_;
    int step_ = 0;
    std::unique_ptr<DeviceScene> dev_;
    bool host_stale_ = false;
    bool host_dirty_ = false;
    uint64_t geom_seen_ = 0;
    void push_if_dirty();
};

}  // namespace ffroom
