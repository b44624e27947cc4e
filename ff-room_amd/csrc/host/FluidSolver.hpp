// FluidSolver.hpp - projection-method solver with the reference's interface
// (cpp/src/core/FluidSolver.hpp:9-84). step() and run() execute on the GPU
// through the C ABI (include/ffroom.h); the Grid's vectors are the host
// mirror of the device state (see Grid.hpp).
#pragma once
#include <functional>
#include <memory>

#include "BoundaryManager.hpp"
#include "Grid.hpp"
#include "PoissonSolver.hpp"

struct ffr_scene;

namespace ffroom {

enum class AdvectionScheme { UPWIND1 = 0, QUICK = 1, LAX_WENDROFF = 2 };

// Arithmetic of the device state and stencils (no counterpart in the reference).
enum class Precision { FP64 = 0, MIXED = 1 };
// Pressure solver choice (AUTO = the reference's: CG, or MG-PCG with use_multigrid).
enum class PoissonMethod { AUTO = 0, CG = 1, MGPCG = 2 };

struct FluidSolverParams {
    double rho = 1.2;
    double nu = 1.5e-5;
    double dt = 0.01;
    int max_steps = 500;
    double convergence_tol = 1.0;
    PoissonSolverParams poisson;
    AdvectionScheme advection = AdvectionScheme::UPWIND1;
    bool use_openmp = false;  // accepted for compatibility; no effect on the GPU
    bool use_multigrid = false;
    bool thermal = false;
    double T_initial = 20.0;
    double T_target = 1e30;
    bool buoyancy = false;
    double g_accel = 9.81;
    double beta = 1.0 / 293.0;
    double k_thermal = 0.026;
    double cp = 1005.0;
    // ---- extensions; the defaults reproduce the reference ----------------
    Precision precision = Precision::FP64;
    PoissonMethod poisson_method = PoissonMethod::AUTO;
    int mg_levels = 0;
    bool mg_bf16 = true;
    double inner_tol = 1e-4;
    int inner_max_iter = 60;
    bool sync_host_every_step = false;  // download the fields after every step()
};

struct StepResult {
    int step;
    double velocity_change;
    double pressure_residual;
    double divergence_max;
    bool converged;
    double T_mean = 0.0;
    int poisson_iterations = 0;  // extension
    bool nonfinite = false;      // extension: NaN/Inf seen in the fields
};

class DeviceScene;

class FluidSolver : public DeviceMirror {
public:
    FluidSolver(Grid &grid, BoundaryManager &bm, FluidSolverParams params = {});
    ~FluidSolver();
    FluidSolver(const FluidSolver &) = delete;
    FluidSolver &operator=(const FluidSolver &) = delete;

    StepResult step();
    StepResult run(std::function<void(const StepResult &)> callback = nullptr);

    double compute_divergence_max() const;
    double compute_T_mean() const;

    const FluidSolverParams &params() const { return params_; }
    int current_step() const { return step_; }

    // device mirror control (no counterpart in the reference)
    void sync_to_host() override;        // download u, v, w, p, T into the Grid if stale
    void mark_host_modified() override;  // upload the Grid's fields before the next step
    // snapshot egress (ffr_snapshot_begin/end): queue a copy of the current
    // state (FFR_SNAP_FIELDS or FFR_SNAP_CELLS) without waiting for it;
    // collect it later into the kind's arrays (returns its step)
    int snapshot_begin(int kind);
    int snapshot_kind(int slot) const { return snap_kind_[slot & 3]; }
    int snapshot_end(int slot, double *a, double *b, double *c, double *d, double *e);
    const Grid &grid() const { return grid_; }

private:
    Grid &grid_;
    BoundaryManager &bm_;
    FluidSolverParams params_;
    int step_ = 0;
    std::unique_ptr<DeviceScene> dev_;
    bool host_stale_ = false;  // device ahead of the Grid vectors
    bool host_dirty_ = false;  // Grid vectors ahead of the device
    uint64_t geom_seen_ = 0;
    int snap_kind_[4] = {0, 0, 0, 0};

    void push_if_dirty();
    void rebuild_device();
};

}  // namespace ffroom
