// PoissonSolver.hpp - pressure solvers with the reference's interface
// (cpp/src/core/PoissonSolver.hpp:14-98). Both run on the GPU:
//   PoissonSolver            unpreconditioned FP64 CG, the reference algorithm
//                            (PoissonSolver.cpp:93-158), x0 = 0, same stopping
//                            test ||r||/||b|| < tol and the same operator.
//   MultigridPoissonSolver   multigrid-preconditioned CG to ||r||/||b|| < tol
//                            with the reference multigrid's convention of
//                            zero pressure on OUTFLOW cells (PoissonSolver.cpp:201).
#pragma once
#include <memory>
#include <vector>

#include "Grid.hpp"

struct ffr_scene;

namespace ffroom {

struct PoissonSolverParams {
    int max_iter = 1000;
    double tol = 1e-6;
    bool use_openmp = false;  // accepted for compatibility; no effect
};

class DeviceScene;

class PoissonSolver {
public:
    explicit PoissonSolver(PoissonSolverParams params = {});
    ~PoissonSolver();
    double solve(Grid &grid, const std::vector<double> &rhs);
    int last_iterations() const { return last_iter_; }
    double last_residual() const { return last_res_; }

protected:
    PoissonSolverParams params_;
    int last_iter_ = 0;
    double last_res_ = 0.0;
    bool multigrid_ = false;
    std::unique_ptr<DeviceScene> cache_;
};

class MultigridPoissonSolver : public PoissonSolver {
public:
    explicit MultigridPoissonSolver(PoissonSolverParams params = {});
};

}  // namespace ffroom
