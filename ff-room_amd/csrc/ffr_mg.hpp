// ffr_mg.hpp - multigrid-preconditioned CG for the pressure equation.
//
// The system solved is the reference operator (PoissonSolver.cpp:22-91)
// restricted to the unknown cells (FLUID and INFLOW): OUTFLOW neighbours are
// Dirichlet (2*scale on the diagonal), SOLID neighbours and walls Neumann.
// Rows of OUTFLOW cells only read unknown neighbours, so their pressure is
// recovered afterwards by back substitution (the value the reference CG
// converges to), or pinned to 0 when use_multigrid asks for the reference
// multigrid's convention (PoissonSolver.cpp:201).
//
// Outer loop (FP64): defect correction x += e with r = b - A x and its norm
// in FP64, stopping at ||r|| / ||b|| < poisson_tol (the reference's test).
// Inner solve (FP32 vectors, FP64 dot accumulation): CG preconditioned by one
// V-cycle, to inner_tol relative to its start.
//
// Hierarchy: 2x2x2 aggregation (an odd extent keeps a one-cell-thick last
// aggregate). Level 0 evaluates its stencil from the cell flags and the
// per-axis coefficients; every coarser level stores the scaled Galerkin
// operator A_{l+1} = (1/2) P^T A_l P of piecewise-constant interpolation P as
// one coupling per cell and axis (to the lower neighbour) plus a diagonal.
// The factor 1/2 makes aggregation agree with re-discretisation on the 2h
// grid. Smoother: red-black Gauss-Seidel, red then black before the coarse
// correction and black then red after it; restriction is P^T and the
// coarsest level runs symmetric red-black sweeps, so the preconditioner is
// symmetric. Levels >= 1 keep z and r in bfloat16 when mg_bf16 is set
// (arithmetic in FP32).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "ffr_device.hpp"

namespace ffr {

constexpr int kMaxLevels = 8;

struct MgLevel {
    int nx, ny, nz;
    uint32_t n;
    FastDiv f_nz, f_nynz;
    // level >= 1: stored operator (FP32)
    uint8_t *active;       // 1 for unknowns
    float *cx, *cy, *cz;   // coupling to the -x/-y/-z neighbour (>= 0), 0 if none
    float *diag, *rdiag;   // 1 for cells with no unknown child
    // vectors: FP32 on level 0, bfloat16 or FP32 on coarser levels
    void *z, *r;
    int bf16;
};

struct Multigrid {
    int levels = 0;
    MgLevel lv[kMaxLevels];
    void *arena = nullptr;
    size_t bytes = 0;
    // level-0 vectors of the inner PCG (FP32) and the outer correction
    float *e, *ri, *pi, *qi;
    int pre_sweeps = 2, level_sweeps = 2, coarse_sweeps = 8;
    int inner_pred[16] = {0};  // inner iterations seen per outer pass (launch prediction)
};

// Scalars of the inner PCG and the outer FP64 defect correction. Values the
// inner iteration `it` reads are double-buffered by parity (see CgState).
struct MgState {
    double rz[2];
    double bnorm, res, rnorm0, inner_tol, outer_tol;
    int solve_done, inner_done, outer, inner_it;
    int total_iters, max_outer, max_inner, pad;
};

}  // namespace ffr
