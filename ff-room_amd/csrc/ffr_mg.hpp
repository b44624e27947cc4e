// ffr_mg.hpp - multigrid-preconditioned CG for the pressure equation.
//
// The system solved is the reference operator (PoissonSolver.cpp:22-91)
// restricted to the unknown cells (FLUID and INFLOW): OUTFLOW neighbours are
// Dirichlet (2*scale on the diagonal), SOLID neighbours and walls Neumann.
// Rows of OUTFLOW cells only read unknown neighbours, so their pressure is
// recovered afterwards by back substitution (the reference CG's converged
// value), or pinned to 0 when use_multigrid asks for the reference
// multigrid's convention (PoissonSolver.cpp:201).
//
// Hierarchy: 2x2x2 aggregation (odd extents keep a thin last aggregate).
// Level 0 evaluates its stencil from the cell flags and the per-axis
// coefficients; every coarser level l+1 stores a 7-point operator
//   A_{l+1} = (1/2) P^T A_l P   (P piecewise-constant interpolation)
// as one coefficient per cell and axis for the face towards the lower
// neighbour plus a diagonal. The 1/2 makes the Galerkin aggregate operator
// agree with a re-discretisation on the 2h grid, so the V-cycle needs no
// separate over-correction. Smoother: red-black Gauss-Seidel, red then
// black before the coarse correction and black then red after it, which
// keeps the preconditioner symmetric; restriction is P^T.
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
    // level >= 1: operator coefficients (FP32)
    float *cx, *cy, *cz;   // coupling to the -x/-y/-z neighbour (>= 0), 0 if none
    float *diag;           // 1 for cells with no unknown child
    float *rdiag;          // 1/diag
    // vectors: FP32 on level 0, BF16 or FP32 on coarser levels
    void *z, *r;
    int bf16;
};

struct Multigrid {
    int levels = 0;
    MgLevel lv[kMaxLevels];
    int tail_level = -1;   // first level solved by the single-workgroup tail kernel
    // level-0 vectors of the inner FP32 PCG
    float *e, *ri, *pi, *qi;
    uint8_t *active0;      // 1 for unknown cells of level 0
};

// Scalars of the inner PCG and the outer FP64 defect correction.
struct MgState {
    double rz, rr, rr0, pq, alpha, beta;
    double bnorm, res_outer, norm_r64;
    int iter, inner_done, outer, solve_done;
    int total_iters, inner_max, outer_max, pad;
    double inner_tol, outer_tol;
};

}  // namespace ffr
