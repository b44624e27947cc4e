// ffr_kernels.hpp - launch wrappers of the gfx950 kernels (host side).
#pragma once
#include <hip/hip_runtime.h>
#include "ffr_device.hpp"

namespace ffr {

// Per-step scalar parameters (FluidSolverParams after the reference's
// derived constants, FluidSolver.cpp:84-85, 315, 333, 377-378, 390).
struct StepParams {
    double dt, nu, dt_rho, rhs_factor, buoy_coeff, T_ref, alpha_T;
    int scheme, buoyancy;
};

// Diagnostics record of one step, written by the kernels, read by the host.
struct StepRec {
    unsigned long long vmax_bits;    // max |u_new - u_old| over faces (bit pattern)
    unsigned long long divmax_bits;  // max |div u| over FLUID cells
    double tsum;                     // sum of T over FLUID cells
    int nonfinite;
    int pad;
};

// State of the device-resident reference CG (PoissonSolver.cpp:93-158).
struct CgState {
    double rr, bnorm, alpha, beta, res, dAd;
    int iter, done, iters, max_iter;
    double tol;
};

// Reduction scratch of one stream: block partials and arrival tickets.
struct RedScratch {
    double *partials;       // >= max grid size
    unsigned int *tickets;  // a few independent counters, zero at rest
};

// Upper bound on blocks per launch: one resident wave of 256-thread blocks
// over 256 CUs at 8 blocks each; larger arrays loop inside each block over a
// contiguous span (block_span), which also bounds the reduction partials.
extern unsigned g_max_blocks;
constexpr unsigned kMaxBlocksCap = 16384;

inline unsigned grid_for(uint64_t n)
{
    uint64_t g = (n + kBlock - 1) / kBlock;
    if (g > g_max_blocks) g = g_max_blocks;
    return (unsigned)(g ? g : 1);
}

// ---- step kernels (step_kernels.hip), R = double or float --------------
template <typename R>
void launch_predict(hipStream_t s, const Geo &g, const StepParams &sp, const R *u, const R *v,
                    const R *w, const R *T, R *us, R *vs, R *ws, StepRec *rec);
template <typename R>
void launch_correct(hipStream_t s, const Geo &g, const StepParams &sp, const R *us, const R *vs,
                    const R *ws, const double *p, R *u, R *v, R *w, StepRec *rec);
template <typename R>
void launch_divmax(hipStream_t s, const Geo &g, const R *u, const R *v, const R *w,
                   StepRec *rec);
template <typename R>
void launch_temperature(hipStream_t s, const Geo &g, const StepParams &sp, const R *u,
                        const R *v, const R *w, const R *T, R *Tn, StepRec *rec,
                        RedScratch red);
template <typename R>
void launch_tsum(hipStream_t s, const Geo &g, const R *T, StepRec *rec, RedScratch red);
template <typename R>
void launch_opening_T(hipStream_t s, const Geo &g, const R *u, const R *v, const R *w, R *T);
template <typename R>
void launch_rhs(hipStream_t s, const Geo &g, const StepParams &sp, const R *us, const R *vs,
                const R *ws, double *b);
template <typename R>
void launch_convert(hipStream_t s, const double *src, R *dst, uint64_t n);
template <typename R>
void launch_convert_back(hipStream_t s, const R *src, double *dst, uint64_t n);

// ---- reference CG in FP64 (cg_kernels.hip) -----------------------------
// b -> (x = 0, r = d = b with SOLID rows zeroed), ||b||; cg->done set when
// ||b|| < 1e-15 or max_iter == 0.
void launch_cg_init(hipStream_t s, const Geo &g, const double *b, double *x, double *r,
                    double *d, CgState *cg, RedScratch red, int max_iter, double tol);
// One CG iteration: Ad = A d (+ d.Ad), x/r update (+ r.r), d = r + beta d.
void launch_cg_iteration(hipStream_t s, const Geo &g, double *x, double *r, double *d,
                         double *Ad, CgState *cg, RedScratch red);
void launch_cg_spmv(hipStream_t s, const Geo &g, const double *d, double *Ad, CgState *cg,
                    RedScratch red);
void launch_cg_update(hipStream_t s, uint32_t n, double *x, double *r, const double *d,
                      const double *Ad, CgState *cg, RedScratch red);
void launch_cg_dir(hipStream_t s, uint32_t n, double *d, const double *r, CgState *cg);
// Plain A*x (PoissonSolver.cpp:22-91) in FP64.
void launch_laplacian(hipStream_t s, const Geo &g, const double *x, double *Ax);

}  // namespace ffr
