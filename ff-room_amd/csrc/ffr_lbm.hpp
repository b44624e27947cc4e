// ffr_lbm.hpp - D3Q19 lattice Boltzmann on a scene (LBMSolver.hpp/.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include "ffr_kernels.hpp"

namespace ffr {

// one outflow or opening patch: the wall (axis, side) and the inclusive
// ranges of its two in-wall axes in order
struct LbmPatch {
    int axis, side, a0, a1, b0, b1;
};

// the lattice state a scene carries once ffr_lbm_attach has run
struct LbmState {
    bool on = false;
    double tau = 0.0, dt = 0.0, tol = 0.0, dx_ref = 0.0;
    double *f = nullptr, *f2 = nullptr;  // 19 * nc each
    double *cu = nullptr;                // cell velocities ux | uy | uz, nc each
    uint32_t *mask = nullptr;            // open neighbours per cell (lbm_kernels.hip)
    // fused form: f holds the next step's populations after inflow and
    // collision (k_lbm_stream_collide); off when a patch copies from an
    // INFLOW cell, where f holds the streamed populations between steps
    bool fused = false;
    void *arena = nullptr;
    size_t bytes = 0;
    int steps = 0;
};

void lbm_launch_rest(hipStream_t s, uint32_t nc, double *f);
void lbm_launch_relax(hipStream_t s, const Geo &g, double *f, double s2l, double inv_tau,
                      bool collide, StepRec *rec);
void lbm_launch_mask(hipStream_t s, const Geo &g, uint32_t *mask);
// streaming, and the moments of the result (cell velocity [m/s] = lattice
// velocity * l2s, p = rho / 3)
void lbm_launch_stream(hipStream_t s, const Geo &g, const uint32_t *mask, const double *f,
                       double *f2, double *cu, double *p, double l2s);
void lbm_launch_patch(hipStream_t s, const Geo &g, double *f, const LbmPatch &pt, double *cu,
                      double *p, double l2s);
// fused form: streaming, moments, then the next step's inflow and collision
void lbm_launch_stream_collide(hipStream_t s, const Geo &g, const uint32_t *mask, const double *f,
                               double *f2, double *cu, double *p, double l2s, double s2l,
                               double inv_tau, StepRec *rec);
// fused form's patch copy (post-collision populations and moments)
void lbm_launch_patch_post(hipStream_t s, const Geo &g, double *f, const LbmPatch &pt, double *cu,
                           double *p);
void lbm_launch_faces(hipStream_t s, const Geo &g, const double *cu, double *u, double *v,
                      double *w, StepRec *rec);

}  // namespace ffr
