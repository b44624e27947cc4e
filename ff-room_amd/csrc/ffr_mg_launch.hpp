// ffr_mg_launch.hpp - launchers of the multigrid and defect-correction kernels.
#pragma once
#include "ffr_kernels.hpp"
#include "ffr_mg.hpp"

namespace ffr {
void mg_launch_build(hipStream_t s, const Geo &g, const MgLevel &F, const MgLevel &C, bool fine0);
void mg_launch_split_codes(hipStream_t s, const Geo &g, const MgLevel &L0);
// `dot` (level 0, red pass only): also write the r.z partials of the PCG
template <typename T>
void mg_launch_smooth(hipStream_t s, const Geo &g, const MgLevel &L, bool l0, int color,
                      int zero_start, const MgState *m, const RedScratch *dot = nullptr);
// level 0 with `scratch` (nc floats), and coarse levels of at least
// kCoarseTwoPassMin cells with it (restrict_two_pass): residual pass +
// child-sum pass; `pass` 1 or 2 launches only that pass of those (0: all)
template <typename TF, typename TC>
void mg_launch_restrict(hipStream_t s, const Geo &g, const MgLevel &F, const MgLevel &C, bool l0,
                        bool red0, const MgState *m, float *scratch = nullptr, int pass = 0);
// the level-0 residual pass of the restriction alone (profiling)
void mg_launch_resid0(hipStream_t s, const Geo &g, const MgLevel &L0, float *res, const MgState *m);
// prolongation fused with the first post-smoothing (black) sweep on F
template <typename TF, typename TC>
void mg_launch_post_black(hipStream_t s, const Geo &g, const MgLevel &F, const MgLevel &C, bool l0,
                          const MgState *m);
// allows the tail kernel kTailLdsMax bytes of dynamic LDS (0 on success)
int mg_tail_allow_lds();
template <typename T>
void mg_launch_tail(hipStream_t s, const TailLevels &tl, size_t lds_bytes, const MgState *m);
template <typename T>
void mg_launch_coarse(hipStream_t s, const Geo &g, const MgLevel &L, bool l0, int sweeps,
                      const MgState *m);
// main form (the V-cycle's last pass is a red level-0 pass): p = z + beta p
// and q = A z + beta q in place; gdot = number of r.z partials
void pcg_launch_dirq(hipStream_t s, const Geo &g, const MgLevel &L0, const float *r, float *p,
                     float *q, MgState *m, RedScratch red, uint32_t gdot);
// general form: p = z + beta p, then q = A p
void pcg_launch_dir(hipStream_t s, uint32_t n, const float *z, float *p, MgState *m, RedScratch red,
                    uint32_t gdot);
void pcg_launch_spmv(hipStream_t s, const Geo &g, const MgLevel &L0, const float *p, float *q,
                     const MgState *m, RedScratch red);
// gq = number of p.q partials (the grid of dirq or spmv)
void pcg_launch_update(hipStream_t s, const Geo &g, const MgLevel &L0, float *e, float *r,
                       const float *p, const float *q, MgState *m, RedScratch red, uint32_t gq);
void pcg_launch_check(hipStream_t s, uint32_t n, MgState *m, RedScratch red);
void pcg_launch_dot_rz(hipStream_t s, uint32_t n, const float *r, const float *z, const MgState *m,
                       RedScratch red);
// warm 0: x starts at 0; 1: x keeps its values (the previous pressure p_n).
// With `out`: the start goes there instead (warm 1: p_n; 2: 2 p_n - h1; 3:
// 3 p_n - 3 h1 + out's old values, extrapolated in time from the pressures of
// the steps before, h1 = p_{n-1}, out = p_{n-2}), and x keeps p_n
void dc_launch_init(hipStream_t s, const Geo &g, const double *b, double *x, int warm,
                    const double *h1, double *out, RedScratch red);
void dc_launch_start(hipStream_t s, const Geo &g, MgState *m, RedScratch red, double outer_tol,
                     int max_outer, int max_inner, double inner_tol, int total_cap);
// one outer pass (x += e, the FP64 residual); `first`: the solve's first pass
// (e is 0 then, and x stays in xa)
void dc_launch_residual(hipStream_t s, const Geo &g, const double *b, double *xa, double *xb,
                        const float *e, float *r32, float *z0, const MgState *m, RedScratch red,
                        bool first);
void dc_launch_check(hipStream_t s, const Geo &g, MgState *m, RedScratch red);
// p from the last outer pass's output (in xa or xb by the pass count); with
// `inplace` that output is p itself (xa == p) and only OUTFLOW rows are written
void dc_launch_finalize(hipStream_t s, const Geo &g, const double *b, const double *xa,
                        const double *xb, double *p, int pin_outflow, bool inplace,
                        const MgState *m);
}  // namespace ffr
