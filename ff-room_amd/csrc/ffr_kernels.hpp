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
// rr is double-buffered by iteration parity: the kernel of iteration `it`
// reads rr[it & 1] while its block 0 writes rr[(it + 1) & 1].
struct CgState {
    double rr[2];
    double bnorm, res;
    int done, iters, max_iter;
    int started;  // iterations begun (bumped by the SpMV, the first kernel of one)
    double tol;
};

// Reduction scratch of one stream: partial sums, kSlots independent slots of
// kMaxBlocksCap doubles (a producer kernel writes one slot, its consumer
// reads it).
struct RedScratch {
    double *partials;
    unsigned int *tickets;  // unused counters kept zero
};
constexpr int kSlots = 6;
enum { SLOT_INIT = 0, SLOT_SPMV = 1, SLOT_UPDATE = 2, SLOT_TSUM = 3, SLOT_MG_A = 4, SLOT_MG_B = 5 };

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

// Kernels that gather neighbours run best as one resident round of blocks
// (8 waves per SIMD): their grid stays at that capacity even when the cap
// above is raised for the streaming kernels.
extern unsigned g_resident_blocks;
inline unsigned grid_resident(uint64_t n)
{
    const unsigned g = grid_for(n);
    return g < g_resident_blocks ? g : g_resident_blocks;
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
// temperature and tsum return the block count of their launch: the number of
// SLOT_TSUM partials they wrote
template <typename R>
unsigned launch_temperature(hipStream_t s, const Geo &g, const StepParams &sp, const R *u,
                            const R *v, const R *w, const R *T, R *Tn, StepRec *rec,
                            RedScratch red);
template <typename R>
unsigned launch_tsum(hipStream_t s, const Geo &g, const R *T, StepRec *rec, RedScratch red);
// rec->tsum := sum of the G SLOT_TSUM partials written by temperature / tsum.
void launch_step_finalize(hipStream_t s, StepRec *rec, RedScratch red, unsigned G);
template <typename R>
void launch_opening_T(hipStream_t s, const Geo &g, const R *u, const R *v, const R *w, R *T);
template <typename R>
void launch_rhs(hipStream_t s, const Geo &g, const StepParams &sp, const R *us, const R *vs,
                const R *ws, double *b);
template <typename R>
void launch_convert(hipStream_t s, const double *src, R *dst, uint64_t n);
template <typename R>
void launch_convert_back(hipStream_t s, const R *src, double *dst, uint64_t n);

// face mask codes of a scene (once, at creation)
void launch_face_codes(hipStream_t s, const Geo &g, uint16_t *fcode);
// Cell flags from marks (type, i0, i1, j0, j1, k0, k1) applied in order, and
// the neighbour codes of the pressure operator from the flags (at creation).
constexpr int kMaxMarks = 64;
struct MarkList {
    int n;
    int m[kMaxMarks][7];
};
void launch_mark_cells(hipStream_t s, int nx, int ny, int nz, const MarkList &ml, uint8_t *flags);
void launch_nb_codes(hipStream_t s, int nx, int ny, int nz, const uint8_t *flags, uint16_t *nbm);
// plain FP64 copy (calibration of memory counters)
void launch_copy64(hipStream_t s, const double *src, double *dst, uint64_t n);

// ---- reference CG in FP64 (cg_kernels.hip) -----------------------------
// b -> (x = 0, r = d = b with SOLID rows zeroed), ||b||; cg->done set when
// ||b|| < 1e-15 or max_iter == 0.
void launch_cg_init(hipStream_t s, const Geo &g, const double *b, double *x, double *r,
                    double *d, CgState *cg, RedScratch red, int max_iter, double tol);
// Iteration `it` of CG: Ad = A d (+ d.Ad), x/r update (+ r.r), d = r + beta d.
void launch_cg_spmv(hipStream_t s, const Geo &g, const double *d, double *Ad, CgState *cg,
                    RedScratch red);
void launch_cg_update(hipStream_t s, uint32_t n, double *x, double *r, const double *d,
                      const double *Ad, CgState *cg, RedScratch red);
void launch_cg_dir(hipStream_t s, uint32_t n, double *d, const double *r, CgState *cg,
                   RedScratch red);
// Plain A*x (PoissonSolver.cpp:22-91) in FP64.
void launch_laplacian(hipStream_t s, const Geo &g, const double *x, double *Ax);

}  // namespace ffr
