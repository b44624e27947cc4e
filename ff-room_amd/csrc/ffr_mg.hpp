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
// Hierarchy: aggregation of 2 cells along each coarsened axis (an odd extent
// keeps a one-cell-thick last aggregate). Each level coarsens the axes whose
// spacing is at most 1/sqrt(2) of the largest (semi-coarsening, so that strongly
// coupled directions of anisotropic cells are coarsened first), or every axis
// when the cells are near-cubic; an axis with fewer than 4 cells is kept.
// Level 0 evaluates its stencil from the cell flags and the per-axis
// coefficients; every coarser level stores the Galerkin operator of
// piecewise-constant interpolation P, P^T A_l P, with the couplings and the
// diagonal part along each axis scaled by 1/2 when that axis was coarsened
// (1 otherwise): this makes aggregation agree with re-discretisation on the
// coarser grid. One coupling per cell and axis (to the lower neighbour), the
// per-axis diagonal parts and their sum are stored. Smoother: red-black Gauss-Seidel, red then black before the coarse
// correction and black then red after it; restriction is P^T and the
// coarsest level runs symmetric red-black sweeps, so the preconditioner is
// symmetric. Levels >= 1 keep z and r in bfloat16 when mg_bf16 is set
// (arithmetic in FP32).
//
// Level-0 FP32 vectors (z, r and the inner PCG's e, p, q) use a red-black
// split layout: cell c (natural index, k fastest) of colour (i+j+k)&1 lives
// at colour*H + (c >> 1), H = ceil(nc/2). Cells 2m and 2m+1 always differ in
// colour (same row when nz is even; otherwise the colour is the parity of c,
// or of c+i when ny is even and planes hold an even count), so each colour
// array is dense and every neighbour of a cell lies in the other colour's
// array at (c +- stride) >> 1. A colour pass of the smoother then streams
// only its own colour's r and codes and the other colour's z. Used slots are
// exactly [0, nc): red [0, H), black [H, nc); the slot nc (odd nc) stays 0.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "ffr_device.hpp"

namespace ffr {

constexpr int kMaxLevels = 12;
// coarse levels of at least this many cells restrict in two passes (the
// residual one thread per fine cell, then the child sums); smaller ones in
// one fused launch; level 0 always in two passes
constexpr uint32_t kCoarseTwoPassMin = 1u << 20;
inline bool restrict_two_pass(bool l0, uint32_t n) { return l0 || n >= kCoarseTwoPassMin; }

struct MgLevel {
    int nx, ny, nz;
    uint32_t n;
    FastDiv f_nz, f_nynz;
    // level >= 1: stored operator (FP32)
    uint8_t *active;       // 1 for unknowns
    float *cx, *cy, *cz;   // coupling to the -x/-y/-z neighbour (>= 0), 0 if none
    float *diag, *rdiag;   // 1 for cells with no unknown child
    float *dgx, *dgy, *dgz;  // diagonal parts along each axis (sum = diag)
    int shx, shy, shz;     // 1 where this level halved that axis of the finer one
    // vectors: FP32 on level 0, bfloat16 or FP32 on coarser levels
    void *z, *r;
    int bf16;
    // level 0: neighbour codes in split layout (NB_COLOR holds c & 1) and H.
    // The parity of the natural index of slot h of colour col is
    //   col ^ (nz even ? (h / (nz/2)) & 1 : 0) ^ (ny even ? (h / (ny*nz/2)) & 1 : 0)
    // (the row, and the plane, of cell 2h and 2h+1 coincide), so the loads of
    // a cell need not wait for its code.
    uint16_t *nbs;
    uint32_t H;
    int par_nz, par_ny;
    FastDiv f_hz, f_hyz;
};

struct Multigrid {
    int levels = 0;
    MgLevel lv[kMaxLevels];
    void *arena = nullptr;
    size_t bytes = 0;
    // level-0 vectors of the inner PCG (FP32) and the outer correction; sc
    // holds the level-0 residual of the V-cycle's restriction
    float *e, *ri, *pi, *qi, *sc;
    // per-axis factors of level 0 on a stretched grid, per index of x | y | z:
    // {w = d / mean width, t of the face below, t of the face above, 0}, with
    // t = 1 / (centre distance * mean width), 0 at the walls
    float4 *mgf;
    int pre_sweeps = 2, level_sweeps = 2, coarse_sweeps = 8;
    // levels tail_first .. levels-1 run as one single-workgroup kernel;
    // levels tail_first .. tail_lds_from-1 keep z, r and the operator in
    // global memory (mg_tail_cells), the others their z and r in LDS (-1:
    // no tail); tail_lds = its LDS bytes; the operators of levels
    // tail_staged .. levels-1 are staged in LDS too
    int tail_first = -1, tail_staged = -1, tail_lds_from = -1;
    size_t tail_lds = 0;
    // workgroups that share the tail's global levels (1: one workgroup), all
    // on XCD tail_xcd (see TailLevels)
    int tail_groups = 1, tail_xcd = 0;
    // dense (generalised) inverse of the coarsest level's operator, FP32,
    // row-major n x n, n <= kCoarsest: the tail's coarse solve is one
    // matrix-vector product (null when levels == 1)
    float *ginv = nullptr;
    // captured graphs: one inner PCG iteration, and the start of an inner
    // solve (V-cycle on r and r.z); built on first use when enabled
    hipGraphExec_t g_iter = nullptr, g_start = nullptr;
    bool graphs_tried = false;
};

// The levels a single workgroup runs from LDS (see k_mg_tail). Level l
// occupies LDS bytes [off[l], off[l] + tail_carve(...).end): a staged level
// holds its operator (cx, cy, cz padded at the top by si, sj and 1 zeros,
// diag, rdiag, active), every level its z (padded by si zeros at both ends,
// so a cell reads its six neighbours without bounds tests: the coupling
// across a wall is 0) and its r (vs bytes each). The first tail level may be unstaged: its
// operator is then read from global memory (it stays in L2) and only z and
// r take LDS. Levels before lds_from (mg_tail_cells) keep everything in
// global memory: their z arrays carry si zero pads at both ends in the arena
// as the LDS copies do, and their passes are ordered by the workgroup's
// barriers like the LDS ones (one workgroup, one CU).
constexpr int kTailThreads = 1024;
// the hierarchy stops at a level of at most this many cells; such a coarsest
// level is solved with a dense inverse
constexpr int kCoarsest = 512;
constexpr size_t kTailLdsDefault = 96 * 1024, kTailLdsMax = 150 * 1024;
// an unstaged first tail level is taken only up to this many cells: beyond
// it, one workgroup's passes cost more than the kernel launches they replace
constexpr uint32_t kTailUnstagedMax = 16384;
// At most this many workgroups share a tail: each takes a whole CU (1024
// threads and the tail's LDS), and four tails running at once (one per
// hardware queue) on one XCD then still fit its 32 CUs, so a group never
// waits for a workgroup that cannot become resident. Automatic choice: one
// workgroup per kTailGroupCells cells of the first (largest) global level.
constexpr int kTailGroupsMax = 8;
constexpr uint32_t kTailGroupCells = 16384;
constexpr unsigned kTailSpinMax = 1u << 21;
struct TailCarve {
    uint32_t cx, cy, cz, diag, rdiag, act, z, r, end;  // byte offsets in the level's region
};
__host__ __device__ inline TailCarve tail_carve(uint32_t n, uint32_t si, uint32_t sj, uint32_t vs,
                                                bool staged)
{
    auto al = [](uint32_t b) { return (b + 15u) & ~15u; };
    TailCarve t{};
    uint32_t o = 0;
    if (staged) {
        t.cx = o; o += al(4u * (n + si));
        t.cy = o; o += al(4u * (n + sj));
        t.cz = o; o += al(4u * (n + 1));
        t.diag = o; o += al(4u * n);
        t.rdiag = o; o += al(4u * n);
        t.act = o; o += al(n);
    }
    t.z = o + vs * si;  // element 0; the pads are z[-si .. 0) and z[n .. n + si)
    o += al(vs * (n + 2 * si));
    t.r = o; o += al(vs * n);
    t.end = o;
    return t;
}
inline size_t tail_scratch_bytes(uint32_t n) { return ((size_t)4 * n + 15) & ~(size_t)15; }
inline size_t tail_level_bytes(const MgLevel &L, size_t vs, bool staged)
{
    return tail_carve(L.n, (uint32_t)(L.ny * L.nz), (uint32_t)L.nz, (uint32_t)vs, staged).end;
}
struct TailLevels {
    MgLevel lv[kMaxLevels];
    uint32_t off[kMaxLevels];
    uint32_t scratch_off; // FP32 residuals of an LDS level being restricted (size of lds_from)
    float *gscratch;      // FP32 residuals of a global level being restricted (>= its n)
    int first, last;
    int lds_from;         // first level with its z and r in LDS (levels before: global)
    int staged_from;      // first level whose operator is in LDS (first or first + 1)
    int sweeps, coarse_sweeps;
    const float *ginv;    // dense inverse of the coarsest operator (n_last^2), or null
    // groups > 1: the global levels' phases are shared by that many
    // workgroups, blocks b of the launch with b % kXcds == xcd (one XCD, so
    // one L2), ordered by a barrier on the counters gsync[0] (arrivals) and
    // gsync[1] (departures, the last one resets both); the LDS levels run on
    // the first of them. A barrier that waits longer than kTailSpinMax polls
    // sets *fault and stops waiting, so every workgroup still ends.
    int groups, xcd;
    unsigned *gsync;
    int *fault;
};
static_assert(sizeof(TailLevels) <= 4096, "kernel arguments are limited to 4 KiB");

// Scalars of the inner PCG and the outer FP64 defect correction, all kept on
// the device so that launches carry no per-iteration arguments (a solve can
// be enqueued ahead of the host, or replayed from a captured graph). rz is
// double-buffered by the parity of inner_it, the iteration count of the
// current inner solve, which only the one-block check kernel advances.
// passes counts FP64 residual evaluations; x ping-pongs between the two
// outer buffers by its parity. hist[k] = iterations of inner solve k.
constexpr int kMgHist = 16;
struct MgState {
    double rz[2];
    double bnorm, res, rnorm0, inner_tol, outer_tol;
    int solve_done, inner_done, outer, inner_it;
    int total_iters, max_outer, max_inner, passes;
    int total_cap, pad;
    int hist[kMgHist];
    // the multigrid tail's group barrier (TailLevels::gsync) and its fault
    // flag (a barrier that timed out); zero between launches
    unsigned gsync[2];
    int tail_fault, pad2;
};

}  // namespace ffr
