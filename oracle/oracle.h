/* oracle.h - FP64 CPU restatement of ff-room's projection time step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 * The product path (the HIP library and the host classes above it) never
 * links, loads or calls anything here.
 *
 * Every function restates the reference algorithm in the same floating-point
 * operation order (compiled with -ffp-contract=off) so that, run serially, it
 * reproduces the reference's bits. Citations are to ddd3h/ff-room:
 *   cpp/src/core/FluidSolver.cpp, PoissonSolver.cpp, BoundaryManager.cpp.
 *
 * Layouts are the reference's (cpp/src/core/Grid.hpp:57-65), k fastest:
 *   u (nx+1,ny,nz)  v (nx,ny+1,nz)  w (nx,ny,nz+1)  p,T,flags (nx,ny,nz).
 */
#ifndef FFROOM_ORACLE_H
#define FFROOM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_FLUID = 0, ORC_SOLID = 1, ORC_INFLOW = 2, ORC_OUTFLOW = 3 };
enum { ORC_UPWIND1 = 0, ORC_QUICK = 1, ORC_LAX_WENDROFF = 2 };

typedef struct { /* FanBC, BoundaryManager.hpp:8-13 */
    int i_min, i_max, j_min, j_max, k_min, k_max;
    double vel[3];
} orc_fan;

typedef struct { /* OutflowBC, BoundaryManager.hpp:17-22 */
    int axis, side, a_min, a_max, b_min, b_max;
} orc_outflow;

typedef struct { /* OpeningBC, BoundaryManager.hpp:27-29 */
    int axis, side, a_min, a_max, b_min, b_max;
    double T_outside;
} orc_opening;

typedef struct {
    int nx, ny, nz;
    const double *dxv, *dyv, *dzv;  /* sizes nx, ny, nz */
    const uint8_t *flags;           /* nx*ny*nz CellType */
    double *u, *v, *w, *p, *T;      /* state, reference layout */
    const orc_fan *fans;         int n_fans;
    const orc_outflow *outflows; int n_outflows;
    const orc_opening *openings; int n_openings;
} orc_state;

typedef struct { /* FluidSolverParams, FluidSolver.hpp:11-38 */
    double rho, nu, dt;
    int    max_steps;
    double convergence_tol;
    int    poisson_max_iter;
    double poisson_tol;
    int    advection;
    int    use_multigrid;
    int    thermal;
    double T_initial, T_target;
    int    buoyancy;
    double g_accel, beta, k_thermal, cp;
    int    nthreads; /* >1: OpenMP over the loops the reference threads */
} orc_params;

typedef struct { /* StepResult, FluidSolver.hpp:40-47 (+ solver iterations) */
    int    step;
    double velocity_change, pressure_residual, divergence_max;
    int    converged;
    double T_mean;
    int    poisson_iters;
} orc_result;

void   orc_default_params(orc_params *p);
/* FluidSolver ctor (FluidSolver.cpp:18-32): T := T_initial, opening T. */
void   orc_init(orc_state *s, const orc_params *p);
/* FluidSolver::step (FluidSolver.cpp:486-550). *step_counter is step_. */
void   orc_step(orc_state *s, const orc_params *p, int *step_counter, orc_result *out);

/* Individual phases, each in place on s. */
void   orc_predictor(orc_state *s, const orc_params *p);          /* :83-309 */
void   orc_buoyancy(orc_state *s, const orc_params *p);           /* :373-386 */
void   orc_apply_noslip(orc_state *s);                            /* BM :60-109 */
void   orc_apply_inflow(orc_state *s);                            /* BM :111-130 */
void   orc_apply_outflow(orc_state *s);                           /* BM :180-220 */
void   orc_apply_opening_temperature(orc_state *s);               /* BM :137-178 */
void   orc_poisson_rhs(const orc_state *s, const orc_params *p, double *rhs); /* :311-330 */
void   orc_correction(orc_state *s, const orc_params *p);         /* :332-371 */
void   orc_advect_temperature(orc_state *s, const orc_params *p); /* :388-456 */
double orc_T_mean(const orc_state *s, const orc_params *p);       /* :458-468 */
double orc_divergence_max(const orc_state *s);                    /* :470-484 */

/* PoissonSolver.cpp:22-91, A*x over the whole grid. */
void   orc_apply_laplacian(const orc_state *s, const double *x, double *Ax, int nthreads);
/* PoissonSolver::solve (CG), PoissonSolver.cpp:93-158; writes s->p. */
double orc_poisson_cg(orc_state *s, const double *rhs, int max_iter, double tol,
                      int nthreads, int *iters);
/* MultigridPoissonSolver::solve, PoissonSolver.cpp:446-505; writes s->p. */
double orc_poisson_mg(orc_state *s, const double *rhs, int max_iter, double tol, int *iters);

/* ---- D3Q19 lattice Boltzmann (LBMSolver.cpp/.hpp) ------------------------
 * BGK collision with one relaxation time, inflow cells held at equilibrium
 * with their fan's velocity, streaming with bounce-back at the domain walls
 * and SOLID cells, zero-gradient copies on outflow and opening patches, and
 * macroscopic fields written to the MAC grid. Populations f[q*nc + cell]. */
typedef struct {
    double tau, dt, dx_ref;     /* relaxation time, physical step, Grid::dx = Lx/Nx */
    double convergence_tol;
    double *f, *f2;             /* 19*nc each (caller-owned) */
    double *ux, *uy, *uz;       /* nc each: cell velocities [m/s] (caller-owned) */
} orc_lbm;
void orc_lbm_init(orc_state *s, orc_lbm *L);                       /* :39-75 */
void orc_lbm_step(orc_state *s, orc_lbm *L, int *step_counter, orc_result *out); /* :403-444 */

#ifdef __cplusplus
}
#endif
#endif
