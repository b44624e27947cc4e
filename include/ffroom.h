/* ffroom.h - C ABI of the MI355X (gfx950) projection-step library.
 *
 * The host C++ classes (Grid, BoundaryManager, PoissonSolver,
 * MultigridPoissonSolver, FluidSolver) and the pybind11 module
 * `_ffroom_core` are thin shims over these calls; a foreign host (ctypes,
 * cgo, JNI) can bind them directly. Plain C types only: no C++ objects, no
 * torch tensors and no exceptions cross this boundary. Every function
 * returns FFR_OK (0) or a negative ffr_status; ffr_last_error() gives the
 * message of the last failure on a context.
 *
 * Host buffers use the reference layout (ddd3h/ff-room cpp/src/core/Grid.hpp:
 * 57-65, k fastest): u (nx+1,ny,nz), v (nx,ny+1,nz), w (nx,ny,nz+1),
 * p/T/cell_type (nx,ny,nz); fields are double, cell_type is uint8.
 *
 * Threading: a context and its scenes are used from one host thread at a time.
 */
#ifndef FFROOM_H
#define FFROOM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFR_ABI_VERSION 1

typedef struct ffr_ctx ffr_ctx;
typedef struct ffr_scene ffr_scene;

typedef enum {
    FFR_OK = 0,
    FFR_E_INVALID = -1,     /* bad argument or descriptor */
    FFR_E_DEVICE = -2,      /* HIP runtime error */
    FFR_E_NOMEM = -3,       /* device allocation or HBM budget exceeded */
    FFR_E_STATE = -4,       /* call not valid in the object's state */
    FFR_E_NUMERIC = -5,     /* non-finite values detected */
    FFR_E_UNSUPPORTED = -6, /* option not available for this scene */
    FFR_E_NOT_FOUND = -7    /* unknown ticket */
} ffr_status;

/* CellType, Grid.hpp:9-14 */
enum { FFR_FLUID = 0, FFR_SOLID = 1, FFR_INFLOW = 2, FFR_OUTFLOW = 3 };
/* AdvectionScheme, FluidSolver.hpp:9 */
enum { FFR_UPWIND1 = 0, FFR_QUICK = 1, FFR_LAX_WENDROFF = 2 };
/* Arithmetic of the state and stencils. */
enum {
    FFR_PREC_FP64 = 0,  /* every field and kernel in FP64: the reference algorithm */
    FFR_PREC_MIXED = 1  /* FP32 fields/stencils, FP32+BF16 multigrid, FP64 outer residual */
};
/* Pressure solver. */
enum {
    FFR_POISSON_AUTO = 0,  /* reference choice: CG, or MG-PCG when use_multigrid */
    FFR_POISSON_CG = 1,    /* unpreconditioned CG, PoissonSolver.cpp:93-158 */
    FFR_POISSON_MGPCG = 2  /* geometric-multigrid-preconditioned CG: an FP64 outer
                              residual of the reference operator around inner
                              solves; on a stretched grid (varying dxv/dyv/dzv)
                              the inner solves run on the operator scaled row by
                              row by the relative cell volume, which is symmetric */
};

/* FanBC, BoundaryManager.hpp:8-13 */
typedef struct {
    int i_min, i_max, j_min, j_max, k_min, k_max;
    double vel[3];
} ffr_fan;

/* OutflowBC, BoundaryManager.hpp:17-22 */
typedef struct {
    int axis, side, a_min, a_max, b_min, b_max;
} ffr_outflow;

/* OpeningBC, BoundaryManager.hpp:27-29 */
typedef struct {
    int axis, side, a_min, a_max, b_min, b_max;
    double T_outside;
} ffr_opening;

/* A box of cells given one type (FFR_SOLID, FFR_INFLOW, FFR_OUTFLOW or
 * FFR_FLUID), inclusive ranges; parts outside the grid are ignored, as the
 * BoundaryManager's markers clamp them (BoundaryManager.cpp:6-44). */
typedef struct {
    int type;
    int i_min, i_max, j_min, j_max, k_min, k_max;
} ffr_mark;
#define FFR_MAX_MARKS 64

/* Geometry and boundary lists of one scene (what Grid + BoundaryManager
 * hold). `outflows` lists every registered outflow patch, openings included,
 * as BoundaryManager::outflows() does (BoundaryManager.cpp:24-44, 132-135).
 * The cell flags come either from `cell_type` or, when it is NULL, from
 * `marks` applied in order on the device (FLUID where no mark applies; the
 * last mark holding a cell wins - the marking order of scene.py:89-106). */
typedef struct {
    int nx, ny, nz;
    const double *dxv, *dyv, *dzv;  /* cell widths, sizes nx, ny, nz */
    const uint8_t *cell_type;       /* nx*ny*nz, or NULL to use marks */
    const ffr_fan *fans;
    int n_fans;
    const ffr_outflow *outflows;
    int n_outflows;
    const ffr_opening *openings;
    int n_openings;
    const ffr_mark *marks;          /* used when cell_type is NULL */
    int n_marks;                    /* 0..FFR_MAX_MARKS */
} ffr_scene_desc;

/* FluidSolverParams (FluidSolver.hpp:11-38) + PoissonSolverParams
 * (PoissonSolver.hpp:14-18), then optional extensions. ffr_params_default()
 * gives the reference defaults and the reference algorithm. */
typedef struct {
    double rho, nu, dt;
    int max_steps;
    double convergence_tol;
    int poisson_max_iter;
    double poisson_tol;
    int advection;
    int use_openmp;    /* accepted for compatibility; no effect on the GPU */
    int use_multigrid; /* pressure at OUTFLOW cells pinned to 0 (PoissonSolver.cpp:201) */
    int thermal;
    double T_initial, T_target;
    int buoyancy;
    double g_accel, beta, k_thermal, cp;
    /* extensions */
    int precision;       /* FFR_PREC_* */
    int poisson_method;  /* FFR_POISSON_* */
    int mg_levels;       /* multigrid levels incl. the finest; 0 = automatic (<= 5) */
    int mg_bf16;         /* store multigrid levels >= 1 in bfloat16 */
    int mg_pre_sweeps;   /* red-black Gauss-Seidel sweeps per side of the V-cycle, level 0 */
    int mg_level_sweeps; /* the same on levels 1 .. coarsest-1 */
    int mg_coarse_sweeps;/* symmetric sweeps on the coarsest level; 0: solve it with a dense
                            inverse when it has at most 512 cells and runs in the
                            one-workgroup tail (8 sweeps otherwise) */
    double inner_tol;    /* mixed: relative residual target of each inner FP32 solve */
    int inner_max_iter;  /* mixed: iteration cap of each inner solve */
    int graphs;          /* replay the inner iteration from a captured HIP graph (1) or
                            launch its kernels one by one (0); results are identical */
    int mg_tail_kb;      /* LDS budget (KiB, <= 150) of the single-workgroup kernel that runs
                            the smallest multigrid levels; 0 = 96 */
    int warm_start;      /* mixed MG-PCG: a step's pressure solve starts from the previous
                            step's pressure (1), from its linear extrapolation in time
                            2 p_n - p_{n-1} (2) or quadratic one 3 p_n - 3 p_{n-1} + p_{n-2}
                            (3, the default) (the first steps after creation or new params
                            use as many past pressures as there are; an uploaded p takes the
                            place of p_n), or from 0 as the reference CG does (0);
                            ffr_poisson_solve always starts from 0 */
    int mg_tail_cells;   /* levels >= 1 of at most this many cells also run inside the
                            single-workgroup kernel, with their vectors and operator in
                            global memory: one launch per V-cycle for all of them instead of
                            several per level (fewer dispatches for small scenes served side
                            by side); 0 = only the levels that fit its LDS budget */
    int mg_tail_groups;  /* workgroups (1..8) that share those global levels, all on one XCD
                            and ordered by a barrier in its L2 (the LDS levels run on one of
                            them); 0 = one per 16384 cells of the largest global level.
                            Results are the same bits for every count */
} ffr_params;

/* StepResult, FluidSolver.hpp:40-47, plus solver statistics. */
typedef struct {
    int step;
    double velocity_change;   /* max |u^{n+1}-u^n| / dt over all faces */
    double pressure_residual; /* relative residual of the pressure solve */
    double divergence_max;    /* max |div u| over FLUID cells */
    int converged;
    double T_mean;
    int poisson_iters;        /* CG / inner PCG iterations of this step */
    int outer_iters;          /* FP64 defect-correction passes (mixed) */
    int nonfinite;            /* 1 when a NaN/Inf appeared in u, v, w or T */
} ffr_step_result;

typedef struct {
    double residual; /* relative residual ||b-Ax|| / ||b|| (SOLID rows excluded) */
    int iterations;
    int outer_iterations;
} ffr_solve_info;

typedef struct {
    char name[64];
    char arch[32];
    int compute_units;
    int xcds;
    size_t hbm_bytes;
    size_t l2_bytes;
    size_t lds_bytes_per_cu;
    int clock_khz;
} ffr_device_info;

/* Per-kernel timing gathered with HIP events on the launching stream. */
typedef struct {
    char name[48];
    long launches;
    double total_ms;
    double bytes_per_launch; /* mean algorithmic bytes per launch */
} ffr_prof_entry;

typedef void (*ffr_step_callback)(const ffr_step_result *, void *user);

/* ---- library / context ------------------------------------------------ */
int ffr_abi_version(void);
const char *ffr_status_string(int status);
void ffr_params_default(ffr_params *p);
int ffr_ctx_create(int device, size_t hbm_budget_bytes, ffr_ctx **out);
void ffr_ctx_destroy(ffr_ctx *ctx);
const char *ffr_last_error(const ffr_ctx *ctx);
int ffr_device_query(ffr_ctx *ctx, ffr_device_info *out);
size_t ffr_ctx_bytes_in_use(const ffr_ctx *ctx);
/* Device blocks of destroyed scenes stay cached in their context for the
 * scenes created next (hipFree would wait for the whole device, stalling the
 * other scenes' streams in a serving loop); they are freed, largest first,
 * when a new block would not fit the budget beside them, and with the
 * context. ffr_ctx_bytes_cached is the device bytes cached; ffr_ctx_trim
 * frees every cached block, pinned block, stream and event and returns the
 * device bytes it freed. No counterpart in the reference, whose scenes live
 * in host memory. */
size_t ffr_ctx_bytes_cached(const ffr_ctx *ctx);
size_t ffr_ctx_trim(ffr_ctx *ctx);

/* ---- scenes ------------------------------------------------------------ */
/* The multigrid hierarchy of a scene that uses one (levels = 0 otherwise):
 * the level count, the first level the single-workgroup kernel runs (-1:
 * none), the first of those with its vectors in LDS (mg_tail_cells puts the
 * ones before it in global memory), and the cells of each level (up to 12
 * entries; NULL to skip). No counterpart in the reference (its multigrid,
 * PoissonSolver.cpp:164-505, has a fixed level list): for tests and tuning. */
int ffr_scene_mg_layout(const ffr_scene *scene, int *levels, int *tail_first, int *tail_lds_from,
                        unsigned int *cells);
/* Workgroups that run the single-workgroup kernel's global levels together
 * (1 when it has none, or one workgroup; 0 without a hierarchy); negative on
 * error. No counterpart in the reference: for tests and tuning. */
int ffr_scene_mg_tail_groups(const ffr_scene *scene);
/* Device bytes a scene with this geometry and these parameters occupies. */
size_t ffr_scene_footprint(const ffr_scene_desc *desc, const ffr_params *params);
/* Copies the geometry to the device and allocates the state (zero fields). */
int ffr_scene_create(ffr_ctx *ctx, const ffr_scene_desc *desc, const ffr_params *params,
                     ffr_scene **out);
void ffr_scene_destroy(ffr_scene *scene);
int ffr_scene_set_params(ffr_scene *scene, const ffr_params *params);
/* The scene's cell flags (nx*ny*nz bytes), however they were given. */
int ffr_scene_cell_type(const ffr_scene *scene, uint8_t *cell_type);
/* Snapshot egress (the captures of SolverBridge.run_animated,
 * solver_bridge.py:172-281, and the fields of SimResult): ffr_snapshot_begin
 * queues a copy of the scene's state as it is after its last step and
 * returns a slot without waiting for it to reach the host; the next steps
 * may be issued at once, and the copy to the host overlaps them.
 *   FFR_SNAP_FIELDS: u, v, w, p, T in the reference layout;
 *   FFR_SNAP_CELLS:  a FieldSnapshot's arrays, computed on the device in FP64
 *                    as the reference computes them on the host
 *                    (solver_bridge.py:224-236): cell-centred velocity
 *                    (nx, ny, nz, 3 interleaved), speed, T.
 * At most FFR_SNAPSHOT_SLOTS snapshots of a scene are pending (FFR_E_STATE
 * beyond); each slot holds device and pinned memory, the device part counted
 * against the context budget, until the scene is destroyed. ffr_snapshot_end
 * waits for slot `slot`, writes its arrays as FP64 - (u, v, w, p, T) for
 * fields, (vel, speed, T, -, -) for cells; NULL skips one - and the step
 * count at capture into *step, and frees the slot. */
#define FFR_SNAPSHOT_SLOTS 4
enum { FFR_SNAP_FIELDS = 0, FFR_SNAP_CELLS = 1 };
int ffr_snapshot_begin(ffr_scene *scene, int kind, int *slot);
int ffr_snapshot_end(ffr_scene *scene, int slot, int *step, double *a, double *b, double *c,
                     double *d, double *e);
/* Host <-> device state copies; a NULL pointer leaves that field alone. */
int ffr_scene_upload(ffr_scene *scene, const double *u, const double *v, const double *w,
                     const double *p, const double *T);
int ffr_scene_download(ffr_scene *scene, double *u, double *v, double *w, double *p,
                       double *T);
int ffr_scene_step_count(const ffr_scene *scene);
int ffr_scene_set_step_count(ffr_scene *scene, int step);

/* ---- the projection step (FluidSolver.cpp:486-560) --------------------- */
int ffr_step(ffr_scene *scene, ffr_step_result *out);
int ffr_run(ffr_scene *scene, int max_steps, ffr_step_callback cb, void *user,
            ffr_step_result *last);
int ffr_divergence_max(ffr_scene *scene, double *out); /* FluidSolver.cpp:470-484 */
int ffr_T_mean(ffr_scene *scene, double *out);         /* FluidSolver.cpp:458-468 */

/* ---- pressure operator and solve (PoissonSolver.cpp:22-158) ------------ */
/* Ax = A x for the scene's flags and spacings (host buffers, nx*ny*nz). */
int ffr_apply_laplacian(ffr_scene *scene, const double *x, double *Ax);
/* Solves A p = rhs with the scene's poisson settings; p is written back. */
int ffr_poisson_solve(ffr_scene *scene, const double *rhs, double *p, ffr_solve_info *info);

/* ---- single phases of the step, on the scene's current state ------------
 * Each takes the scene's current u, v, w as the predictor's output u*, v*, w*
 * and runs the step's own kernel for one phase (between steps only):
 * ffr_build_rhs: rhs = -(rho/dt) div(u*) on non-SOLID cells, 0 on SOLID
 *   (FluidSolver::build_poisson_rhs, FluidSolver.cpp:311-330), nx*ny*nz.
 * ffr_apply_correction: p (host, nx*ny*nz) becomes the scene's pressure, and
 *   u, v, w := u* - (dt/rho) grad p on interior faces, then the no-slip, fan
 *   and outflow masks (FluidSolver::apply_correction, FluidSolver.cpp:332-371,
 *   then BoundaryManager::apply_noslip / apply_inflow / apply_outflow as
 *   FluidSolver::step applies them after the solve, FluidSolver.cpp:511-515);
 *   *velocity_change (may be NULL) = max |u - u*| over all faces. */
int ffr_build_rhs(ffr_scene *scene, double *rhs);
int ffr_apply_correction(ffr_scene *scene, const double *p, double *velocity_change);

/* ---- a pressure solve that allocates nothing -----------------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same solve and result as ffr_poisson_solve on a
 * scene whose pressure solver is the CG form (FFR_POISSON_CG, or AUTO without
 * use_multigrid; FFR_E_UNSUPPORTED otherwise), with the staged right-hand
 * side, the CG vectors, the reduction partials and the iteration scalars in
 * a fixed device array of the library (one per GPU, about 32 MiB, held until
 * the process ends) instead of the scene's memory. Scenes of more than
 * FFR_NOALLOC_MAX_CELLS cells are rejected with FFR_E_INVALID. Calls on one
 * GPU share its array: the library serialises them within the process, and
 * calls that overlapped on one GPU would race. rhs and p are host buffers
 * (nx*ny*nz); p may be NULL; the solution is also the scene's pressure. */
#define FFR_NOALLOC_MAX_CELLS (1u << 20)
int ffr_poisson_solve_noalloc(ffr_scene *scene, const double *rhs, double *p,
                              ffr_solve_info *info);

/* ---- lattice Boltzmann (LBMSolver, LBMSolver.hpp:17-39) ------------------
 * The reference's second solver method: D3Q19, BGK with relaxation time tau,
 * bounce-back at walls and SOLID cells with the reference's direction
 * pairing, fans as equilibrium inflow, zero-gradient copies on outflow and
 * opening patches. ffr_lbm_attach gives an FP64 scene a lattice at rest
 * (LBMSolver's constructor, LBMSolver.cpp:39-75; dx_ref is Grid::dx = Lx/Nx,
 * the unit between lattice and physical velocity); ffr_lbm_step advances it
 * one step (LBMSolver::step, :403-444) and leaves u, v, w (faces) and p =
 * rho/3 (cells that are not SOLID) in the scene, results equal to the
 * reference's bits. ffr_step is the projection step and does not touch the
 * lattice. */
int ffr_lbm_attach(ffr_scene *scene, double tau, double dt, double convergence_tol,
                   double dx_ref);
int ffr_lbm_step(ffr_scene *scene, ffr_step_result *out);

/* ---- batching and serving ---------------------------------------------- */
/* Continuous batching of many scenes on the context's device (no counterpart
 * in the reference, which runs one scene per process). A ticket asks for up
 * to `max_steps` steps of a scene, ending early when a step reports
 * `converged` (FluidSolver::run, FluidSolver.cpp:552-560). At most `slots`
 * tickets (default 16) are active; later ones wait in submission order and
 * are admitted as active ones finish. Each scene runs on its own stream and
 * every ffr_poll / ffr_drain pass advances all active scenes without
 * blocking, so their kernels overlap on the device. A scene's results are the
 * same alone and in a batch (every reduction has a fixed order). HBM
 * admission: ffr_scene_create returns FFR_E_NOMEM beyond the context budget;
 * destroy finished scenes to admit new ones. */
int ffr_serve_config(ffr_ctx *ctx, int slots);
/* Host threads (no counterpart in the reference; 1..64, default 8) that run
 * the active tickets: a scene's step is a few hundred kernel launches and
 * host checks, and issuing them for many scenes from one thread limits the
 * served steps per second. With more than one thread each active ticket is
 * run by one worker thread at a time (waiting on the device between its
 * launches), and ffr_progress / ffr_poll / ffr_drain admit, collect and
 * report; with 1, those calls advance every ticket themselves. Results do
 * not depend on the count. Gangs (below) are advanced by the calling
 * thread. */
int ffr_serve_threads(ffr_ctx *ctx, int threads);
/* Gangs (no counterpart in the reference): with gangs >= 1, the active
 * mixed-precision MG-PCG scenes are spread over `gangs` groups, and each
 * group advances its scenes in rounds of merged launches: one launch runs the
 * same kernel of every member whose step is at that kernel, each member with
 * its own grid and arguments, so a round costs the launches of its longest
 * member instead of the sum. Results are bit-identical to gangs = 0 (every
 * scene on its own stream, the default); other scenes (FP64 form, plain CG)
 * keep their own streams. Set it before submitting. */
int ffr_serve_gangs(ffr_ctx *ctx, int gangs);
/* Totals over the context's gangs since ffr_serve_gangs. */
typedef struct ffr_gang_stats {
    uint64_t rounds;      /* rounds issued */
    uint64_t launches;    /* merged launches issued */
    uint64_t ops;         /* operations the members recorded (launches and scalar
                             copies): what their own streams would have issued */
    uint64_t longest_ops; /* the longest member's operations, summed over rounds */
    double host_us;       /* host time recording, merging and issuing rounds */
} ffr_gang_stats;
int ffr_serve_gang_stats(ffr_ctx *ctx, ffr_gang_stats *out);
int ffr_submit(ffr_ctx *ctx, ffr_scene *scene, int max_steps, uint64_t *ticket);
/* Advances all active tickets once; *done is set when this ticket finished
 * and *last receives its latest step result. Returns the ticket's error code
 * once it has finished with one. */
int ffr_poll(ffr_ctx *ctx, uint64_t ticket, ffr_step_result *last, int *done);
/* Advances every active ticket once (non-blocking); *steps_completed = the
 * number of steps that finished in this pass. */
int ffr_progress(ffr_ctx *ctx, int *steps_completed);
/* Runs every submitted ticket to completion. */
int ffr_drain(ffr_ctx *ctx);
/* Steps completed so far, without advancing anything. */
int ffr_ticket_info(ffr_ctx *ctx, uint64_t ticket, int *steps, int *done);
/* Forgets a finished ticket. */
int ffr_release(ffr_ctx *ctx, uint64_t ticket);

/* ---- launch tuning (no counterpart in the reference) ----------------------- */
/* Upper bound on the blocks of one grid-strided launch (process-wide; 256 to
 * 16384, default 2048). Results do not depend on it: every reduction keeps a
 * fixed order for a given value, and the value is part of the tuning. */
int ffr_set_max_blocks(int max_blocks);
int ffr_get_max_blocks(void);

/* ---- profiling ---------------------------------------------------------- */
int ffr_prof_enable(ffr_ctx *ctx, int on);
int ffr_prof_reset(ffr_ctx *ctx);
int ffr_prof_read(ffr_ctx *ctx, ffr_prof_entry *out, int cap, int *n_out);
/* Times `reps` launches of one kernel of the step on the scene's current
 * buffers, each between the begin and end timestamps of its own dispatch
 * (the figure a kernel trace reports): the mean kernel duration, and the
 * kernel's algorithmic bytes per launch. Kernels: "predict", "correct",
 * "temperature" (thermal scenes), "rhs",
 * "cg_spmv", "cg_update", "cg_dir", "mg_smooth_l0", "mg_resid0",
 * "pcg_dirq", "pcg_update", "dc_residual", the lattice's "lbm_relax",
 * "lbm_stream", "lbm_faces" (after ffr_lbm_attach; timing
 * advances the lattice), and "copy" (a plain FP64 copy
 * of nx*ny*nz doubles, to calibrate memory counters). The timed kernels overwrite
 * solver scratch; "correct" rewrites u, v, w with the values the last step gave
 * them: call between steps, not inside one. */
int ffr_prof_kernel(ffr_scene *scene, const char *name, int reps, double *mean_ms,
                    double *bytes_per_launch);

#ifdef __cplusplus
}
#endif
#endif
