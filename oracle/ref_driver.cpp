// ref_driver.cpp - C shim over the reference's own C++ solver classes.
//
// TEST INFRASTRUCTURE ONLY. oracle/Makefile compiles this file together
// with the reference's Grid.cpp, BoundaryManager.cpp, PoissonSolver.cpp and
// FluidSolver.cpp (and LBMSolver.cpp), read in place from the reference tree, into
// oracle/_ref/libffref.so. Tests use it to pin the C restatement (oracle.c)
// and to write golden fixtures; the benchmark may time it as the CPU
// baseline. Nothing in the product links it.
#include "core/Grid.hpp"
#include "core/BoundaryManager.hpp"
#include "core/PoissonSolver.hpp"
#include "core/FluidSolver.hpp"
#include "core/LBMSolver.hpp"

#include <cstring>
#include <memory>
#include <vector>

using namespace ffroom;

namespace {
struct RefHandle {
    std::unique_ptr<Grid> grid;
    BoundaryManager bm;
    std::unique_ptr<FluidSolver> solver;
    std::unique_ptr<LBMSolver> lbm;
};

void copy_out(const std::vector<double> &src, double *dst)
{
    if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(double));
}
void copy_in(std::vector<double> &dst, const double *src)
{
    if (src) std::memcpy(dst.data(), src, dst.size() * sizeof(double));
}
}  // namespace

extern "C" {

typedef struct {
    double rho, nu, dt;
    int max_steps;
    double convergence_tol;
    int poisson_max_iter;
    double poisson_tol;
    int advection, use_multigrid, thermal;
    double T_initial, T_target;
    int buoyancy;
    double g_accel, beta, k_thermal, cp;
    int use_openmp;
} ref_params;

typedef struct {
    int step;
    double velocity_change, pressure_residual, divergence_max;
    int converged;
    double T_mean;
} ref_result;

void *ref_create(int nx, int ny, int nz, double lx, double ly, double lz)
{
    auto *h = new RefHandle;
    h->grid.reset(new Grid(nx, ny, nz, lx, ly, lz));
    return h;
}

void ref_destroy(void *hp) { delete static_cast<RefHandle *>(hp); }

void ref_add_solid_box(void *hp, int i0, int i1, int j0, int j1, int k0, int k1)
{
    auto *h = static_cast<RefHandle *>(hp);
    h->bm.add_solid_box(*h->grid, i0, i1, j0, j1, k0, k1);
}

void ref_add_fan(void *hp, int i0, int i1, int j0, int j1, int k0, int k1,
                 double vx, double vy, double vz)
{
    auto *h = static_cast<RefHandle *>(hp);
    FanBC f;
    f.i_min = i0; f.i_max = i1; f.j_min = j0; f.j_max = j1; f.k_min = k0; f.k_max = k1;
    f.vel = {vx, vy, vz};
    h->bm.add_fan(*h->grid, f);
}

void ref_add_outflow(void *hp, int axis, int side, int a0, int a1, int b0, int b1)
{
    auto *h = static_cast<RefHandle *>(hp);
    OutflowBC o;
    o.axis = axis; o.side = side; o.a_min = a0; o.a_max = a1; o.b_min = b0; o.b_max = b1;
    h->bm.add_outflow(*h->grid, o);
}

void ref_add_opening(void *hp, int axis, int side, int a0, int a1, int b0, int b1, double T_out)
{
    auto *h = static_cast<RefHandle *>(hp);
    OpeningBC o;
    o.axis = axis; o.side = side; o.a_min = a0; o.a_max = a1; o.b_min = b0; o.b_max = b1;
    o.T_outside = T_out;
    h->bm.add_opening(*h->grid, o);
}

void ref_get_flags(void *hp, unsigned char *out)
{
    auto *h = static_cast<RefHandle *>(hp);
    for (size_t i = 0; i < h->grid->cell_type.size(); i++)
        out[i] = static_cast<unsigned char>(h->grid->cell_type[i]);
}

void ref_set_flags(void *hp, const unsigned char *in)
{
    auto *h = static_cast<RefHandle *>(hp);
    for (size_t i = 0; i < h->grid->cell_type.size(); i++)
        h->grid->cell_type[i] = static_cast<CellType>(in[i]);
}

// Grid::set_face_coords (Grid.cpp:39-47): xs, ys, zs of sizes nx+1, ny+1, nz+1
void ref_set_face_coords(void *hp, const double *xs, const double *ys, const double *zs)
{
    auto *h = static_cast<RefHandle *>(hp);
    const Grid &g = *h->grid;
    h->grid->set_face_coords(std::vector<double>(xs, xs + g.Nx + 1),
                             std::vector<double>(ys, ys + g.Ny + 1),
                             std::vector<double>(zs, zs + g.Nz + 1));
}

void ref_make_solver(void *hp, const ref_params *p)
{
    auto *h = static_cast<RefHandle *>(hp);
    FluidSolverParams fp;
    fp.rho = p->rho; fp.nu = p->nu; fp.dt = p->dt; fp.max_steps = p->max_steps;
    fp.convergence_tol = p->convergence_tol;
    fp.poisson.max_iter = p->poisson_max_iter;
    fp.poisson.tol = p->poisson_tol;
    fp.advection = static_cast<AdvectionScheme>(p->advection);
    fp.use_multigrid = p->use_multigrid != 0;
    fp.thermal = p->thermal != 0;
    fp.T_initial = p->T_initial; fp.T_target = p->T_target;
    fp.buoyancy = p->buoyancy != 0;
    fp.g_accel = p->g_accel; fp.beta = p->beta; fp.k_thermal = p->k_thermal; fp.cp = p->cp;
    fp.use_openmp = p->use_openmp != 0;
    h->solver.reset(new FluidSolver(*h->grid, h->bm, fp));
}

void ref_step(void *hp, ref_result *out)
{
    auto *h = static_cast<RefHandle *>(hp);
    StepResult r = h->solver->step();
    out->step = r.step;
    out->velocity_change = r.velocity_change;
    out->pressure_residual = r.pressure_residual;
    out->divergence_max = r.divergence_max;
    out->converged = r.converged ? 1 : 0;
    out->T_mean = r.T_mean;
}

// FluidSolver::run (FluidSolver.cpp:552-560); returns the last StepResult.
void ref_run(void *hp, ref_result *out)
{
    auto *h = static_cast<RefHandle *>(hp);
    StepResult r = h->solver->run();
    out->step = r.step;
    out->velocity_change = r.velocity_change;
    out->pressure_residual = r.pressure_residual;
    out->divergence_max = r.divergence_max;
    out->converged = r.converged ? 1 : 0;
    out->T_mean = r.T_mean;
}

void ref_get_fields(void *hp, double *u, double *v, double *w, double *p, double *T)
{
    auto *h = static_cast<RefHandle *>(hp);
    copy_out(h->grid->u, u); copy_out(h->grid->v, v); copy_out(h->grid->w, w);
    copy_out(h->grid->p, p); copy_out(h->grid->T, T);
}

void ref_set_fields(void *hp, const double *u, const double *v, const double *w,
                    const double *p, const double *T)
{
    auto *h = static_cast<RefHandle *>(hp);
    copy_in(h->grid->u, u); copy_in(h->grid->v, v); copy_in(h->grid->w, w);
    copy_in(h->grid->p, p); copy_in(h->grid->T, T);
}

double ref_divergence_max(void *hp) { return static_cast<RefHandle *>(hp)->solver->compute_divergence_max(); }

// PoissonSolver / MultigridPoissonSolver::solve on the handle's grid.
double ref_poisson_solve(void *hp, const double *rhs, int use_mg, int max_iter, double tol,
                         int *iters)
{
    auto *h = static_cast<RefHandle *>(hp);
    PoissonSolverParams pp;
    pp.max_iter = max_iter;
    pp.tol = tol;
    std::vector<double> b(rhs, rhs + h->grid->p.size());
    double res;
    if (use_mg) {
        MultigridPoissonSolver s(pp);
        res = s.solve(*h->grid, b);
        if (iters) *iters = s.last_iterations();
    } else {
        PoissonSolver s(pp);
        res = s.solve(*h->grid, b);
        if (iters) *iters = s.last_iterations();
    }
    return res;
}

// LBMSolver (LBMSolver.hpp:24-39) on the handle's grid and boundary lists.
void ref_make_lbm(void *hp, double tau, int max_steps, double convergence_tol, double dt)
{
    auto *h = static_cast<RefHandle *>(hp);
    h->lbm.reset(new LBMSolver(*h->grid, h->bm, tau, max_steps, convergence_tol, dt));
}

void ref_lbm_step(void *hp, ref_result *out)
{
    auto *h = static_cast<RefHandle *>(hp);
    StepResult r = h->lbm->step();
    out->step = r.step;
    out->velocity_change = r.velocity_change;
    out->pressure_residual = r.pressure_residual;
    out->divergence_max = r.divergence_max;
    out->converged = r.converged ? 1 : 0;
    out->T_mean = r.T_mean;
}

void ref_apply_noslip(void *hp) { auto *h = static_cast<RefHandle *>(hp); h->bm.apply_noslip(*h->grid); }
void ref_apply_inflow(void *hp) { auto *h = static_cast<RefHandle *>(hp); h->bm.apply_inflow(*h->grid); }
void ref_apply_outflow(void *hp) { auto *h = static_cast<RefHandle *>(hp); h->bm.apply_outflow(*h->grid); }
void ref_apply_opening_temperature(void *hp)
{
    auto *h = static_cast<RefHandle *>(hp);
    h->bm.apply_opening_temperature(*h->grid);
}

}  // extern "C"
