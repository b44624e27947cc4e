/* oracle.c - FP64 CPU restatement of ff-room's projection time step.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the CPU-baseline port);
 * nothing in the product links or loads it. Built with -ffp-contract=off so
 * that each expression rounds exactly as the reference's does when that is
 * also built without FMA contraction.
 *
 * Citations: ddd3h/ff-room cpp/src/core/{FluidSolver,PoissonSolver,
 * BoundaryManager}.cpp, line numbers given per function.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Flat indices of Grid.hpp:59-65 (k fastest). Negative j or k wrap into the
 * previous row exactly like the reference's int arithmetic does. */
#define IU(s, i, j, k) ((long)(i) * ((s)->ny * (s)->nz) + (long)(j) * (s)->nz + (k))
#define IV(s, i, j, k) ((long)(i) * (((s)->ny + 1) * (s)->nz) + (long)(j) * (s)->nz + (k))
#define IW(s, i, j, k) ((long)(i) * ((s)->ny * ((s)->nz + 1)) + (long)(j) * ((s)->nz + 1) + (k))
#define IC(s, i, j, k) IU(s, i, j, k)

static long n_u(const orc_state *s) { return (long)(s->nx + 1) * s->ny * s->nz; }
static long n_v(const orc_state *s) { return (long)s->nx * (s->ny + 1) * s->nz; }
static long n_w(const orc_state *s) { return (long)s->nx * s->ny * (s->nz + 1); }
static long n_c(const orc_state *s) { return (long)s->nx * s->ny * s->nz; }

void orc_default_params(orc_params *p)
{ /* FluidSolver.hpp:11-38, PoissonSolver.hpp:14-18 */
    memset(p, 0, sizeof *p);
    p->rho = 1.2; p->nu = 1.5e-5; p->dt = 0.01; p->max_steps = 500;
    p->convergence_tol = 1.0; p->poisson_max_iter = 1000; p->poisson_tol = 1e-6;
    p->advection = ORC_UPWIND1; p->T_initial = 20.0; p->T_target = 1e30;
    p->g_accel = 9.81; p->beta = 1.0 / 293.0; p->k_thermal = 0.026; p->cp = 1005.0;
    p->nthreads = 1;
}

/* ---- one-direction advection operators, FluidSolver.cpp:47-81 ---------- */
static inline double adv_upwind(double m, double c, double pl, double vel, double h)
{
    if (vel > 0.0) return vel * (c - m) / h;
    return vel * (pl - c) / h;
}

/* QUICK as the reference writes it, including its u<0 branch (:64-69). */
static inline double adv_quick(double mm, double m, double c, double pl, double vel, double h)
{
    double right, left;
    if (vel >= 0.0) {
        right = (6.0 * c + 3.0 * pl - m) / 8.0;
        left  = (6.0 * m + 3.0 * c - mm) / 8.0;
    } else {
        right = (6.0 * pl + 3.0 * c - mm) / 8.0;
        left  = (6.0 * c + 3.0 * pl - m) / 8.0;
    }
    return vel * (right - left) / h;
}

static inline double adv_lw(double m, double c, double pl, double vel, double dt, double h)
{
    double cr = vel * dt / h;
    double fr = vel * (0.5 * (c + pl) - 0.5 * cr * (pl - c));
    double fl = vel * (0.5 * (m + c) - 0.5 * cr * (c - m));
    return (fr - fl) / h;
}

/* Combine the three directions for one face value (the scheme switch of
 * :132-155 and the diffusion/update of :157-162). q* are the QUICK far
 * upwind values, only read when quick_ok. */
static inline double face_update(int scheme, int quick_ok, const double nb[6],
                                 const double q[3], double c, const double vel[3],
                                 const double h[3], double dt, double nu)
{
    double adv;
    if (scheme == ORC_QUICK && quick_ok) {
        adv = adv_quick(q[0], nb[0], c, nb[1], vel[0], h[0])
            + adv_quick(q[1], nb[2], c, nb[3], vel[1], h[1])
            + adv_quick(q[2], nb[4], c, nb[5], vel[2], h[2]);
    } else if (scheme == ORC_LAX_WENDROFF) {
        adv = adv_lw(nb[0], c, nb[1], vel[0], dt, h[0])
            + adv_lw(nb[2], c, nb[3], vel[1], dt, h[1])
            + adv_lw(nb[4], c, nb[5], vel[2], dt, h[2]);
    } else {
        adv = adv_upwind(nb[0], c, nb[1], vel[0], h[0])
            + adv_upwind(nb[2], c, nb[3], vel[1], h[1])
            + adv_upwind(nb[4], c, nb[5], vel[2], h[2]);
    }
    double i2x = 1.0 / (h[0] * h[0]);
    double i2y = 1.0 / (h[1] * h[1]);
    double i2z = 1.0 / (h[2] * h[2]);
    double dif = nu * (i2x * (nb[0] - 2 * c + nb[1]) +
                       i2y * (nb[2] - 2 * c + nb[3]) +
                       i2z * (nb[4] - 2 * c + nb[5]));
    return c + dt * (-adv + dif);
}

/* compute_intermediate, FluidSolver.cpp:83-309. Writes u*, v*, w* into the
 * state (the reference computes into *_tmp and copies back, :306-308). */
void orc_predictor(orc_state *s, const orc_params *p)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double dt = p->dt, nu = p->nu;
    const int sch = p->advection;
    double *us = (double *)malloc(sizeof(double) * n_u(s));
    double *vs = (double *)malloc(sizeof(double) * n_v(s));
    double *ws = (double *)malloc(sizeof(double) * n_w(s));
    const double *u = s->u, *v = s->v, *w = s->w;
    const int nt = p->nthreads > 1 ? p->nthreads : 1;
    (void)nt;

    /* x-faces, :90-163 */
#ifdef _OPENMP
#pragma omp parallel for collapse(3) schedule(static) num_threads(nt) if (nt > 1)
#endif
    for (int i = 1; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                double h[3] = {s->dxv[i], s->dyv[j], s->dzv[k]};
                double c = u[IU(s, i, j, k)];
                double nb[6], q[3] = {0, 0, 0}, vel[3];
                nb[0] = i > 1 ? u[IU(s, i - 1, j, k)] : 0.0;
                nb[1] = i < nx - 1 ? u[IU(s, i + 1, j, k)] : 0.0;
                nb[2] = j > 0 ? u[IU(s, i, j - 1, k)] : 0.0;
                nb[3] = j < ny - 1 ? u[IU(s, i, j + 1, k)] : 0.0;
                nb[4] = k > 0 ? u[IU(s, i, j, k - 1)] : 0.0;
                nb[5] = k < nz - 1 ? u[IU(s, i, j, k + 1)] : 0.0;
                vel[0] = c;
                vel[1] = (j > 0) ? 0.5 * (v[IV(s, i - 1, j, k)] + v[IV(s, i, j, k)]) : 0.0;
                vel[2] = (k > 0) ? 0.5 * (w[IW(s, i - 1, j, k)] + w[IW(s, i, j, k)]) : 0.0;
                int qok = i >= 2 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2;
                if (sch == ORC_QUICK && qok) {
                    q[0] = u[IU(s, i - 2, j, k)];
                    q[1] = u[IU(s, i, j - 2, k)]; /* wraps at j=1 (reference :139) */
                    q[2] = u[IU(s, i, j, k - 2)]; /* wraps at k=1 (reference :141) */
                }
                us[IU(s, i, j, k)] = face_update(sch, qok, nb, q, c, vel, h, dt, nu);
            }
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
            us[IU(s, 0, j, k)] = u[IU(s, 0, j, k)];
            us[IU(s, nx, j, k)] = u[IU(s, nx, j, k)];
        }

    /* y-faces, :171-236 */
#ifdef _OPENMP
#pragma omp parallel for collapse(3) schedule(static) num_threads(nt) if (nt > 1)
#endif
    for (int i = 0; i < nx; i++)
        for (int j = 1; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                double h[3] = {s->dxv[i], s->dyv[j], s->dzv[k]};
                double c = v[IV(s, i, j, k)];
                double nb[6], q[3] = {0, 0, 0}, vel[3];
                vel[0] = (i > 0) ? 0.5 * (u[IU(s, i, j - 1, k)] + u[IU(s, i, j, k)]) : 0.0;
                nb[0] = i > 0 ? v[IV(s, i - 1, j, k)] : 0.0;
                nb[1] = i < nx - 1 ? v[IV(s, i + 1, j, k)] : 0.0;
                nb[2] = j > 1 ? v[IV(s, i, j - 1, k)] : 0.0;
                nb[3] = j < ny - 1 ? v[IV(s, i, j + 1, k)] : 0.0;
                vel[1] = c;
                vel[2] = (k > 0) ? 0.5 * (w[IW(s, i, j - 1, k)] + w[IW(s, i, j, k)]) : 0.0;
                nb[4] = k > 0 ? v[IV(s, i, j, k - 1)] : 0.0;
                nb[5] = k < nz - 1 ? v[IV(s, i, j, k + 1)] : 0.0;
                int qok = i >= 1 && i <= nx - 2 && j >= 2 && j <= ny - 2 && k >= 1 && k <= nz - 2;
                if (sch == ORC_QUICK && qok) {
                    q[0] = v[IV(s, i - 1, j, k)]; /* adjacent, not 2-away (:208) */
                    q[1] = v[IV(s, i, j - 2, k)];
                    q[2] = v[IV(s, i, j, k - 1)]; /* adjacent (:211) */
                }
                vs[IV(s, i, j, k)] = face_update(sch, qok, nb, q, c, vel, h, dt, nu);
            }
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) {
            vs[IV(s, i, 0, k)] = v[IV(s, i, 0, k)];
            vs[IV(s, i, ny, k)] = v[IV(s, i, ny, k)];
        }

    /* z-faces, :238-303 */
#ifdef _OPENMP
#pragma omp parallel for collapse(3) schedule(static) num_threads(nt) if (nt > 1)
#endif
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 1; k < nz; k++) {
                double h[3] = {s->dxv[i], s->dyv[j], s->dzv[k]};
                double c = w[IW(s, i, j, k)];
                double nb[6], q[3] = {0, 0, 0}, vel[3];
                vel[0] = (i > 0) ? 0.5 * (u[IU(s, i, j, k - 1)] + u[IU(s, i, j, k)]) : 0.0;
                nb[0] = i > 0 ? w[IW(s, i - 1, j, k)] : 0.0;
                nb[1] = i < nx - 1 ? w[IW(s, i + 1, j, k)] : 0.0;
                vel[1] = (j > 0) ? 0.5 * (v[IV(s, i, j, k - 1)] + v[IV(s, i, j, k)]) : 0.0;
                nb[2] = j > 0 ? w[IW(s, i, j - 1, k)] : 0.0;
                nb[3] = j < ny - 1 ? w[IW(s, i, j + 1, k)] : 0.0;
                nb[4] = k > 1 ? w[IW(s, i, j, k - 1)] : 0.0;
                nb[5] = k < nz - 1 ? w[IW(s, i, j, k + 1)] : 0.0;
                vel[2] = c;
                int qok = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 2 && k <= nz - 2;
                if (sch == ORC_QUICK && qok) {
                    q[0] = w[IW(s, i - 1, j, k)]; /* adjacent (:275) */
                    q[1] = w[IW(s, i, j - 1, k)]; /* adjacent (:276) */
                    q[2] = w[IW(s, i, j, k - 2)];
                }
                ws[IW(s, i, j, k)] = face_update(sch, qok, nb, q, c, vel, h, dt, nu);
            }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            ws[IW(s, i, j, 0)] = w[IW(s, i, j, 0)];
            ws[IW(s, i, j, nz)] = w[IW(s, i, j, nz)];
        }

    memcpy(s->u, us, sizeof(double) * n_u(s));
    memcpy(s->v, vs, sizeof(double) * n_v(s));
    memcpy(s->w, ws, sizeof(double) * n_w(s));
    free(us); free(vs); free(ws);
}

/* apply_buoyancy, FluidSolver.cpp:373-386 */
void orc_buoyancy(orc_state *s, const orc_params *p)
{
    const double coeff = p->g_accel * p->beta * p->dt;
    const double tref = p->T_initial;
    for (int i = 0; i < s->nx; i++)
        for (int j = 0; j < s->ny; j++)
            for (int k = 1; k < s->nz; k++) {
                double lo = s->T[IC(s, i, j, k - 1)], hi = s->T[IC(s, i, j, k)];
                s->w[IW(s, i, j, k)] += coeff * (0.5 * (lo + hi) - tref);
            }
}

/* apply_noslip, BoundaryManager.cpp:60-109 */
void orc_apply_noslip(orc_state *s)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const uint8_t *f = s->flags;
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
            if (f[IC(s, 0, j, k)] != ORC_OUTFLOW) s->u[IU(s, 0, j, k)] = 0.0;
        }
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
            if (f[IC(s, nx - 1, j, k)] != ORC_OUTFLOW) s->u[IU(s, nx, j, k)] = 0.0;
        }
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) {
            if (f[IC(s, i, 0, k)] != ORC_OUTFLOW) s->v[IV(s, i, 0, k)] = 0.0;
        }
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) {
            if (f[IC(s, i, ny - 1, k)] != ORC_OUTFLOW) s->v[IV(s, i, ny, k)] = 0.0;
        }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            if (f[IC(s, i, j, 0)] != ORC_OUTFLOW) s->w[IW(s, i, j, 0)] = 0.0;
        }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            if (f[IC(s, i, j, nz - 1)] != ORC_OUTFLOW) s->w[IW(s, i, j, nz)] = 0.0;
        }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                if (f[IC(s, i, j, k)] != ORC_SOLID) continue;
                s->u[IU(s, i, j, k)] = 0.0; s->u[IU(s, i + 1, j, k)] = 0.0;
                s->v[IV(s, i, j, k)] = 0.0; s->v[IV(s, i, j + 1, k)] = 0.0;
                s->w[IW(s, i, j, k)] = 0.0; s->w[IW(s, i, j, k + 1)] = 0.0;
            }
}

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* apply_inflow, BoundaryManager.cpp:111-130; fans in order, last one wins */
void orc_apply_inflow(orc_state *s)
{
    for (int n = 0; n < s->n_fans; n++) {
        const orc_fan *fn = &s->fans[n];
        for (int i = imax(0, fn->i_min); i <= imin(s->nx - 1, fn->i_max); i++)
            for (int j = imax(0, fn->j_min); j <= imin(s->ny - 1, fn->j_max); j++)
                for (int k = imax(0, fn->k_min); k <= imin(s->nz - 1, fn->k_max); k++) {
                    if (fn->vel[0] != 0.0) {
                        s->u[IU(s, i, j, k)] = fn->vel[0];
                        s->u[IU(s, i + 1, j, k)] = fn->vel[0];
                    }
                    if (fn->vel[1] != 0.0) {
                        s->v[IV(s, i, j, k)] = fn->vel[1];
                        s->v[IV(s, i, j + 1, k)] = fn->vel[1];
                    }
                    if (fn->vel[2] != 0.0) {
                        s->w[IW(s, i, j, k)] = fn->vel[2];
                        s->w[IW(s, i, j, k + 1)] = fn->vel[2];
                    }
                }
    }
}

/* apply_outflow, BoundaryManager.cpp:180-220: boundary face := interior face */
void orc_apply_outflow(orc_state *s)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    for (int n = 0; n < s->n_outflows; n++) {
        const orc_outflow *o = &s->outflows[n];
        for (int a = o->a_min; a <= o->a_max; a++)
            for (int b = o->b_min; b <= o->b_max; b++) {
                if (o->axis == 0) {
                    if (o->side == 1) s->u[IU(s, nx, a, b)] = s->u[IU(s, nx - 1, a, b)];
                    else              s->u[IU(s, 0, a, b)] = s->u[IU(s, 1, a, b)];
                } else if (o->axis == 1) {
                    if (o->side == 1) s->v[IV(s, a, ny, b)] = s->v[IV(s, a, ny - 1, b)];
                    else              s->v[IV(s, a, 0, b)] = s->v[IV(s, a, 1, b)];
                } else {
                    if (o->side == 1) s->w[IW(s, a, b, nz)] = s->w[IW(s, a, b, nz - 1)];
                    else              s->w[IW(s, a, b, 0)] = s->w[IW(s, a, b, 1)];
                }
            }
    }
}

/* apply_opening_temperature, BoundaryManager.cpp:137-178 */
void orc_apply_opening_temperature(orc_state *s)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    for (int n = 0; n < s->n_openings; n++) {
        const orc_opening *o = &s->openings[n];
        const int sd = o->side;
        for (int a = o->a_min; a <= o->a_max; a++)
            for (int b = o->b_min; b <= o->b_max; b++) {
                double vf;
                long wall, inner;
                if (o->axis == 0) {
                    vf = s->u[IU(s, sd ? nx : 0, a, b)];
                    wall = IC(s, sd ? nx - 1 : 0, a, b);
                    inner = IC(s, sd ? nx - 2 : 1, a, b);
                } else if (o->axis == 1) {
                    vf = s->v[IV(s, a, sd ? ny : 0, b)];
                    wall = IC(s, a, sd ? ny - 1 : 0, b);
                    inner = IC(s, a, sd ? ny - 2 : 1, b);
                } else {
                    vf = s->w[IW(s, a, b, sd ? nz : 0)];
                    wall = IC(s, a, b, sd ? nz - 1 : 0);
                    inner = IC(s, a, b, sd ? nz - 2 : 1);
                }
                int incoming = sd ? (vf <= 0.0) : (vf >= 0.0);
                s->T[wall] = incoming ? o->T_outside : s->T[inner];
            }
    }
}

/* build_poisson_rhs, FluidSolver.cpp:311-330 */
void orc_poisson_rhs(const orc_state *s, const orc_params *p, double *rhs)
{
    const double factor = -(p->rho / p->dt);
    for (int i = 0; i < s->nx; i++)
        for (int j = 0; j < s->ny; j++)
            for (int k = 0; k < s->nz; k++) {
                long c = IC(s, i, j, k);
                if (s->flags[c] == ORC_SOLID) { rhs[c] = 0.0; continue; }
                double div = (s->u[IU(s, i + 1, j, k)] - s->u[IU(s, i, j, k)]) / s->dxv[i]
                           + (s->v[IV(s, i, j + 1, k)] - s->v[IV(s, i, j, k)]) / s->dyv[j]
                           + (s->w[IW(s, i, j, k + 1)] - s->w[IW(s, i, j, k)]) / s->dzv[k];
                rhs[c] = factor * div;
            }
}

/* Per-axis Laplacian coefficients, PoissonSolver.cpp:34-49: sl[i], sr[i]. */
static void lap_coeffs(const double *d, int n, double *sl, double *sr)
{
    for (int i = 0; i < n; i++) {
        double hl = (i > 0) ? 0.5 * (d[i - 1] + d[i]) : d[i];
        double hr = (i < n - 1) ? 0.5 * (d[i] + d[i + 1]) : d[i];
        sl[i] = 1.0 / (hl * d[i]);
        sr[i] = 1.0 / (hr * d[i]);
    }
}

/* The 7-point row of PoissonSolver.cpp:52-87 for cell (i,j,k) of an
 * (nx,ny,nz) level with flags f: accumulates diag and off in the
 * reference's neighbour order (-x,+x,-y,+y,-z,+z). */
static inline void lap_row(int nx, int ny, int nz, const uint8_t *f, const double *x,
                           int i, int j, int k, const double sc[6],
                           double *diag_out, double *off_out)
{
    double diag = 0.0, off = 0.0;
    const int ni[6] = {i - 1, i + 1, i, i, i, i};
    const int nj[6] = {j, j, j - 1, j + 1, j, j};
    const int nk[6] = {k, k, k, k, k - 1, k + 1};
    for (int q = 0; q < 6; q++) {
        if (ni[q] < 0 || ni[q] >= nx || nj[q] < 0 || nj[q] >= ny || nk[q] < 0 || nk[q] >= nz)
            continue;
        long nc = ((long)ni[q] * ny + nj[q]) * nz + nk[q];
        uint8_t t = f[nc];
        if (t == ORC_OUTFLOW) diag += 2.0 * sc[q];
        else if (t == ORC_SOLID) { /* Neumann */ }
        else { diag += sc[q]; off += sc[q] * x[nc]; }
    }
    if (diag < 1e-30) diag = 1.0;
    *diag_out = diag;
    *off_out = off;
}

static void laplacian_level(int nx, int ny, int nz, const double *dxv, const double *dyv,
                            const double *dzv, const uint8_t *f, const double *x, double *Ax,
                            int nthreads)
{
    double *c = (double *)malloc(sizeof(double) * 2 * (nx + ny + nz));
    double *slx = c, *srx = c + nx, *sly = c + 2 * nx, *sry = sly + ny;
    double *slz = sry + ny, *srz = slz + nz;
    lap_coeffs(dxv, nx, slx, srx);
    lap_coeffs(dyv, ny, sly, sry);
    lap_coeffs(dzv, nz, slz, srz);
    (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 1 ? nthreads : 1) if (nthreads > 1)
#endif
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                long cc = ((long)i * ny + j) * nz + k;
                if (f[cc] == ORC_SOLID) { Ax[cc] = x[cc]; continue; }
                double sc[6] = {slx[i], srx[i], sly[j], sry[j], slz[k], srz[k]};
                double diag, off;
                lap_row(nx, ny, nz, f, x, i, j, k, sc, &diag, &off);
                Ax[cc] = diag * x[cc] - off;
            }
    free(c);
}

void orc_apply_laplacian(const orc_state *s, const double *x, double *Ax, int nthreads)
{
    laplacian_level(s->nx, s->ny, s->nz, s->dxv, s->dyv, s->dzv, s->flags, x, Ax, nthreads);
}

/* PoissonSolver::solve, PoissonSolver.cpp:93-158: unpreconditioned CG from
 * x0 = 0 with SOLID rows of the right-hand side zeroed. */
double orc_poisson_cg(orc_state *s, const double *rhs, int max_iter, double tol,
                      int nthreads, int *iters)
{
    const long n = n_c(s);
    const int nt = nthreads > 1 ? nthreads : 1;
    (void)nt;
    double *x = s->p;
    double *r = (double *)malloc(sizeof(double) * n);
    double *d = (double *)malloc(sizeof(double) * n);
    double *Ad = (double *)malloc(sizeof(double) * n);
    for (long i = 0; i < n; i++) {
        x[i] = 0.0;
        r[i] = (s->flags[i] == ORC_SOLID) ? 0.0 : rhs[i];
        d[i] = r[i];
    }
    double rr = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : rr) schedule(static) num_threads(nt) if (nt > 1)
#endif
    for (long i = 0; i < n; i++) rr += r[i] * r[i];
    double bnorm = sqrt(rr);
    double res = 0.0;
    int it_done = 0;
    if (bnorm < 1e-15) {
        if (iters) *iters = 0;
        free(r); free(d); free(Ad);
        return 0.0;
    }
    int converged = 0;
    for (int it = 0; it < max_iter; it++) {
        orc_apply_laplacian(s, d, Ad, nthreads);
        double dAd = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : dAd) schedule(static) num_threads(nt) if (nt > 1)
#endif
        for (long i = 0; i < n; i++) dAd += d[i] * Ad[i];
        if (fabs(dAd) < 1e-30) break;
        double alpha = rr / dAd;
        double rr_new = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : rr_new) schedule(static) num_threads(nt) if (nt > 1)
#endif
        for (long i = 0; i < n; i++) {
            x[i] += alpha * d[i];
            r[i] -= alpha * Ad[i];
            rr_new += r[i] * r[i];
        }
        res = sqrt(rr_new) / bnorm;
        if (res < tol) { it_done = it + 1; converged = 1; break; }
        double beta = rr_new / rr;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
#endif
        for (long i = 0; i < n; i++) d[i] = r[i] + beta * d[i];
        rr = rr_new;
    }
    if (!converged) { it_done = max_iter; res = sqrt(rr) / bnorm; }
    if (iters) *iters = it_done;
    free(r); free(d); free(Ad);
    return res;
}

/* ---- reference multigrid, PoissonSolver.cpp:164-505 -------------------- */
typedef struct {
    int nx, ny, nz;
    double *dxv, *dyv, *dzv;
    uint8_t *f;
} mg_level;

/* smooth, :172-239: lexicographic Gauss-Seidel, OUTFLOW pinned at 0 */
static void mg_smooth(const mg_level *L, double *x, const double *b, int sweeps)
{
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    double *c = (double *)malloc(sizeof(double) * 2 * (nx + ny + nz));
    double *slx = c, *srx = c + nx, *sly = c + 2 * nx, *sry = sly + ny;
    double *slz = sry + ny, *srz = slz + nz;
    lap_coeffs(L->dxv, nx, slx, srx);
    lap_coeffs(L->dyv, ny, sly, sry);
    lap_coeffs(L->dzv, nz, slz, srz);
    for (int sw = 0; sw < sweeps; sw++)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++) {
                    long cc = ((long)i * ny + j) * nz + k;
                    if (L->f[cc] == ORC_SOLID) continue;
                    if (L->f[cc] == ORC_OUTFLOW) { x[cc] = 0.0; continue; }
                    double sc[6] = {slx[i], srx[i], sly[j], sry[j], slz[k], srz[k]};
                    double diag, off;
                    lap_row(nx, ny, nz, L->f, x, i, j, k, sc, &diag, &off);
                    x[cc] = (b[cc] + off) / diag;
                }
    free(c);
}

/* compute_residual, :244-309: zero on SOLID and OUTFLOW rows */
static void mg_residual(const mg_level *L, const double *x, const double *b, double *r)
{
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    double *c = (double *)malloc(sizeof(double) * 2 * (nx + ny + nz));
    double *slx = c, *srx = c + nx, *sly = c + 2 * nx, *sry = sly + ny;
    double *slz = sry + ny, *srz = slz + nz;
    lap_coeffs(L->dxv, nx, slx, srx);
    lap_coeffs(L->dyv, ny, sly, sry);
    lap_coeffs(L->dzv, nz, slz, srz);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                long cc = ((long)i * ny + j) * nz + k;
                if (L->f[cc] == ORC_SOLID || L->f[cc] == ORC_OUTFLOW) { r[cc] = 0.0; continue; }
                double sc[6] = {slx[i], srx[i], sly[j], sry[j], slz[k], srz[k]};
                double diag, off;
                lap_row(nx, ny, nz, L->f, x, i, j, k, sc, &diag, &off);
                r[cc] = b[cc] - (diag * x[cc] - off);
            }
    free(c);
}

/* v_cycle, :358-441 */
static void mg_vcycle(const mg_level *L, double *x, const double *b, int level, int max_level)
{
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    if (level == max_level || (nx <= 4 && ny <= 4 && nz <= 4)) {
        mg_smooth(L, x, b, 20);
        return;
    }
    mg_smooth(L, x, b, 2);
    long nf = (long)nx * ny * nz;
    double *res = (double *)malloc(sizeof(double) * nf);
    mg_residual(L, x, b, res);

    mg_level C;
    C.nx = imax(1, nx / 2); C.ny = imax(1, ny / 2); C.nz = imax(1, nz / 2);
    C.dxv = (double *)malloc(sizeof(double) * C.nx);
    C.dyv = (double *)malloc(sizeof(double) * C.ny);
    C.dzv = (double *)malloc(sizeof(double) * C.nz);
    for (int i = 0; i < C.nx; i++) C.dxv[i] = L->dxv[2 * i] + (2 * i + 1 < nx ? L->dxv[2 * i + 1] : 0.0);
    for (int j = 0; j < C.ny; j++) C.dyv[j] = L->dyv[2 * j] + (2 * j + 1 < ny ? L->dyv[2 * j + 1] : 0.0);
    for (int k = 0; k < C.nz; k++) C.dzv[k] = L->dzv[2 * k] + (2 * k + 1 < nz ? L->dzv[2 * k + 1] : 0.0);
    long nc = (long)C.nx * C.ny * C.nz;
    C.f = (uint8_t *)malloc(nc);
    for (int ic = 0; ic < C.nx; ic++)
        for (int jc = 0; jc < C.ny; jc++)
            for (int kc = 0; kc < C.nz; kc++) {
                int outflow = 0, all_solid = 1;
                for (int a = 0; a < 2; a++)
                    for (int bb = 0; bb < 2; bb++)
                        for (int cidx = 0; cidx < 2; cidx++) {
                            int fi = imin(2 * ic + a, nx - 1);
                            int fj = imin(2 * jc + bb, ny - 1);
                            int fk = imin(2 * kc + cidx, nz - 1);
                            uint8_t t = L->f[((long)fi * ny + fj) * nz + fk];
                            if (t == ORC_OUTFLOW) outflow = 1;
                            if (t != ORC_SOLID) all_solid = 0;
                        }
                long cc = ((long)ic * C.ny + jc) * C.nz + kc;
                C.f[cc] = outflow ? ORC_OUTFLOW : (all_solid ? ORC_SOLID : ORC_FLUID);
            }
    /* injection, :314-330 */
    double *rc = (double *)calloc(nc, sizeof(double));
    for (int ic = 0; ic < C.nx; ic++)
        for (int jc = 0; jc < C.ny; jc++)
            for (int kc = 0; kc < C.nz; kc++) {
                int fi = 2 * ic, fj = 2 * jc, fk = 2 * kc;
                long cc = ((long)ic * C.ny + jc) * C.nz + kc;
                rc[cc] = (fi < nx && fj < ny && fk < nz) ? res[((long)fi * ny + fj) * nz + fk] : 0.0;
            }
    double *ec = (double *)calloc(nc, sizeof(double));
    mg_vcycle(&C, ec, rc, level + 1, max_level);
    /* piecewise-constant prolongation and correction, :338-353, :431-437 */
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                int ic = imin(i / 2, C.nx - 1), jc = imin(j / 2, C.ny - 1), kc = imin(k / 2, C.nz - 1);
                long cc = ((long)i * ny + j) * nz + k;
                double e = ec[((long)ic * C.ny + jc) * C.nz + kc];
                if (L->f[cc] != ORC_OUTFLOW) x[cc] += e;
            }
    mg_smooth(L, x, b, 2);
    free(res); free(rc); free(ec); free(C.f); free(C.dxv); free(C.dyv); free(C.dzv);
}

double orc_poisson_mg(orc_state *s, const double *rhs, int max_iter, double tol, int *iters)
{ /* MultigridPoissonSolver::solve, :446-505 */
    const long n = n_c(s);
    double *x = s->p;
    for (long i = 0; i < n; i++) x[i] = 0.0;
    double bn = 0.0;
    for (long i = 0; i < n; i++)
        if (s->flags[i] != ORC_SOLID) bn += rhs[i] * rhs[i];
    bn = sqrt(bn);
    if (bn < 1e-15) { if (iters) *iters = 0; return 0.0; }
    int md = imin(s->nx, imin(s->ny, s->nz));
    int max_level = 0;
    while (md > 4) { md /= 2; max_level++; }
    if (max_level < 1) max_level = 1;
    mg_level L = {s->nx, s->ny, s->nz, (double *)s->dxv, (double *)s->dyv, (double *)s->dzv,
                  (uint8_t *)s->flags};
    double *rv = (double *)malloc(sizeof(double) * n);
    double rel = 0.0;
    for (int it = 0; it < max_iter; it++) {
        mg_vcycle(&L, x, rhs, 0, max_level);
        mg_residual(&L, x, rhs, rv);
        double rr = 0.0;
        for (long i = 0; i < n; i++) rr += rv[i] * rv[i];
        rel = sqrt(rr) / bn;
        if (rel < tol) { if (iters) *iters = it + 1; free(rv); return rel; }
    }
    mg_residual(&L, x, rhs, rv);
    double rr = 0.0;
    for (long i = 0; i < n; i++) rr += rv[i] * rv[i];
    if (iters) *iters = max_iter;
    free(rv);
    return sqrt(rr) / bn;
}

/* apply_correction, FluidSolver.cpp:332-371 */
void orc_correction(orc_state *s, const orc_params *p)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double dtr = p->dt / p->rho;
    const int nt = p->nthreads > 1 ? p->nthreads : 1;
    (void)nt;
#ifdef _OPENMP
#pragma omp parallel for collapse(3) schedule(static) num_threads(nt) if (nt > 1)
#endif
    for (int i = 1; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                double h = 0.5 * (s->dxv[i - 1] + s->dxv[i]);
                double g = (s->p[IC(s, i, j, k)] - s->p[IC(s, i - 1, j, k)]) / h;
                s->u[IU(s, i, j, k)] -= dtr * g;
            }
#ifdef _OPENMP
#pragma omp parallel for collapse(3) schedule(static) num_threads(nt) if (nt > 1)
#endif
    for (int i = 0; i < nx; i++)
        for (int j = 1; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                double h = 0.5 * (s->dyv[j - 1] + s->dyv[j]);
                double g = (s->p[IC(s, i, j, k)] - s->p[IC(s, i, j - 1, k)]) / h;
                s->v[IV(s, i, j, k)] -= dtr * g;
            }
#ifdef _OPENMP
#pragma omp parallel for collapse(3) schedule(static) num_threads(nt) if (nt > 1)
#endif
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 1; k < nz; k++) {
                double h = 0.5 * (s->dzv[k - 1] + s->dzv[k]);
                double g = (s->p[IC(s, i, j, k)] - s->p[IC(s, i, j, k - 1)]) / h;
                s->w[IW(s, i, j, k)] -= dtr * g;
            }
}

/* advect_temperature, FluidSolver.cpp:388-456 */
void orc_advect_temperature(orc_state *s, const orc_params *p)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double dt = p->dt;
    const double alpha = p->k_thermal / (p->rho * p->cp);
    const int quick = p->advection == ORC_QUICK, lw = p->advection == ORC_LAX_WENDROFF;
    const long n = n_c(s);
    double *Tn = (double *)malloc(sizeof(double) * n);
    const double *T = s->T;
    const int nt = p->nthreads > 1 ? p->nthreads : 1;
    (void)nt;
#ifdef _OPENMP
#pragma omp parallel for collapse(3) schedule(static) num_threads(nt) if (nt > 1)
#endif
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                long c = IC(s, i, j, k);
                if (s->flags[c] != ORC_FLUID) { Tn[c] = T[c]; continue; }
                const double hx = s->dxv[i], hy = s->dyv[j], hz = s->dzv[k];
                const double Tc = T[c];
                const double uc = 0.5 * (s->u[IU(s, i, j, k)] + s->u[IU(s, i + 1, j, k)]);
                const double vc = 0.5 * (s->v[IV(s, i, j, k)] + s->v[IV(s, i, j + 1, k)]);
                const double wc = 0.5 * (s->w[IW(s, i, j, k)] + s->w[IW(s, i, j, k + 1)]);
/* neighbour T with Neumann (T = Tc) outside the domain and at SOLID, :417-422 */
#define TN(a, b, cc_)                                                                   \
    (((a) < 0 || (a) >= nx || (b) < 0 || (b) >= ny || (cc_) < 0 || (cc_) >= nz)         \
         ? Tc                                                                           \
         : (s->flags[IC(s, a, b, cc_)] == ORC_SOLID ? Tc : T[IC(s, a, b, cc_)]))
                const double xm = TN(i - 1, j, k), xp = TN(i + 1, j, k);
                const double ym = TN(i, j - 1, k), yp = TN(i, j + 1, k);
                const double zm = TN(i, j, k - 1), zp = TN(i, j, k + 1);
                double adv;
                if (quick && i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2) {
                    const double xmm = TN(i - 2, j, k), ymm = TN(i, j - 2, k), zmm = TN(i, j, k - 2);
                    adv = adv_quick(xmm, xm, Tc, xp, uc, hx)
                        + adv_quick(ymm, ym, Tc, yp, vc, hy)
                        + adv_quick(zmm, zm, Tc, zp, wc, hz);
                } else if (lw) {
                    adv = adv_lw(xm, Tc, xp, uc, dt, hx)
                        + adv_lw(ym, Tc, yp, vc, dt, hy)
                        + adv_lw(zm, Tc, zp, wc, dt, hz);
                } else {
                    adv = adv_upwind(xm, Tc, xp, uc, hx)
                        + adv_upwind(ym, Tc, yp, vc, hy)
                        + adv_upwind(zm, Tc, zp, wc, hz);
                }
#undef TN
                double dif = alpha * ((xm - 2 * Tc + xp) / (hx * hx) +
                                      (ym - 2 * Tc + yp) / (hy * hy) +
                                      (zm - 2 * Tc + zp) / (hz * hz));
                Tn[c] = Tc + dt * (-adv + dif);
            }
    memcpy(s->T, Tn, sizeof(double) * n);
    free(Tn);
}

double orc_T_mean(const orc_state *s, const orc_params *p)
{ /* compute_T_mean, FluidSolver.cpp:458-468 */
    double sum = 0.0;
    long cnt = 0;
    const long n = n_c(s);
    for (long c = 0; c < n; c++)
        if (s->flags[c] == ORC_FLUID) { sum += s->T[c]; ++cnt; }
    return cnt > 0 ? sum / (double)cnt : p->T_initial;
}

double orc_divergence_max(const orc_state *s)
{ /* compute_divergence_max, FluidSolver.cpp:470-484 */
    double mx = 0.0;
    for (int i = 0; i < s->nx; i++)
        for (int j = 0; j < s->ny; j++)
            for (int k = 0; k < s->nz; k++) {
                if (s->flags[IC(s, i, j, k)] != ORC_FLUID) continue;
                double d = fabs((s->u[IU(s, i + 1, j, k)] - s->u[IU(s, i, j, k)]) / s->dxv[i]
                              + (s->v[IV(s, i, j + 1, k)] - s->v[IV(s, i, j, k)]) / s->dyv[j]
                              + (s->w[IW(s, i, j, k + 1)] - s->w[IW(s, i, j, k)]) / s->dzv[k]);
                if (d > mx) mx = d;
            }
    return mx;
}

void orc_init(orc_state *s, const orc_params *p)
{ /* FluidSolver ctor, FluidSolver.cpp:18-32 */
    const long n = n_c(s);
    for (long c = 0; c < n; c++) s->T[c] = p->T_initial;
    if (p->thermal) orc_apply_opening_temperature(s);
}

static double max_abs_diff(const double *a, const double *b, long n, double acc)
{
    for (long i = 0; i < n; i++) {
        double d = fabs(a[i] - b[i]);
        if (d > acc) acc = d;
    }
    return acc;
}

void orc_step(orc_state *s, const orc_params *p, int *step_counter, orc_result *out)
{ /* FluidSolver::step, FluidSolver.cpp:486-550 */
    const long nu = n_u(s), nv = n_v(s), nw = n_w(s), n = n_c(s);
    double *u0 = (double *)malloc(sizeof(double) * nu);
    double *v0 = (double *)malloc(sizeof(double) * nv);
    double *w0 = (double *)malloc(sizeof(double) * nw);
    memcpy(u0, s->u, sizeof(double) * nu);
    memcpy(v0, s->v, sizeof(double) * nv);
    memcpy(w0, s->w, sizeof(double) * nw);

    orc_predictor(s, p);
    if (p->buoyancy) orc_buoyancy(s, p);
    orc_apply_noslip(s);
    orc_apply_inflow(s);

    double *rhs = (double *)malloc(sizeof(double) * n);
    orc_poisson_rhs(s, p, rhs);
    int iters = 0;
    double pres = p->use_multigrid
        ? orc_poisson_mg(s, rhs, p->poisson_max_iter, p->poisson_tol, &iters)
        : orc_poisson_cg(s, rhs, p->poisson_max_iter, p->poisson_tol, p->nthreads, &iters);
    free(rhs);

    orc_correction(s, p);
    orc_apply_noslip(s);
    orc_apply_inflow(s);
    orc_apply_outflow(s);

    double tmean = p->T_initial;
    if (p->thermal) {
        orc_apply_opening_temperature(s);
        orc_advect_temperature(s, p);
        orc_apply_opening_temperature(s);
        tmean = orc_T_mean(s, p);
    }

    double dv = 0.0;
    dv = max_abs_diff(s->u, u0, nu, dv);
    dv = max_abs_diff(s->v, v0, nv, dv);
    dv = max_abs_diff(s->w, w0, nw, dv);
    dv /= p->dt;
    free(u0); free(v0); free(w0);

    (*step_counter)++;
    int conv;
    if (p->thermal && p->T_target < 1e29)
        conv = (p->T_target >= p->T_initial) ? (tmean >= p->T_target) : (tmean <= p->T_target);
    else
        conv = dv < p->convergence_tol;

    out->step = *step_counter;
    out->velocity_change = dv;
    out->pressure_residual = pres;
    out->divergence_max = orc_divergence_max(s);
    out->converged = conv;
    out->T_mean = tmean;
    out->poisson_iters = iters;
}


/* ---- D3Q19 lattice Boltzmann, LBMSolver.cpp -----------------------------
 * A restatement, not a transcription: streaming is written as a pull. The
 * reference pushes into a zeroed buffer (:151-190): a population moves to
 * the neighbour in its direction, or, when that neighbour is outside the
 * domain or SOLID, is added to the same cell in the direction OPP[q]. Its
 * OPP table (:28) reverses the axis directions but, for the twelve diagonal
 * ones, only one component (7 (+1,+1,0) <-> 8 (-1,+1,0), ...), so a target
 * population of a cell may receive the streamed value from upstream, the
 * bounced value from the cell itself, both (summed) or neither (zero); the
 * pull below sums exactly those. A SOLID cell receives nothing. */
static const int LBM_E[19][3] = { /* directions, LBMSolver.cpp:14-16 */
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1},
    {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};
static const int LBM_OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

static double lbm_w(int q) { return q == 0 ? 1.0 / 3.0 : (q < 7 ? 1.0 / 18.0 : 1.0 / 36.0); }

/* equilibrium, LBMSolver.hpp:85-90 (cs2 = 1/3 as a double) */
static double lbm_feq(int q, double rho, double ux, double uy, double uz)
{
    const double cs2 = 1.0 / 3.0;
    const double eu = LBM_E[q][0] * ux + LBM_E[q][1] * uy + LBM_E[q][2] * uz;
    const double usq = ux * ux + uy * uy + uz * uz;
    return lbm_w(q) * rho * (1.0 + eu / cs2 + eu * eu / (2.0 * cs2 * cs2) - usq / (2.0 * cs2));
}

/* density and velocity of one cell's populations (:114-125, :306-317) */
static void lbm_moments(const double *f, long nc, long c, double *rho, double *ux, double *uy,
                        double *uz)
{
    double r = 0.0, x = 0.0, y = 0.0, z = 0.0;
    for (int q = 0; q < 19; q++) {
        const double fq = f[q * nc + c];
        r += fq;
        x += fq * LBM_E[q][0];
        y += fq * LBM_E[q][1];
        z += fq * LBM_E[q][2];
    }
    if (r > 1e-14) {
        x /= r;
        y /= r;
        z /= r;
    }
    *rho = r; *ux = x; *uy = y; *uz = z;
}

static int lbm_in(const orc_state *s, int i, int j, int k)
{
    return i >= 0 && i < s->nx && j >= 0 && j < s->ny && k >= 0 && k < s->nz;
}

/* inflow cells at equilibrium with their fan's velocity, fans in order (:81-100) */
static void lbm_inflow(const orc_state *s, const orc_lbm *L)
{
    const long nc = n_c(s);
    const double s2l = L->dt / L->dx_ref;
    for (int n = 0; n < s->n_fans; n++) {
        const orc_fan *fn = &s->fans[n];
        const double vx = fn->vel[0] * s2l, vy = fn->vel[1] * s2l, vz = fn->vel[2] * s2l;
        for (int i = fn->i_min; i <= fn->i_max; i++)
            for (int j = fn->j_min; j <= fn->j_max; j++)
                for (int k = fn->k_min; k <= fn->k_max; k++) {
                    if (!lbm_in(s, i, j, k)) continue;
                    const long c = IC(s, i, j, k);
                    if (s->flags[c] != ORC_INFLOW) continue;
                    for (int q = 0; q < 19; q++) L->f[q * nc + c] = lbm_feq(q, 1.0, vx, vy, vz);
                }
    }
}

/* BGK on FLUID and OUTFLOW cells (:106-144) */
static void lbm_collide(const orc_state *s, orc_lbm *L)
{
    const long nc = n_c(s);
    const double inv_tau = 1.0 / L->tau;
    for (long c = 0; c < nc; c++) {
        if (s->flags[c] == ORC_SOLID || s->flags[c] == ORC_INFLOW) continue;
        double rho, ux, uy, uz;
        lbm_moments(L->f, nc, c, &rho, &ux, &uy, &uz);
        for (int q = 0; q < 19; q++) {
            double *fq = &L->f[q * nc + c];
            *fq += inv_tau * (lbm_feq(q, rho, ux, uy, uz) - *fq);
        }
    }
}

/* streaming as a pull into f2, then the buffers swap (:151-190) */
static void lbm_stream(const orc_state *s, orc_lbm *L)
{
    const long nc = n_c(s);
    for (int i = 0; i < s->nx; i++)
        for (int j = 0; j < s->ny; j++)
            for (int k = 0; k < s->nz; k++) {
                const long c = IC(s, i, j, k);
                for (int q = 0; q < 19; q++) {
                    double v = 0.0;
                    if (s->flags[c] != ORC_SOLID) {
                        /* streamed in from upstream */
                        const int ui = i - LBM_E[q][0], uj = j - LBM_E[q][1], uk = k - LBM_E[q][2];
                        if (lbm_in(s, ui, uj, uk) && s->flags[IC(s, ui, uj, uk)] != ORC_SOLID)
                            v += L->f[q * nc + IC(s, ui, uj, uk)];
                        /* this cell's population OPP[q], bounced at a wall or SOLID */
                        const int b = LBM_OPP[q];
                        const int bi = i + LBM_E[b][0], bj = j + LBM_E[b][1], bk = k + LBM_E[b][2];
                        if (!lbm_in(s, bi, bj, bk) || s->flags[IC(s, bi, bj, bk)] == ORC_SOLID)
                            v += L->f[b * nc + c];
                    }
                    L->f2[q * nc + c] = v;
                }
            }
    double *t = L->f;
    L->f = L->f2;
    L->f2 = t;
}

/* zero-gradient copy of one patch's OUTFLOW cells from one cell inside (:196-282) */
static void lbm_outflow_patch(const orc_state *s, orc_lbm *L, int axis, int side, int a0, int a1,
                              int b0, int b1)
{
    const long nc = n_c(s);
    const int n_ax = axis == 0 ? s->nx : (axis == 1 ? s->ny : s->nz);
    const int wall = side ? n_ax - 1 : 0, src = side ? n_ax - 2 : 1;
    if (src < 0 || src >= n_ax) return;
    for (int a = a0; a <= a1; a++)
        for (int b = b0; b <= b1; b++) {
            int cw[3], cs[3];
            if (axis == 0) { cw[0] = wall; cw[1] = a; cw[2] = b; }
            else if (axis == 1) { cw[0] = a; cw[1] = wall; cw[2] = b; }
            else { cw[0] = a; cw[1] = b; cw[2] = wall; }
            if (!lbm_in(s, cw[0], cw[1], cw[2])) continue;
            cs[0] = cw[0]; cs[1] = cw[1]; cs[2] = cw[2];
            cs[axis] = src;
            const long co = IC(s, cw[0], cw[1], cw[2]);
            if (s->flags[co] != ORC_OUTFLOW) continue;
            const long ci = IC(s, cs[0], cs[1], cs[2]);
            for (int q = 0; q < 19; q++) L->f[q * nc + co] = L->f[q * nc + ci];
        }
}

/* cell velocities, p = rho/3 on non-SOLID cells, MAC faces (:288-376) */
static void lbm_macroscopic(orc_state *s, orc_lbm *L)
{
    const long nc = n_c(s);
    const double l2s = L->dx_ref / L->dt;
    for (long c = 0; c < nc; c++) {
        if (s->flags[c] == ORC_SOLID) {
            L->ux[c] = L->uy[c] = L->uz[c] = 0.0;
            continue;
        }
        double rho, ux, uy, uz;
        lbm_moments(L->f, nc, c, &rho, &ux, &uy, &uz);
        L->ux[c] = ux * l2s;
        L->uy[c] = uy * l2s;
        L->uz[c] = uz * l2s;
        s->p[c] = rho * (1.0 / 3.0);
    }
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    for (int i = 0; i <= nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                const double lo = L->ux[IC(s, i > 0 ? i - 1 : 0, j, k)];
                const double hi = L->ux[IC(s, i < nx ? i : nx - 1, j, k)];
                s->u[IU(s, i, j, k)] = (i == 0) ? hi : (i == nx ? lo : 0.5 * (lo + hi));
            }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j <= ny; j++)
            for (int k = 0; k < nz; k++) {
                const double lo = L->uy[IC(s, i, j > 0 ? j - 1 : 0, k)];
                const double hi = L->uy[IC(s, i, j < ny ? j : ny - 1, k)];
                s->v[IV(s, i, j, k)] = (j == 0) ? hi : (j == ny ? lo : 0.5 * (lo + hi));
            }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k <= nz; k++) {
                const double lo = L->uz[IC(s, i, j, k > 0 ? k - 1 : 0)];
                const double hi = L->uz[IC(s, i, j, k < nz ? k : nz - 1)];
                s->w[IW(s, i, j, k)] = (k == 0) ? hi : (k == nz ? lo : 0.5 * (lo + hi));
            }
}

void orc_lbm_init(orc_state *s, orc_lbm *L)
{ /* constructor and initialize_f (:39-75): rest equilibrium, then inflow */
    const long nc = n_c(s);
    for (int q = 0; q < 19; q++)
        for (long c = 0; c < nc; c++) L->f[q * nc + c] = lbm_w(q);
    lbm_inflow(s, L);
}

void orc_lbm_step(orc_state *s, orc_lbm *L, int *step_counter, orc_result *out)
{ /* LBMSolver::step (:403-444) */
    const long nu = n_u(s), nv = n_v(s), nw = n_w(s);
    double *prev = (double *)malloc(sizeof(double) * (size_t)(nu + nv + nw));
    memcpy(prev, s->u, sizeof(double) * nu);
    memcpy(prev + nu, s->v, sizeof(double) * nv);
    memcpy(prev + nu + nv, s->w, sizeof(double) * nw);
    lbm_inflow(s, L);
    lbm_collide(s, L);
    lbm_stream(s, L);
    for (int n = 0; n < s->n_outflows; n++) {
        const orc_outflow *o = &s->outflows[n];
        lbm_outflow_patch(s, L, o->axis, o->side, o->a_min, o->a_max, o->b_min, o->b_max);
    }
    for (int n = 0; n < s->n_openings; n++) {
        const orc_opening *o = &s->openings[n];
        lbm_outflow_patch(s, L, o->axis, o->side, o->a_min, o->a_max, o->b_min, o->b_max);
    }
    lbm_macroscopic(s, L);
    double dv = 0.0;
    for (long e = 0; e < nu; e++) { const double d = fabs(s->u[e] - prev[e]); if (d > dv) dv = d; }
    for (long e = 0; e < nv; e++) { const double d = fabs(s->v[e] - prev[nu + e]); if (d > dv) dv = d; }
    for (long e = 0; e < nw; e++) { const double d = fabs(s->w[e] - prev[nu + nv + e]); if (d > dv) dv = d; }
    free(prev);
    dv /= L->dt;
    (*step_counter)++;
    memset(out, 0, sizeof *out);
    out->step = *step_counter;
    out->velocity_change = dv;
    out->pressure_residual = 0.0;
    out->divergence_max = orc_divergence_max(s);
    out->converged = dv < L->convergence_tol;
    out->T_mean = 0.0;
}
