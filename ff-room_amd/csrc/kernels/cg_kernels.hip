// cg_kernels.hip - the reference pressure solve (PoissonSolver.cpp:22-158) in
// FP64, device-resident: every scalar of the CG recurrence (rr, alpha,
// beta, the stopping test) is computed by the last block of a deterministic
// two-pass reduction, so the host only launches and never waits inside the
// iteration. A kernel that finds cg->done set returns at once, which lets
// the host enqueue iterations ahead of the convergence test.
#include "../ffr_kernels.hpp"

#pragma clang fp contract(off)

namespace ffr {

// Row c of A (PoissonSolver.cpp:52-87): neighbours in the reference order
// -x,+x,-y,+y,-z,+z; OUTFLOW adds 2*scale to the diagonal, SOLID and the
// domain walls add nothing, other cells couple with -scale.
__device__ __forceinline__ double lap_row(const Geo &g, const double *x, int i, int j, int k,
                                          uint32_t c)
{
    const uint8_t *f = g.flags;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const uint32_t si = (uint32_t)ny * nz, sj = nz;
    const double *L = g.lap;
    double diag = 0.0, off = 0.0;
#define FFR_FACE(cond, nc, sc)                          \
    if (cond) {                                         \
        const uint8_t t = f[nc];                        \
        const double s_ = (sc);                         \
        if (t == OUTFLOW) diag += 2.0 * s_;             \
        else if (t != SOLID) { diag += s_; off += s_ * x[nc]; } \
    }
    FFR_FACE(i > 0, c - si, L[i])
    FFR_FACE(i < nx - 1, c + si, L[nx + i])
    FFR_FACE(j > 0, c - sj, L[2 * nx + j])
    FFR_FACE(j < ny - 1, c + sj, L[2 * nx + ny + j])
    FFR_FACE(k > 0, c - 1, L[2 * nx + 2 * ny + k])
    FFR_FACE(k < nz - 1, c + 1, L[2 * nx + 2 * ny + nz + k])
#undef FFR_FACE
    if (diag < 1e-30) diag = 1.0;
    return diag * x[c] - off;
}

__device__ __forceinline__ void decode_cell(const Geo &g, uint32_t c, int &i, int &j, int &k)
{
    const uint32_t ii = fdiv(c, g.f_nynz), rem = c - ii * (g.ny * g.nz);
    const uint32_t jj = fdiv(rem, g.f_nz);
    i = ii;
    j = jj;
    k = rem - jj * g.nz;
}

__global__ void __launch_bounds__(kBlock) k_laplacian(Geo g, const double *x, double *Ax)
{
    uint32_t sb, se;
    block_span(g.nc, xcd_block(), sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        if (g.flags[c] == SOLID) {
            Ax[c] = x[c];
            continue;
        }
        int i, j, k;
        decode_cell(g, c, i, j, k);
        Ax[c] = lap_row(g, x, i, j, k, c);
    }
}

void launch_laplacian(hipStream_t s, const Geo &g, const double *x, double *Ax)
{
    hipLaunchKernelGGL(k_laplacian, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, x, Ax);
}

// x = 0, r = d = b (SOLID rows 0), rr = r.r (PoissonSolver.cpp:96-117)
__global__ void __launch_bounds__(kBlock) k_cg_init(Geo g, const double *b, double *x, double *r,
                                                    double *d, CgState *cg, RedScratch red,
                                                    int max_iter, double tol)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t blk = xcd_block();
    double rr = 0.0;
    uint32_t sb, se;
    block_span(g.nc, blk, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        const double rv = (g.flags[c] == SOLID) ? 0.0 : b[c];
        x[c] = 0.0;
        r[c] = rv;
        d[c] = rv;
        rr += rv * rv;
    }
    double total;
    if (grid_sum_last(rr, red.partials, &red.tickets[1], blk, lds, &total)) {
        cg->max_iter = max_iter;
        cg->tol = tol;
        const double bn = sqrt(total);
        cg->rr = total;
        cg->bnorm = bn;
        cg->iter = 0;
        cg->alpha = cg->beta = 0.0;
        if (bn < 1e-15) {
            cg->done = 1;
            cg->iters = 0;
            cg->res = 0.0;
        } else if (cg->max_iter <= 0) {
            cg->done = 1;
            cg->iters = cg->max_iter;
            cg->res = sqrt(total) / bn;
        } else {
            cg->done = 0;
            cg->iters = 0;
            cg->res = 1.0;
        }
    }
}

void launch_cg_init(hipStream_t s, const Geo &g, const double *b, double *x, double *r,
                    double *d, CgState *cg, RedScratch red, int max_iter, double tol)
{
    hipLaunchKernelGGL(k_cg_init, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, b, x, r, d, cg,
                       red, max_iter, tol);
}

// Ad = A d and dAd = d.Ad; alpha = rr / dAd, or stop when |dAd| < 1e-30 (:120-129)
__global__ void __launch_bounds__(kBlock) k_cg_spmv(Geo g, const double *d, double *Ad,
                                                    CgState *cg, RedScratch red)
{
    __shared__ double lds[kBlock / kWave];
    if (cg->done) return;
    const uint32_t blk = xcd_block();
    double dq = 0.0;
    uint32_t sb, se;
    block_span(g.nc, blk, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        double a;
        const double dc = d[c];
        if (g.flags[c] == SOLID) {
            a = dc;
        } else {
            int i, j, k;
            decode_cell(g, c, i, j, k);
            a = lap_row(g, d, i, j, k, c);
        }
        Ad[c] = a;
        dq += dc * a;
    }
    double dAd;
    if (grid_sum_last(dq, red.partials, &red.tickets[2], blk, lds, &dAd)) {
        cg->dAd = dAd;
        if (fabs(dAd) < 1e-30) {  // breakdown: the reference reports max_iter (:127, 155-157)
            cg->done = 1;
            cg->iters = cg->max_iter;
            cg->res = sqrt(cg->rr) / cg->bnorm;
        } else {
            cg->alpha = cg->rr / dAd;
        }
    }
}

void launch_cg_spmv(hipStream_t s, const Geo &g, const double *d, double *Ad, CgState *cg,
                    RedScratch red)
{
    hipLaunchKernelGGL(k_cg_spmv, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, d, Ad, cg, red);
}

// x += alpha d, r -= alpha Ad, rr_new = r.r; convergence test and beta (:130-152)
__global__ void __launch_bounds__(kBlock) k_cg_update(uint32_t n, double *x, double *r,
                                                      const double *d, const double *Ad,
                                                      CgState *cg, RedScratch red)
{
    __shared__ double lds[kBlock / kWave];
    if (cg->done) return;
    const double alpha = cg->alpha;
    const uint32_t blk = xcd_block();
    double rr = 0.0;
    uint32_t sb, se;
    block_span(n, blk, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        x[c] += alpha * d[c];
        const double rv = r[c] - alpha * Ad[c];
        r[c] = rv;
        rr += rv * rv;
    }
    double rr_new;
    if (grid_sum_last(rr, red.partials, &red.tickets[3], blk, lds, &rr_new)) {
        const double res = sqrt(rr_new) / cg->bnorm;
        const int it = cg->iter + 1;
        if (res < cg->tol) {
            cg->done = 1;
            cg->iters = it;
            cg->res = res;
        } else {
            cg->beta = rr_new / cg->rr;
            cg->rr = rr_new;
            cg->iter = it;
            if (it >= cg->max_iter) {
                cg->done = 1;
                cg->iters = cg->max_iter;
                cg->res = sqrt(rr_new) / cg->bnorm;
            }
        }
    }
}

void launch_cg_update(hipStream_t s, uint32_t n, double *x, double *r, const double *d,
                      const double *Ad, CgState *cg, RedScratch red)
{
    hipLaunchKernelGGL(k_cg_update, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, r, d, Ad, cg,
                       red);
}

// d = r + beta d (:147-151)
__global__ void __launch_bounds__(kBlock) k_cg_dir(uint32_t n, double *d, const double *r,
                                                   CgState *cg)
{
    if (cg->done) return;
    const double beta = cg->beta;
    uint32_t sb, se;
    block_span(n, xcd_block(), sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) d[c] = r[c] + beta * d[c];
}

void launch_cg_dir(hipStream_t s, uint32_t n, double *d, const double *r, CgState *cg)
{
    hipLaunchKernelGGL(k_cg_dir, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d, r, cg);
}

void launch_cg_iteration(hipStream_t s, const Geo &g, double *x, double *r, double *d,
                         double *Ad, CgState *cg, RedScratch red)
{
    launch_cg_spmv(s, g, d, Ad, cg, red);
    launch_cg_update(s, g.nc, x, r, d, Ad, cg, red);
    launch_cg_dir(s, g.nc, d, r, cg);
}

}  // namespace ffr
