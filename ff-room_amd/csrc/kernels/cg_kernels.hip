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

__device__ __forceinline__ double *slot(RedScratch red, int k) { return red.partials + k * kMaxBlocksCap; }

// x = 0, r = d = b (SOLID rows 0), partials of r.r (PoissonSolver.cpp:96-117)
__global__ void __launch_bounds__(kBlock) k_cg_init(Geo g, const double *b, double *x, double *r,
                                                    double *d, RedScratch red)
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
    block_partial(rr, slot(red, SLOT_INIT), blk, lds);
}

// ||b||, early exits of :113-117 and max_iter <= 0 (one block)
__global__ void __launch_bounds__(kBlock) k_cg_start(CgState *cg, RedScratch red, uint32_t G,
                                                     int max_iter, double tol)
{
    __shared__ double lds[kBlock / kWave];
    const double total = sum_partials(slot(red, SLOT_INIT), G, lds);
    if (threadIdx.x != 0) return;
    const double bn = sqrt(total);
    cg->max_iter = max_iter;
    cg->tol = tol;
    cg->rr[0] = total;
    cg->bnorm = bn;
    if (bn < 1e-15) {
        cg->done = 1;
        cg->iters = 0;
        cg->res = 0.0;
    } else if (max_iter <= 0) {
        cg->done = 1;
        cg->iters = max_iter;
        cg->res = sqrt(total) / bn;
    } else {
        cg->done = 0;
        cg->iters = 0;
        cg->res = 1.0;
    }
}

void launch_cg_init(hipStream_t s, const Geo &g, const double *b, double *x, double *r,
                    double *d, CgState *cg, RedScratch red, int max_iter, double tol)
{
    const unsigned G = grid_for(g.nc);
    hipLaunchKernelGGL(k_cg_init, dim3(G), dim3(kBlock), 0, s, g, b, x, r, d, red);
    hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(kBlock), 0, s, cg, red, G, max_iter, tol);
}

// Ad = A d and partials of d.Ad (:120-126)
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
    block_partial(dq, slot(red, SLOT_SPMV), blk, lds);
}

void launch_cg_spmv(hipStream_t s, const Geo &g, const double *d, double *Ad, CgState *cg,
                    RedScratch red)
{
    hipLaunchKernelGGL(k_cg_spmv, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, d, Ad, cg, red);
}

// alpha = rr / d.Ad (stop when |d.Ad| < 1e-30, :127-129); x += alpha d,
// r -= alpha Ad and partials of r.r (:130-138)
__global__ void __launch_bounds__(kBlock) k_cg_update(uint32_t n, double *x, double *r,
                                                      const double *d, const double *Ad,
                                                      CgState *cg, RedScratch red, int it)
{
    __shared__ double lds[kBlock / kWave];
    if (cg->done) return;
    const uint32_t blk = xcd_block();
    const double dAd = sum_partials(slot(red, SLOT_SPMV), gridDim.x, lds);
    const double rr_old = cg->rr[it & 1];
    if (fabs(dAd) < 1e-30) {  // breakdown: the reference reports max_iter (:127, 155-157)
        if (blk == 0 && threadIdx.x == 0) {
            cg->done = 1;
            cg->iters = cg->max_iter;
            cg->res = sqrt(rr_old) / cg->bnorm;
        }
        return;
    }
    const double alpha = rr_old / dAd;
    double rr = 0.0;
    uint32_t sb, se;
    block_span(n, blk, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        x[c] += alpha * d[c];
        const double rv = r[c] - alpha * Ad[c];
        r[c] = rv;
        rr += rv * rv;
    }
    block_partial(rr, slot(red, SLOT_UPDATE), blk, lds);
}

void launch_cg_update(hipStream_t s, uint32_t n, double *x, double *r, const double *d,
                      const double *Ad, CgState *cg, RedScratch red, int it)
{
    hipLaunchKernelGGL(k_cg_update, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, r, d, Ad, cg,
                       red, it);
}

// res = sqrt(rr_new)/||b|| and the stopping test (:140-145); beta, d = r + beta d (:147-152)
__global__ void __launch_bounds__(kBlock) k_cg_dir(uint32_t n, double *d, const double *r,
                                                   CgState *cg, RedScratch red, int it)
{
    __shared__ double lds[kBlock / kWave];
    if (cg->done) return;
    const uint32_t blk = xcd_block();
    const double rr_new = sum_partials(slot(red, SLOT_UPDATE), gridDim.x, lds);
    const double res = sqrt(rr_new) / cg->bnorm;
    const bool lead = blk == 0 && threadIdx.x == 0;
    if (res < cg->tol) {
        if (lead) {
            cg->done = 1;
            cg->iters = it + 1;
            cg->res = res;
        }
        return;
    }
    const double beta = rr_new / cg->rr[it & 1];
    uint32_t sb, se;
    block_span(n, blk, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) d[c] = r[c] + beta * d[c];
    if (lead) {
        cg->rr[(it + 1) & 1] = rr_new;
        if (it + 1 >= cg->max_iter) {
            cg->done = 1;
            cg->iters = cg->max_iter;
            cg->res = sqrt(rr_new) / cg->bnorm;
        }
    }
}

void launch_cg_dir(hipStream_t s, uint32_t n, double *d, const double *r, CgState *cg,
                   RedScratch red, int it)
{
    hipLaunchKernelGGL(k_cg_dir, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d, r, cg, red, it);
}

}  // namespace ffr
