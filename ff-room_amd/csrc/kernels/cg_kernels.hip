// cg_kernels.hip - the reference pressure solve (PoissonSolver.cpp:22-158) in
// FP64, device-resident: every scalar of the CG recurrence (rr, alpha,
// beta, the stopping test) is computed by the last block of a deterministic
// two-pass reduction, so the host only launches and never waits inside the
// iteration. A kernel that finds cg->done set returns at once, which lets
// the host enqueue iterations ahead of the convergence test.
#include "../ffr_kernels.hpp"

#pragma clang fp contract(off)

namespace ffr {

// Row c of A (PoissonSolver.cpp:52-87) from the cell's neighbour code:
// faces in the reference order -x,+x,-y,+y,-z,+z; an unknown neighbour
// couples with -scale, an OUTFLOW neighbour adds 2*scale to the diagonal,
// walls and SOLID add nothing. UNI: every face of an axis has the same
// scale 1/h^2 (uniform spacing), so no index decode is needed.
template <bool UNI>
__device__ __forceinline__ double lap_row(const Geo &g, const double *__restrict__ x, uint32_t c,
                                          uint16_t m)
{
    const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz;
    double sxl, sxr, syl, syr, szl, szr;
    if (UNI) {
        sxl = sxr = g.sxd;
        syl = syr = g.syd;
        szl = szr = g.szd;
    } else {
        const uint32_t i = fdiv(c, g.f_nynz), rem = c - i * si;
        const uint32_t j = fdiv(rem, g.f_nz), k = rem - j * sj;
        const double *L = g.lap;
        sxl = L[i]; sxr = L[g.nx + i];
        syl = L[2 * g.nx + j]; syr = L[2 * g.nx + g.ny + j];
        szl = L[2 * g.nx + 2 * g.ny + k]; szr = L[2 * g.nx + 2 * g.ny + g.nz + k];
    }
    double diag = 0.0, off = 0.0;
#define FFR_FACE(q, nc, sc)                                                   \
    if (m & (NB_UNK0 << q)) { diag += (sc); off += (sc) * x[nc]; }           \
    else if (m & (NB_OUT0 << q)) diag += 2.0 * (sc);
    FFR_FACE(0, c - si, sxl)
    FFR_FACE(1, c + si, sxr)
    FFR_FACE(2, c - sj, syl)
    FFR_FACE(3, c + sj, syr)
    FFR_FACE(4, c - 1, szl)
    FFR_FACE(5, c + 1, szr)
#undef FFR_FACE
    if (diag < 1e-30) diag = 1.0;
    return diag * x[c] - off;
}

template <bool UNI>
__global__ void __launch_bounds__(kBlock) k_laplacian(Geo g, const double *__restrict__ x,
                                                      double *__restrict__ Ax)
{
    uint32_t sb, se;
    block_span(g.nc, xcd_block(), sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        const uint16_t m = g.nbm[c];
        Ax[c] = (m & NB_SELF_SOLID) ? x[c] : lap_row<UNI>(g, x, c, m);
    }
}

void launch_laplacian(hipStream_t s, const Geo &g, const double *x, double *Ax)
{
    if (g.uniform)
        FFR_LAUNCH(k_laplacian<true>, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, x, Ax);
    else
        FFR_LAUNCH(k_laplacian<false>, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, x, Ax);
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
    cg->started = 0;
}

void launch_cg_init(hipStream_t s, const Geo &g, const double *b, double *x, double *r,
                    double *d, CgState *cg, RedScratch red, int max_iter, double tol)
{
    const unsigned G = grid_for(g.nc);
    FFR_LAUNCH(k_cg_init, dim3(G), dim3(kBlock), 0, s, g, b, x, r, d, red);
    FFR_LAUNCH(k_cg_start, dim3(1), dim3(kBlock), 0, s, cg, red, G, max_iter, tol);
}

// Ad = A d and partials of d.Ad (:120-126). Each thread takes two cells (of
// the chunks xcd_chunks gives its block) and issues all their loads (code, d and the six
// neighbours, through buffer resources: out of range reads give 0) before
// using any; the row is lap_row's, face by face in the same order, with the
// values of faces that do not couple dropped by the code.
template <bool UNI>
__global__ void __launch_bounds__(kBlock) k_cg_spmv(Geo g, const double *__restrict__ d,
                                                    double *__restrict__ Ad, CgState *cg,
                                                    RedScratch red)
{
    constexpr int U = 2;
    __shared__ double lds[kBlock / kWave];
    if (cg->done) return;
    const uint32_t blk = xcd_block();
    const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz;
    const auto bd = buf_rsrc(d, 8u * g.nc);
    double dq = 0.0;
    uint32_t ch0, cst, cend;
    xcd_chunks(g.nc, blk, gridDim.x, ch0, cst, cend);
    for (uint32_t ch = ch0; ch < cend; ch += U * cst) {
        uint16_t m[U];
        double dc[U], v[U][6];
        uint32_t cc[U];
        bool in[U];
#pragma unroll
        for (int t = 0; t < U; t++) {
            const uint32_t ct = (ch + t * cst) * kBlock + threadIdx.x;
            in[t] = ch + t * cst < cend && ct < g.nc;
            const uint32_t c = in[t] ? ct : 0u;
            cc[t] = c;
            m[t] = g.nbm[c];
            dc[t] = d[c];
            const uint32_t q[6] = {c - si, c + si, c - sj, c + sj, c - 1u, c + 1u};
#pragma unroll
            for (int f = 0; f < 6; f++) v[t][f] = buf_ldd(bd, q[f]);
        }
#pragma unroll
        for (int t = 0; t < U; t++) {
            if (!in[t]) continue;
            const uint32_t c = cc[t];
            double sc[6];
            if (UNI) {
                sc[0] = sc[1] = g.sxd;
                sc[2] = sc[3] = g.syd;
                sc[4] = sc[5] = g.szd;
            } else {
                const uint32_t i = fdiv(c, g.f_nynz), rem = c - i * si;
                const uint32_t j = fdiv(rem, g.f_nz), k = rem - j * sj;
                const double *L = g.lap;
                sc[0] = L[i]; sc[1] = L[g.nx + i];
                sc[2] = L[2 * g.nx + j]; sc[3] = L[2 * g.nx + g.ny + j];
                sc[4] = L[2 * g.nx + 2 * g.ny + k]; sc[5] = L[2 * g.nx + 2 * g.ny + g.nz + k];
            }
            double a;
            if (m[t] & NB_SELF_SOLID) {
                a = dc[t];
            } else {
                double diag = 0.0, off = 0.0;
#pragma unroll
                for (int q = 0; q < 6; q++) {
                    if (m[t] & (NB_UNK0 << q)) { diag += sc[q]; off += sc[q] * v[t][q]; }
                    else if (m[t] & (NB_OUT0 << q)) diag += 2.0 * sc[q];
                }
                if (diag < 1e-30) diag = 1.0;
                a = diag * dc[t] - off;
            }
            Ad[c] = a;
            dq += dc[t] * a;
        }
    }
    // no other block of this kernel reads the counter
    if (blk == 0 && threadIdx.x == 0) cg->started += 1;
    block_partial(dq, slot(red, SLOT_SPMV), blk, lds);
}

void launch_cg_spmv(hipStream_t s, const Geo &g, const double *d, double *Ad, CgState *cg,
                    RedScratch red)
{
    if (g.uniform)
        FFR_LAUNCH(k_cg_spmv<true>, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, d, Ad, cg, red);
    else
        FFR_LAUNCH(k_cg_spmv<false>, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, d, Ad, cg, red);
}

// alpha = rr / d.Ad (stop when |d.Ad| < 1e-30, :127-129); x += alpha d,
// r -= alpha Ad and partials of r.r (:130-138)
__global__ void __launch_bounds__(kBlock) k_cg_update(uint32_t n, double *__restrict__ x,
                                                      double *__restrict__ r,
                                                      const double *__restrict__ d,
                                                      const double *__restrict__ Ad,
                                                      CgState *cg, RedScratch red)
{
    __shared__ double lds[kBlock / kWave];
    if (cg->done) return;
    const int it = cg->started - 1;  // this iteration's index
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
                      const double *Ad, CgState *cg, RedScratch red)
{
    FFR_LAUNCH(k_cg_update, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, r, d, Ad, cg, red);
}

// res = sqrt(rr_new)/||b|| and the stopping test (:140-145); beta, d = r + beta d (:147-152)
__global__ void __launch_bounds__(kBlock) k_cg_dir(uint32_t n, double *__restrict__ d,
                                                   const double *__restrict__ r,
                                                   CgState *cg, RedScratch red)
{
    __shared__ double lds[kBlock / kWave];
    if (cg->done) return;
    const int it = cg->started - 1;
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
                   RedScratch red)
{
    FFR_LAUNCH(k_cg_dir, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d, r, cg, red);
}

}  // namespace ffr
