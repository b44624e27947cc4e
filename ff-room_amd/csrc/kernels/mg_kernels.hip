// mg_kernels.hip - kernels of the multigrid-preconditioned pressure solve
// (see ffr_mg.hpp for the method). Level-0 vectors are FP32; coarser levels
// store z and r as TC (float or bfloat16) and compute in FP32. Dot products
// accumulate in FP64 and are reduced deterministically (block partials, fixed
// order). Every kernel of the inner iteration returns at once when the
// solve or the inner iteration has stopped, so the host may enqueue ahead.
#include "../ffr_kernels.hpp"
#include "../ffr_mg.hpp"
#include "../ffr_mg_launch.hpp"

namespace ffr {

__device__ __forceinline__ float ldv(const float *p, uint32_t i) { return p[i]; }
__device__ __forceinline__ float ldv(const __hip_bfloat16 *p, uint32_t i) { return __bfloat162float(p[i]); }
__device__ __forceinline__ void stv(float *p, uint32_t i, float v) { p[i] = v; }
__device__ __forceinline__ void stv(__hip_bfloat16 *p, uint32_t i, float v) { p[i] = __float2bfloat16(v); }

__device__ __forceinline__ bool unknown0(uint8_t f) { return f == FLUID || f == INFLOW; }
__device__ __forceinline__ double *pslot(RedScratch red, int k) { return red.partials + k * kMaxBlocksCap; }

__device__ __forceinline__ bool stopped(const MgState *m) { return m->solve_done || m->inner_done; }

__device__ __forceinline__ void decode(uint32_t c, const FastDiv &fnynz, const FastDiv &fnz, int ny,
                                       int nz, int &i, int &j, int &k)
{
    const uint32_t ii = fdiv(c, fnynz), rem = c - ii * (uint32_t)(ny * nz);
    const uint32_t jj = fdiv(rem, fnz);
    i = ii;
    j = jj;
    k = rem - jj * nz;
}

// ---- level-0 operator in FP32 from the neighbour code (uniform spacing):
// the reference row (PoissonSolver.cpp:52-87) of an unknown cell.
__device__ __forceinline__ float diag0m(const Geo &g, uint16_t m)
{
    const float cx = (float)(__popc(m & (3u << 1)) + 2 * __popc(m & (3u << 7)));
    const float cy = (float)(__popc(m & (3u << 3)) + 2 * __popc(m & (3u << 9)));
    const float cz = (float)(__popc(m & (3u << 5)) + 2 * __popc(m & (3u << 11)));
    const float d = g.sx * cx + g.sy * cy + g.sz * cz;
    return d < 1e-30f ? 1.f : d;
}

template <typename T>
__device__ __forceinline__ float off0m(const Geo &g, const T *__restrict__ x, uint32_t c, uint16_t m)
{
    const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz;
    float ox = 0.f, oy = 0.f, oz = 0.f;
    if (m & (NB_UNK0 << 0)) ox += ldv(x, c - si);
    if (m & (NB_UNK0 << 1)) ox += ldv(x, c + si);
    if (m & (NB_UNK0 << 2)) oy += ldv(x, c - sj);
    if (m & (NB_UNK0 << 3)) oy += ldv(x, c + sj);
    if (m & (NB_UNK0 << 4)) oz += ldv(x, c - 1);
    if (m & (NB_UNK0 << 5)) oz += ldv(x, c + 1);
    return g.sx * ox + g.sy * oy + g.sz * oz;
}

// stored operator of a coarse level: returns the neighbour sum
template <typename T>
__device__ __forceinline__ float rowc(const MgLevel &L, const T *x, int i, int j, int k, uint32_t c)
{
    const uint32_t si = (uint32_t)L.ny * L.nz, sj = L.nz;
    float of = 0.f;
    if (i > 0) of += L.cx[c] * ldv(x, c - si);
    if (i < L.nx - 1) of += L.cx[c + si] * ldv(x, c + si);
    if (j > 0) of += L.cy[c] * ldv(x, c - sj);
    if (j < L.ny - 1) of += L.cy[c + sj] * ldv(x, c + sj);
    if (k > 0) of += L.cz[c] * ldv(x, c - 1);
    if (k < L.nz - 1) of += L.cz[c + 1] * ldv(x, c + 1);
    return of;
}

// ---- hierarchy construction: A_{l+1} = 1/2 P^T A_l P ----------------------
// coupling of a fine cell to its lower neighbour along `axis` (0 unless both unknown)
__device__ __forceinline__ float coupling0(const Geo &g, int axis, int i, int j, int k)
{
    const uint16_t m = g.nbm[cidx(g, i, j, k)];
    if (!(m & NB_SELF) || !(m & (NB_UNK0 << (2 * axis)))) return 0.f;
    return axis == 0 ? g.sx : (axis == 1 ? g.sy : g.sz);
}
__device__ __forceinline__ float couplingc(const MgLevel &F, int axis, int i, int j, int k)
{
    const uint32_t c = ((uint32_t)i * F.ny + j) * F.nz + k;
    return axis == 0 ? F.cx[c] : (axis == 1 ? F.cy[c] : F.cz[c]);
}

template <bool FINE0>
__global__ void __launch_bounds__(kBlock) k_mg_build(Geo g, MgLevel F, MgLevel C)
{
    uint32_t sb, se;
    block_span(C.n, xcd_block(), sb, se);
    for (uint32_t I = sb + threadIdx.x; I < se; I += kBlock) {
        int ic, jc, kc;
        decode(I, C.f_nynz, C.f_nz, C.ny, C.nz, ic, jc, kc);
        const int i0 = 2 * ic, i1 = min(2 * ic + 1, F.nx - 1);
        const int j0 = 2 * jc, j1 = min(2 * jc + 1, F.ny - 1);
        const int k0 = 2 * kc, k1 = min(2 * kc + 1, F.nz - 1);
        float dsum = 0.f, inner = 0.f, cxs = 0.f, cys = 0.f, czs = 0.f;
        bool any = false;
        for (int a = i0; a <= i1; a++)
            for (int b = j0; b <= j1; b++)
                for (int d = k0; d <= k1; d++) {
                    float dg;
                    bool act;
                    if (FINE0) {
                        const uint16_t m = g.nbm[cidx(g, a, b, d)];
                        act = (m & NB_SELF) != 0;
                        dg = act ? diag0m(g, m) : 0.f;
                    } else {
                        const uint32_t c = ((uint32_t)a * F.ny + b) * F.nz + d;
                        act = F.active[c] != 0;
                        dg = act ? F.diag[c] : 0.f;
                    }
                    if (!act) continue;
                    any = true;
                    dsum += dg;
                    auto cp = [&](int axis, int x, int y, int z) {
                        return FINE0 ? coupling0(g, axis, x, y, z) : couplingc(F, axis, x, y, z);
                    };
                    // couplings inside the aggregate (counted from the upper cell)
                    if (a == i1 && i1 > i0) inner += cp(0, a, b, d);
                    if (b == j1 && j1 > j0) inner += cp(1, a, b, d);
                    if (d == k1 && k1 > k0) inner += cp(2, a, b, d);
                    // couplings across the aggregate's lower faces
                    if (a == i0 && ic > 0) cxs += cp(0, a, b, d);
                    if (b == j0 && jc > 0) cys += cp(1, a, b, d);
                    if (d == k0 && kc > 0) czs += cp(2, a, b, d);
                }
        if (any) {
            float dg = 0.5f * (dsum - 2.f * inner);
            if (!(dg > 1e-30f)) dg = 1.f;  // a pocket with no Dirichlet cell: keep finite
            C.active[I] = 1;
            C.diag[I] = dg;
            C.rdiag[I] = 1.f / dg;
            C.cx[I] = 0.5f * cxs;
            C.cy[I] = 0.5f * cys;
            C.cz[I] = 0.5f * czs;
        } else {
            C.active[I] = 0;
            C.diag[I] = 1.f;
            C.rdiag[I] = 1.f;
            C.cx[I] = C.cy[I] = C.cz[I] = 0.f;
        }
    }
}

// ---- red-black Gauss-Seidel pass of one colour ------------------------------
// z := (r + sum coupling * z_nb) / diag on unknowns of `color`; with
// zero_start the pass begins from z = 0 and also clears the other colour.
template <typename T, bool L0>
__global__ void __launch_bounds__(kBlock) k_mg_smooth(Geo g, MgLevel L, T *__restrict__ z,
                                                      const T *__restrict__ r, int color,
                                                      int zero_start, const MgState *m)
{
    if (stopped(m)) return;
    uint32_t sb, se;
    block_span(L.n, xcd_block(), sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        if (L0) {
            const uint16_t nb = g.nbm[c];
            const int col = (nb & NB_COLOR) ? 1 : 0;
            if (col != color) {
                if (zero_start) stv(z, c, 0.f);
                continue;
            }
            if (!(nb & NB_SELF)) {
                stv(z, c, 0.f);
                continue;
            }
            const float of = zero_start ? 0.f : off0m(g, z, c, nb);
            stv(z, c, (ldv(r, c) + of) / diag0m(g, nb));
            continue;
        }
        int i, j, k;
        decode(c, L.f_nynz, L.f_nz, L.ny, L.nz, i, j, k);
        if (((i + j + k) & 1) != color) {
            if (zero_start) stv(z, c, 0.f);
            continue;
        }
        if (!L.active[c]) {
            stv(z, c, 0.f);
            continue;
        }
        const float of = zero_start ? 0.f : rowc<T>(L, z, i, j, k, c);
        stv(z, c, (ldv(r, c) + of) * L.rdiag[c]);
    }
}

// ---- fused residual + restriction: rc = P^T (r - A z) ------------------------
template <typename TF, typename TC, bool L0>
__global__ void __launch_bounds__(kBlock) k_mg_restrict(Geo g, MgLevel F, MgLevel C, const TF *z,
                                                        const TF *r, TC *rc, const MgState *m)
{
    if (stopped(m)) return;
    uint32_t sb, se;
    block_span(C.n, xcd_block(), sb, se);
    for (uint32_t I = sb + threadIdx.x; I < se; I += kBlock) {
        int ic, jc, kc;
        decode(I, C.f_nynz, C.f_nz, C.ny, C.nz, ic, jc, kc);
        const int i1 = min(2 * ic + 1, F.nx - 1), j1 = min(2 * jc + 1, F.ny - 1),
                  k1 = min(2 * kc + 1, F.nz - 1);
        float sum = 0.f;
        for (int a = 2 * ic; a <= i1; a++)
            for (int b = 2 * jc; b <= j1; b++)
                for (int d = 2 * kc; d <= k1; d++) {
                    const uint32_t c = ((uint32_t)a * F.ny + b) * F.nz + d;
                    if (L0) {
                        const uint16_t nb = g.nbm[c];
                        if (!(nb & NB_SELF)) continue;
                        sum += ldv(r, c) - (diag0m(g, nb) * ldv(z, c) - off0m(g, z, c, nb));
                    } else {
                        if (!F.active[c]) continue;
                        const float of = rowc<TF>(F, z, a, b, d, c);
                        sum += ldv(r, c) - (F.diag[c] * ldv(z, c) - of);
                    }
                }
        stv(rc, I, sum);
    }
}

// ---- prolongation + correction: z += P zc ------------------------------------
template <typename TF, typename TC, bool L0>
__global__ void __launch_bounds__(kBlock) k_mg_prolong(Geo g, MgLevel F, MgLevel C, TF *z,
                                                       const TC *zc, const MgState *m)
{
    if (stopped(m)) return;
    uint32_t sb, se;
    block_span(F.n, xcd_block(), sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        const bool act = L0 ? (g.nbm[c] & NB_SELF) != 0 : F.active[c] != 0;
        if (!act) continue;
        int i, j, k;
        decode(c, F.f_nynz, F.f_nz, F.ny, F.nz, i, j, k);
        const uint32_t P = ((uint32_t)(i >> 1) * C.ny + (j >> 1)) * C.nz + (k >> 1);
        stv(z, c, ldv(z, c) + ldv(zc, P));
    }
}

// ---- coarsest level: symmetric red-black sweeps from z = 0 (one block) -----
template <typename T, bool L0>
__global__ void __launch_bounds__(1024) k_mg_coarse(Geo g, MgLevel L, T *z, const T *r, int sweeps,
                                                    const MgState *m)
{
    if (stopped(m)) return;
    for (uint32_t c = threadIdx.x; c < L.n; c += blockDim.x) stv(z, c, 0.f);
    __syncthreads();
    // (red, black)^sweeps then (black, red)^sweeps: a symmetric operator
    for (int pass = 0; pass < 4 * sweeps; pass++) {
        const int half = pass / (2 * sweeps);
        const int color = ((pass & 1) ^ half) & 1;  // R,B,...  then B,R,...
        for (uint32_t c = threadIdx.x; c < L.n; c += blockDim.x) {
            int i, j, k;
            decode(c, L.f_nynz, L.f_nz, L.ny, L.nz, i, j, k);
            if (((i + j + k) & 1) != color) continue;
            if (L0) {
                const uint16_t nb = g.nbm[c];
                if (!(nb & NB_SELF)) continue;
                stv(z, c, (ldv(r, c) + off0m(g, z, c, nb)) / diag0m(g, nb));
            } else {
                if (!L.active[c]) continue;
                stv(z, c, (ldv(r, c) + rowc<T>(L, z, i, j, k, c)) * L.rdiag[c]);
            }
        }
        __syncthreads();
    }
}

// ---- inner PCG on level 0 (FP32 vectors, FP64 dots) -------------------------
// q = A p, partials of p.q
__global__ void __launch_bounds__(kBlock) k_pcg_spmv(Geo g, const float *p, float *q,
                                                     const MgState *m, RedScratch red)
{
    __shared__ double lds[kBlock / kWave];
    if (stopped(m)) return;
    const uint32_t blk = xcd_block();
    double acc = 0.0;
    uint32_t sb, se;
    block_span(g.nc, blk, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        float qv = 0.f;
        const uint16_t nb = g.nbm[c];
        if (nb & NB_SELF) {
            const float pc = p[c];
            qv = diag0m(g, nb) * pc - off0m(g, p, c, nb);
            acc += (double)pc * (double)qv;
        }
        q[c] = qv;
    }
    block_partial(acc, pslot(red, SLOT_SPMV), blk, lds);
}

// alpha = rz / p.q; e += alpha p; r -= alpha q; partials of r.r
__global__ void __launch_bounds__(kBlock) k_pcg_update(uint32_t n, float *e, float *r,
                                                       const float *p, const float *q, MgState *m,
                                                       RedScratch red, int it)
{
    __shared__ double lds[kBlock / kWave];
    if (stopped(m)) return;
    const uint32_t blk = xcd_block();
    const double pq = sum_partials(pslot(red, SLOT_SPMV), gridDim.x, lds);
    if (!(pq > 0.0)) {  // breakdown or a non-positive operator: end the inner solve
        if (blk == 0 && threadIdx.x == 0) m->inner_done = 1;
        return;
    }
    const float alpha = (float)(m->rz[it & 1] / pq);
    double acc = 0.0;
    uint32_t sb, se;
    block_span(n, blk, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        e[c] += alpha * p[c];
        const float rv = r[c] - alpha * q[c];
        r[c] = rv;
        acc += (double)rv * (double)rv;
    }
    block_partial(acc, pslot(red, SLOT_UPDATE), blk, lds);
}

// stopping test of the inner solve (one block)
__global__ void __launch_bounds__(kBlock) k_pcg_check(MgState *m, RedScratch red, uint32_t G, int it)
{
    __shared__ double lds[kBlock / kWave];
    if (stopped(m)) return;
    const double rr = sum_partials(pslot(red, SLOT_UPDATE), G, lds);
    if (threadIdx.x != 0) return;
    m->inner_it = it + 1;
    m->total_iters += 1;
    if (sqrt(rr) <= m->inner_tol * m->rnorm0 || it + 1 >= m->max_inner || !(rr == rr))
        m->inner_done = 1;
}

// partials of r.z (after the V-cycle)
__global__ void __launch_bounds__(kBlock) k_pcg_dot_rz(uint32_t n, const float *r, const float *z,
                                                       const MgState *m, RedScratch red)
{
    __shared__ double lds[kBlock / kWave];
    if (stopped(m)) return;
    const uint32_t blk = xcd_block();
    double acc = 0.0;
    uint32_t sb, se;
    block_span(n, blk, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) acc += (double)r[c] * (double)z[c];
    block_partial(acc, pslot(red, SLOT_MG_A), blk, lds);
}

// first direction: p = z, rz[0] = r.z, e = 0
__global__ void __launch_bounds__(kBlock) k_pcg_start(uint32_t n, const float *z, float *p, float *e,
                                                      MgState *m, RedScratch red)
{
    __shared__ double lds[kBlock / kWave];
    if (stopped(m)) return;
    const uint32_t blk = xcd_block();
    const double rz = sum_partials(pslot(red, SLOT_MG_A), gridDim.x, lds);
    uint32_t sb, se;
    block_span(n, blk, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        p[c] = z[c];
        e[c] = 0.f;
    }
    if (blk == 0 && threadIdx.x == 0) m->rz[0] = rz;
}

// beta = rz_new / rz; p = z + beta p
__global__ void __launch_bounds__(kBlock) k_pcg_dir(uint32_t n, const float *z, float *p, MgState *m,
                                                    RedScratch red, int it)
{
    __shared__ double lds[kBlock / kWave];
    if (stopped(m)) return;
    const uint32_t blk = xcd_block();
    const double rz_new = sum_partials(pslot(red, SLOT_MG_A), gridDim.x, lds);
    const float beta = (float)(rz_new / m->rz[it & 1]);
    uint32_t sb, se;
    block_span(n, blk, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) p[c] = z[c] + beta * p[c];
    if (blk == 0 && threadIdx.x == 0) m->rz[(it + 1) & 1] = rz_new;
}

// ---- outer FP64 defect correction -------------------------------------------
// x = 0, e = 0, partials of ||b||^2 over non-SOLID cells (PoissonSolver.cpp:101-113)
__global__ void __launch_bounds__(kBlock) k_dc_init(Geo g, const double *b, double *x, float *e,
                                                    RedScratch red)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t blk = xcd_block();
    double acc = 0.0;
    uint32_t sb, se;
    block_span(g.nc, blk, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        x[c] = 0.0;
        e[c] = 0.f;
        if (g.flags[c] != SOLID) acc += b[c] * b[c];
    }
    block_partial(acc, pslot(red, SLOT_INIT), blk, lds);
}

__global__ void __launch_bounds__(kBlock) k_dc_start(MgState *m, RedScratch red, uint32_t G,
                                                     double outer_tol, int max_outer, int max_inner,
                                                     double inner_tol)
{
    __shared__ double lds[kBlock / kWave];
    const double bb = sum_partials(pslot(red, SLOT_INIT), G, lds);
    if (threadIdx.x != 0) return;
    m->bnorm = sqrt(bb);
    m->outer_tol = outer_tol;
    m->inner_tol = inner_tol;
    m->max_outer = max_outer;
    m->max_inner = max_inner;
    m->outer = 0;
    m->total_iters = 0;
    m->inner_it = 0;
    m->inner_done = 1;
    m->solve_done = m->bnorm < 1e-15 ? 1 : 0;
    m->res = 0.0;
}

// FP64 reference row (PoissonSolver.cpp:52-87) on a uniform grid from the
// neighbour code; neighbour values come from `val` (x + e for the outer
// residual, x for the OUTFLOW back substitution).
template <typename F>
__device__ __forceinline__ void row64(const Geo &g, uint32_t c, uint16_t m, F val, double &diag,
                                      double &off)
{
    const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz;
    double dg = 0.0, of = 0.0;
#define FFR_FACE(q, nc, sc)                                                   \
    if (m & (NB_UNK0 << q)) { dg += (sc); of += (sc) * val(nc); }            \
    else if (m & (NB_OUT0 << q)) dg += 2.0 * (sc);
    FFR_FACE(0, c - si, g.sxd)
    FFR_FACE(1, c + si, g.sxd)
    FFR_FACE(2, c - sj, g.syd)
    FFR_FACE(3, c + sj, g.syd)
    FFR_FACE(4, c - 1, g.szd)
    FFR_FACE(5, c + 1, g.szd)
#undef FFR_FACE
    diag = dg < 1e-30 ? 1.0 : dg;
    off = of;
}

// x_out = x_in + e on unknowns; r = b - A x_out (FP64); r32 = (float) r;
// partials of ||r||^2
__global__ void __launch_bounds__(kBlock) k_dc_residual(Geo g, const double *__restrict__ b,
                                                        const double *__restrict__ xin,
                                                        const float *__restrict__ e,
                                                        double *__restrict__ xout,
                                                        float *__restrict__ r32, const MgState *m,
                                                        RedScratch red)
{
    __shared__ double lds[kBlock / kWave];
    if (m->solve_done) return;
    const uint32_t blk = xcd_block();
    double acc = 0.0;
    uint32_t sb, se;
    block_span(g.nc, blk, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        const uint16_t nb = g.nbm[c];
        if (!(nb & NB_SELF)) {
            xout[c] = 0.0;
            r32[c] = 0.f;
            continue;
        }
        double diag, off;
        row64(g, c, nb, [&](uint32_t q) { return xin[q] + (double)e[q]; }, diag, off);
        const double xc = xin[c] + (double)e[c];
        const double rv = b[c] - (diag * xc - off);
        xout[c] = xc;
        r32[c] = (float)rv;
        acc += rv * rv;
    }
    block_partial(acc, pslot(red, SLOT_UPDATE), blk, lds);
}

// outer stopping test: ||r|| / ||b|| < tol (one block)
__global__ void __launch_bounds__(kBlock) k_dc_check(MgState *m, RedScratch red, uint32_t G)
{
    __shared__ double lds[kBlock / kWave];
    if (m->solve_done) return;
    const double rr = sum_partials(pslot(red, SLOT_UPDATE), G, lds);
    if (threadIdx.x != 0) return;
    const double rn = sqrt(rr);
    m->res = rn / m->bnorm;
    if (m->res < m->outer_tol || m->outer >= m->max_outer || !(rr == rr)) {
        m->solve_done = 1;
        return;
    }
    m->outer += 1;
    m->rnorm0 = rn;
    m->inner_it = 0;
    m->inner_done = 0;
}

// p: unknowns from x; OUTFLOW rows solved from their neighbours (or pinned
// to 0 for the reference multigrid's convention); SOLID cells 0.
__global__ void __launch_bounds__(kBlock) k_dc_finalize(Geo g, const double *__restrict__ b,
                                                        const double *x, double *p,
                                                        int pin_outflow)
{
    uint32_t sb, se;
    block_span(g.nc, xcd_block(), sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        const uint16_t nb = g.nbm[c];
        if (nb & NB_SELF) {
            p[c] = x[c];
            continue;
        }
        if ((nb & NB_SELF_SOLID) || pin_outflow) {
            p[c] = 0.0;
            continue;
        }
        double diag, off;
        row64(g, c, nb, [&](uint32_t q) { return x[q]; }, diag, off);
        p[c] = (b[c] + off) / diag;
    }
}

// ---- launchers ------------------------------------------------------------------
void mg_launch_build(hipStream_t s, const Geo &g, const MgLevel &F, const MgLevel &C, bool fine0)
{
    if (fine0)
        FFR_LAUNCH(k_mg_build<true>, dim3(grid_for(C.n)), dim3(kBlock), 0, s, g, F, C);
    else
        FFR_LAUNCH(k_mg_build<false>, dim3(grid_for(C.n)), dim3(kBlock), 0, s, g, F, C);
}

template <typename T>
void mg_launch_smooth(hipStream_t s, const Geo &g, const MgLevel &L, bool l0, int color,
                      int zero_start, const MgState *m)
{
    T *z = (T *)L.z;
    const T *r = (const T *)L.r;
    if (l0)
        FFR_LAUNCH((k_mg_smooth<T, true>), dim3(grid_for(L.n)), dim3(kBlock), 0, s, g, L, z, r,
                           color, zero_start, m);
    else
        FFR_LAUNCH((k_mg_smooth<T, false>), dim3(grid_for(L.n)), dim3(kBlock), 0, s, g, L, z,
                           r, color, zero_start, m);
}

template <typename TF, typename TC>
void mg_launch_restrict(hipStream_t s, const Geo &g, const MgLevel &F, const MgLevel &C, bool l0,
                        const MgState *m)
{
    if (l0)
        FFR_LAUNCH((k_mg_restrict<TF, TC, true>), dim3(grid_for(C.n)), dim3(kBlock), 0, s, g, F,
                           C, (const TF *)F.z, (const TF *)F.r, (TC *)C.r, m);
    else
        FFR_LAUNCH((k_mg_restrict<TF, TC, false>), dim3(grid_for(C.n)), dim3(kBlock), 0, s, g,
                           F, C, (const TF *)F.z, (const TF *)F.r, (TC *)C.r, m);
}

template <typename TF, typename TC>
void mg_launch_prolong(hipStream_t s, const Geo &g, const MgLevel &F, const MgLevel &C, bool l0,
                       const MgState *m)
{
    if (l0)
        FFR_LAUNCH((k_mg_prolong<TF, TC, true>), dim3(grid_for(F.n)), dim3(kBlock), 0, s, g, F,
                           C, (TF *)F.z, (const TC *)C.z, m);
    else
        FFR_LAUNCH((k_mg_prolong<TF, TC, false>), dim3(grid_for(F.n)), dim3(kBlock), 0, s, g,
                           F, C, (TF *)F.z, (const TC *)C.z, m);
}

template <typename T>
void mg_launch_coarse(hipStream_t s, const Geo &g, const MgLevel &L, bool l0, int sweeps,
                      const MgState *m)
{
    if (l0)
        FFR_LAUNCH((k_mg_coarse<T, true>), dim3(1), dim3(1024), 0, s, g, L, (T *)L.z,
                           (const T *)L.r, sweeps, m);
    else
        FFR_LAUNCH((k_mg_coarse<T, false>), dim3(1), dim3(1024), 0, s, g, L, (T *)L.z,
                           (const T *)L.r, sweeps, m);
}

template void mg_launch_smooth<float>(hipStream_t, const Geo &, const MgLevel &, bool, int, int, const MgState *);
template void mg_launch_smooth<__hip_bfloat16>(hipStream_t, const Geo &, const MgLevel &, bool, int, int, const MgState *);
template void mg_launch_restrict<float, float>(hipStream_t, const Geo &, const MgLevel &, const MgLevel &, bool, const MgState *);
template void mg_launch_restrict<float, __hip_bfloat16>(hipStream_t, const Geo &, const MgLevel &, const MgLevel &, bool, const MgState *);
template void mg_launch_restrict<__hip_bfloat16, __hip_bfloat16>(hipStream_t, const Geo &, const MgLevel &, const MgLevel &, bool, const MgState *);
template void mg_launch_prolong<float, float>(hipStream_t, const Geo &, const MgLevel &, const MgLevel &, bool, const MgState *);
template void mg_launch_prolong<float, __hip_bfloat16>(hipStream_t, const Geo &, const MgLevel &, const MgLevel &, bool, const MgState *);
template void mg_launch_prolong<__hip_bfloat16, __hip_bfloat16>(hipStream_t, const Geo &, const MgLevel &, const MgLevel &, bool, const MgState *);
template void mg_launch_coarse<float>(hipStream_t, const Geo &, const MgLevel &, bool, int, const MgState *);
template void mg_launch_coarse<__hip_bfloat16>(hipStream_t, const Geo &, const MgLevel &, bool, int, const MgState *);

void pcg_launch_spmv(hipStream_t s, const Geo &g, const float *p, float *q, const MgState *m, RedScratch red)
{
    FFR_LAUNCH(k_pcg_spmv, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, p, q, m, red);
}
void pcg_launch_update(hipStream_t s, uint32_t n, float *e, float *r, const float *p, const float *q,
                       MgState *m, RedScratch red, int it)
{
    FFR_LAUNCH(k_pcg_update, dim3(grid_for(n)), dim3(kBlock), 0, s, n, e, r, p, q, m, red, it);
}
void pcg_launch_check(hipStream_t s, uint32_t n, MgState *m, RedScratch red, int it)
{
    FFR_LAUNCH(k_pcg_check, dim3(1), dim3(kBlock), 0, s, m, red, grid_for(n), it);
}
void pcg_launch_dot_rz(hipStream_t s, uint32_t n, const float *r, const float *z, const MgState *m,
                       RedScratch red)
{
    FFR_LAUNCH(k_pcg_dot_rz, dim3(grid_for(n)), dim3(kBlock), 0, s, n, r, z, m, red);
}
void pcg_launch_start(hipStream_t s, uint32_t n, const float *z, float *p, float *e, MgState *m,
                      RedScratch red)
{
    FFR_LAUNCH(k_pcg_start, dim3(grid_for(n)), dim3(kBlock), 0, s, n, z, p, e, m, red);
}
void pcg_launch_dir(hipStream_t s, uint32_t n, const float *z, float *p, MgState *m, RedScratch red, int it)
{
    FFR_LAUNCH(k_pcg_dir, dim3(grid_for(n)), dim3(kBlock), 0, s, n, z, p, m, red, it);
}
void dc_launch_init(hipStream_t s, const Geo &g, const double *b, double *x, float *e, RedScratch red)
{
    FFR_LAUNCH(k_dc_init, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, b, x, e, red);
}
void dc_launch_start(hipStream_t s, const Geo &g, MgState *m, RedScratch red, double outer_tol,
                     int max_outer, int max_inner, double inner_tol)
{
    FFR_LAUNCH(k_dc_start, dim3(1), dim3(kBlock), 0, s, m, red, grid_for(g.nc), outer_tol,
                       max_outer, max_inner, inner_tol);
}
void dc_launch_residual(hipStream_t s, const Geo &g, const double *b, const double *xin, const float *e,
                        double *xout, float *r32, const MgState *m, RedScratch red)
{
    FFR_LAUNCH(k_dc_residual, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, b, xin, e, xout, r32,
                       m, red);
}
void dc_launch_check(hipStream_t s, const Geo &g, MgState *m, RedScratch red)
{
    FFR_LAUNCH(k_dc_check, dim3(1), dim3(kBlock), 0, s, m, red, grid_for(g.nc));
}
void dc_launch_finalize(hipStream_t s, const Geo &g, const double *b, const double *x, double *p,
                        int pin_outflow)
{
    FFR_LAUNCH(k_dc_finalize, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, b, x, p, pin_outflow);
}

}  // namespace ffr
