// lbm_kernels.hip - D3Q19 lattice Boltzmann (LBMSolver.cpp) on an FP64 scene.
//
// Populations live as f[q * nc + cell] (one contiguous array per direction,
// so each direction's loads and stores are coalesced), double-buffered. A
// step is: inflow cells set to their fan's equilibrium and BGK collision on
// FLUID and OUTFLOW cells (one pass, in place); streaming as a pull into the
// other buffer; the zero-gradient copies of the outflow and opening patches,
// one launch per patch in the reference's order; cell moments; MAC faces with
// the velocity change. All arithmetic is FP64 without contraction, in the
// reference's operation order, so results equal its bits. Streaming reads a
// per-cell mask of open neighbours (built once: bit d set when the cell in
// direction d is inside the domain and not SOLID, bit 31 when the cell is
// SOLID, bit 30 when it is INFLOW) instead of 38 flag bytes, and also forms
// the cell's moments; a patch copy recomputes the moments of the cells it
// overwrites.
//
// Fused form (k_lbm_stream_collide): the next step's inflow and collision
// act on each cell's streamed values alone, so the streaming pass applies
// them at once and stores the next step's pre-streaming populations, and the
// separate relaxation pass disappears. A patch copy then copies the source
// cell's post-collision values and moments, which equal the collision of the
// copied populations when the source cell collides (FLUID, OUTFLOW) or is
// SOLID (zero populations stay zero); scenes where a patch copies from an
// INFLOW cell keep the separate passes.
//
// Streaming: the reference pushes into a zeroed buffer (LBMSolver.cpp:151-
// 190) and bounces a population that would leave the domain or enter a
// SOLID cell into the same cell's direction OPP[q]. Its OPP table (:28)
// reverses the axis directions but only one component of the twelve diagonal
// ones (7 (+1,+1,0) <-> 8 (-1,+1,0), ...), so a target may receive the
// streamed value, the bounced value, both or neither; the pull sums exactly
// those contributions.
#include "../ffr_kernels.hpp"
#include "../ffr_lbm.hpp"

#pragma clang fp contract(off)

namespace ffr {

namespace {
__device__ __forceinline__ int lbm_ex(int q)
{
    constexpr int e[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    return e[q];
}
__device__ __forceinline__ int lbm_ey(int q)
{
    constexpr int e[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
    return e[q];
}
__device__ __forceinline__ int lbm_ez(int q)
{
    constexpr int e[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    return e[q];
}
// the reference's pairing of directions for bounce-back (see above)
__device__ __forceinline__ int lbm_opp(int q)
{
    constexpr int o[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
    return o[q];
}
// the true reverse of direction q (for the upstream neighbour)
__device__ __forceinline__ int lbm_rev(int q)
{
    constexpr int r[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
    return r[q];
}
constexpr uint32_t kSolidBit = 1u << 31, kInflowBit = 1u << 30;
__device__ __forceinline__ double lbm_w(int q)
{
    return q == 0 ? 1.0 / 3.0 : (q < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
}

// equilibrium, LBMSolver.hpp:85-90 (cs2 = 1/3 as a double)
__device__ __forceinline__ double feq(int q, double rho, double ux, double uy, double uz)
{
    const double cs2 = 1.0 / 3.0;
    const double eu = lbm_ex(q) * ux + lbm_ey(q) * uy + lbm_ez(q) * uz;
    const double usq = ux * ux + uy * uy + uz * uz;
    return lbm_w(q) * rho * (1.0 + eu / cs2 + eu * eu / (2.0 * cs2 * cs2) - usq / (2.0 * cs2));
}

// density and velocity of cell c's populations (LBMSolver.cpp:114-125)
__device__ __forceinline__ void moments(const double *f, uint32_t nc, uint32_t c, double &rho,
                                        double &ux, double &uy, double &uz)
{
    double r = 0.0, x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const double fq = f[(size_t)q * nc + c];
        r += fq;
        x += fq * lbm_ex(q);
        y += fq * lbm_ey(q);
        z += fq * lbm_ez(q);
    }
    if (r > 1e-14) {
        x /= r;
        y /= r;
        z /= r;
    }
    rho = r;
    ux = x;
    uy = y;
    uz = z;
}

__device__ __forceinline__ void cell_ijk(const Geo &g, uint32_t c, int &i, int &j, int &k)
{
    const uint32_t ii = fdiv(c, g.f_nynz), rem = c - ii * (uint32_t)(g.ny * g.nz);
    const uint32_t jj = fdiv(rem, g.f_nz);
    i = ii;
    j = jj;
    k = rem - jj * g.nz;
}
}  // namespace

__global__ void __launch_bounds__(kBlock) k_lbm_rest(uint32_t nc, double *f)
{
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock)
#pragma unroll
        for (int q = 0; q < 19; q++) f[(size_t)q * nc + c] = lbm_w(q);
}

// the velocity (lattice units) of the last fan whose box holds cell c (the
// fans applied in order, LBMSolver.cpp:81-100); false when none does
__device__ __forceinline__ bool inflow_of(const Geo &g, uint32_t c, double s2l, double &vx,
                                          double &vy, double &vz)
{
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    for (int n = g.n_fans - 1; n >= 0; n--) {
        const FanDev &fn = g.fans[n];
        if (i < fn.i0 || i > fn.i1 || j < fn.j0 || j > fn.j1 || k < fn.k0 || k > fn.k1) continue;
        vx = fn.vel[0] * s2l;
        vy = fn.vel[1] * s2l;
        vz = fn.vel[2] * s2l;
        return true;
    }
    return false;
}

// Inflow cells at the equilibrium of the last fan whose box holds them (the
// fans applied in order, LBMSolver.cpp:81-100); with `collide`, BGK on FLUID
// and OUTFLOW cells (:106-144). Block 0 clears the step's diagnostics.
__global__ void __launch_bounds__(kBlock) k_lbm_relax(Geo g, double *f, double s2l, double inv_tau,
                                                       int collide, StepRec *rec)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        rec->vmax_bits = 0ull;
        rec->divmax_bits = 0ull;
        rec->nonfinite = 0;
    }
    const uint32_t nc = g.nc;
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
        const uint8_t t = g.flags[c];
        if (t == INFLOW) {
            double vx, vy, vz;
            if (inflow_of(g, c, s2l, vx, vy, vz))
#pragma unroll
                for (int q = 0; q < 19; q++) f[(size_t)q * nc + c] = feq(q, 1.0, vx, vy, vz);
            continue;
        }
        if (!collide || t == SOLID) continue;
        double rho, ux, uy, uz;
        moments(f, nc, c, rho, ux, uy, uz);
#pragma unroll
        for (int q = 0; q < 19; q++) {
            double &fq = f[(size_t)q * nc + c];
            fq += inv_tau * (feq(q, rho, ux, uy, uz) - fq);
        }
    }
}

// open-neighbour masks (see the file comment)
__global__ void __launch_bounds__(kBlock) k_lbm_mask(Geo g, uint32_t *mask)
{
    const uint32_t nc = g.nc;
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
        int i, j, k;
        cell_ijk(g, c, i, j, k);
        uint32_t m = g.flags[c] == SOLID ? kSolidBit : (g.flags[c] == INFLOW ? kInflowBit : 0u);
#pragma unroll
        for (int d = 0; d < 19; d++) {
            const int a = i + lbm_ex(d), b = j + lbm_ey(d), e = k + lbm_ez(d);
            if (a >= 0 && a < g.nx && b >= 0 && b < g.ny && e >= 0 && e < g.nz &&
                g.flags[((uint32_t)a * g.ny + b) * g.nz + e] != SOLID)
                m |= 1u << d;
        }
        mask[c] = m;
    }
}

// moments of a cell's new populations -> cell velocity [m/s] and p = rho/3
// (LBMSolver.cpp:288-326)
__device__ __forceinline__ void store_moments(uint32_t nc, uint32_t c, double r, double x, double y,
                                              double z, double l2s, double *cu, double *p)
{
    if (r > 1e-14) {
        x /= r;
        y /= r;
        z /= r;
    }
    cu[c] = x * l2s;
    cu[nc + c] = y * l2s;
    cu[2ull * nc + c] = z * l2s;
    p[c] = r * (1.0 / 3.0);
}

// streaming with bounce-back, pulled into f2 (see the file comment), and
// the moments of the result; SOLID cells get zero populations and velocity
// and keep their p
__global__ void __launch_bounds__(kBlock) k_lbm_stream(Geo g, const uint32_t *__restrict__ mask,
                                                        const double *__restrict__ f,
                                                        double *__restrict__ f2, double *cu, double *p,
                                                        double l2s)
{
    const uint32_t nc = g.nc, nynz = (uint32_t)g.ny * g.nz, nz = g.nz;
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
        const uint32_t m = mask[c];
        if (m & kSolidBit) {
#pragma unroll
            for (int q = 0; q < 19; q++) f2[(size_t)q * nc + c] = 0.0;
            cu[c] = 0.0;
            cu[nc + c] = 0.0;
            cu[2ull * nc + c] = 0.0;
            continue;
        }
        double r = 0.0, x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
        for (int q = 0; q < 19; q++) {
            double v = 0.0;
            if (m & (1u << lbm_rev(q))) {  // upstream neighbour c - e_q is open
                const int off = lbm_ex(q) * (int)nynz + lbm_ey(q) * (int)nz + lbm_ez(q);
                v += f[(size_t)q * nc + (uint32_t)((int)c - off)];
            }
            const int b = lbm_opp(q);
            if (!(m & (1u << b))) v += f[(size_t)b * nc + c];  // bounced at a wall or SOLID
            f2[(size_t)q * nc + c] = v;
            r += v;
            x += v * lbm_ex(q);
            y += v * lbm_ey(q);
            z += v * lbm_ez(q);
        }
        store_moments(nc, c, r, x, y, z, l2s, cu, p);
    }
}

// Fused form: streaming and moments as k_lbm_stream, then the next step's
// inflow equilibrium (INFLOW cells) or BGK collision (FLUID, OUTFLOW) on the
// streamed values, stored in f2; block 0 clears the step's diagnostics.
// The collision needs all 19 streamed values after the moments: kLbmLds of
// them wait in the thread's own LDS slots (no barrier: each thread reads
// back only what it wrote), the rest in registers, and the loads and the
// stores go out kLbmGroup directions at a time (scheduling barriers), so
// eight waves per SIMD fit 64 VGPRs with a few spilled values (36 B per lane,
// against 76 before). 10 x 256 x 8 B = 20 KiB per block, 160 KiB for the
// eight blocks of a CU. Measured on MI355X (mean of 20 launches): c2 370 ->
// 323 us, c3 2793 -> 2461 us; with 11..14 values in LDS (no spills, but six
// or seven waves by LDS or VGPRs) 408-417 us on c2.
constexpr int kLbmLds = 10, kLbmGroup = 4;
__global__ void __launch_bounds__(kBlock, 8) k_lbm_stream_collide(Geo g, const uint32_t *__restrict__ mask,
                                                                const double *__restrict__ f,
                                                                double *__restrict__ f2, double *cu,
                                                                double *p, double l2s, double s2l,
                                                                double inv_tau, StepRec *rec)
{
    __shared__ double sv[kLbmLds][kBlock];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        rec->vmax_bits = 0ull;
        rec->divmax_bits = 0ull;
        rec->nonfinite = 0;
    }
    const uint32_t nc = g.nc, nynz = (uint32_t)g.ny * g.nz, nz = g.nz;
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
        const uint32_t m = mask[c];
        if (m & kSolidBit) {  // streams nothing in, collides nothing
#pragma unroll
            for (int q = 0; q < 19; q++) f2[(size_t)q * nc + c] = 0.0;
            cu[c] = 0.0;
            cu[nc + c] = 0.0;
            cu[2ull * nc + c] = 0.0;
            continue;
        }
        double v[19 - kLbmLds];
        double r = 0.0, x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
        for (int q = 0; q < 19; q++) {
            double a = 0.0;
            if (m & (1u << lbm_rev(q))) {
                const int off = lbm_ex(q) * (int)nynz + lbm_ey(q) * (int)nz + lbm_ez(q);
                a += f[(size_t)q * nc + (uint32_t)((int)c - off)];
            }
            const int b = lbm_opp(q);
            if (!(m & (1u << b))) a += f[(size_t)b * nc + c];
            if (q < kLbmLds) sv[q][threadIdx.x] = a;
            else v[q - kLbmLds] = a;
            // at most a group of directions' loads in flight at once, so
            // their data and addresses fit the registers
            if (q % kLbmGroup == kLbmGroup - 1) __builtin_amdgcn_sched_barrier(0);
            r += a;
            x += a * lbm_ex(q);
            y += a * lbm_ey(q);
            z += a * lbm_ez(q);
        }
        if (r > 1e-14) {  // LBMSolver.cpp:114-125 and :288-326 divide alike
            x /= r;
            y /= r;
            z /= r;
        }
        cu[c] = x * l2s;
        cu[nc + c] = y * l2s;
        cu[2ull * nc + c] = z * l2s;
        p[c] = r * (1.0 / 3.0);
        if (m & kInflowBit) {
            double vx, vy, vz;
            if (inflow_of(g, c, s2l, vx, vy, vz)) {
#pragma unroll
                for (int q = 0; q < 19; q++) f2[(size_t)q * nc + c] = feq(q, 1.0, vx, vy, vz);
            } else {
#pragma unroll
                for (int q = 0; q < 19; q++)
                    f2[(size_t)q * nc + c] = q < kLbmLds ? sv[q][threadIdx.x] : v[q - kLbmLds];
            }
            continue;
        }
#pragma unroll
        for (int q = 0; q < 19; q++) {
            const double vq = q < kLbmLds ? sv[q][threadIdx.x] : v[q - kLbmLds];
            f2[(size_t)q * nc + c] = vq + inv_tau * (feq(q, r, x, y, z) - vq);
            if (q % kLbmGroup == kLbmGroup - 1) __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// Zero-gradient copy of one patch's OUTFLOW cells from the cell one step
// inside along the patch axis (LBMSolver.cpp:196-282); a, b the patch's two
// in-wall axes in order.
__global__ void __launch_bounds__(kBlock) k_lbm_patch(Geo g, double *f, LbmPatch pt, double *cu,
                                                       double *p, double l2s)
{
    const uint32_t na = (uint32_t)(pt.a1 - pt.a0 + 1), nb = (uint32_t)(pt.b1 - pt.b0 + 1);
    const uint32_t n = na * nb, nc = g.nc;
    const int n_ax = pt.axis == 0 ? g.nx : (pt.axis == 1 ? g.ny : g.nz);
    const int wall = pt.side ? n_ax - 1 : 0, src = pt.side ? n_ax - 2 : 1;
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
        const int a = pt.a0 + (int)(t / nb), b = pt.b0 + (int)(t % nb);
        int cw[3], cs[3];
        if (pt.axis == 0) { cw[0] = wall; cw[1] = a; cw[2] = b; }
        else if (pt.axis == 1) { cw[0] = a; cw[1] = wall; cw[2] = b; }
        else { cw[0] = a; cw[1] = b; cw[2] = wall; }
        if (cw[0] < 0 || cw[0] >= g.nx || cw[1] < 0 || cw[1] >= g.ny || cw[2] < 0 || cw[2] >= g.nz)
            continue;
        cs[0] = cw[0];
        cs[1] = cw[1];
        cs[2] = cw[2];
        cs[pt.axis] = src;
        const uint32_t co = ((uint32_t)cw[0] * g.ny + cw[1]) * g.nz + cw[2];
        if (g.flags[co] != OUTFLOW) continue;
        const uint32_t ci = ((uint32_t)cs[0] * g.ny + cs[1]) * g.nz + cs[2];
        double r = 0.0, x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
        for (int q = 0; q < 19; q++) {
            const double v = f[(size_t)q * nc + ci];
            f[(size_t)q * nc + co] = v;
            r += v;
            x += v * lbm_ex(q);
            y += v * lbm_ey(q);
            z += v * lbm_ez(q);
        }
        store_moments(nc, co, r, x, y, z, l2s, cu, p);
    }
}

// Fused form's patch copy: the source cell's post-collision populations and
// its moments (zero for a SOLID source, whose populations are zero).
__global__ void __launch_bounds__(kBlock) k_lbm_patch_post(Geo g, double *f, LbmPatch pt, double *cu,
                                                            double *p)
{
    const uint32_t na = (uint32_t)(pt.a1 - pt.a0 + 1), nb = (uint32_t)(pt.b1 - pt.b0 + 1);
    const uint32_t n = na * nb, nc = g.nc;
    const int n_ax = pt.axis == 0 ? g.nx : (pt.axis == 1 ? g.ny : g.nz);
    const int wall = pt.side ? n_ax - 1 : 0, src = pt.side ? n_ax - 2 : 1;
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
        const int a = pt.a0 + (int)(t / nb), b = pt.b0 + (int)(t % nb);
        int cw[3], cs[3];
        if (pt.axis == 0) { cw[0] = wall; cw[1] = a; cw[2] = b; }
        else if (pt.axis == 1) { cw[0] = a; cw[1] = wall; cw[2] = b; }
        else { cw[0] = a; cw[1] = b; cw[2] = wall; }
        if (cw[0] < 0 || cw[0] >= g.nx || cw[1] < 0 || cw[1] >= g.ny || cw[2] < 0 || cw[2] >= g.nz)
            continue;
        cs[0] = cw[0];
        cs[1] = cw[1];
        cs[2] = cw[2];
        cs[pt.axis] = src;
        const uint32_t co = ((uint32_t)cw[0] * g.ny + cw[1]) * g.nz + cw[2];
        if (g.flags[co] != OUTFLOW) continue;
        const uint32_t ci = ((uint32_t)cs[0] * g.ny + cs[1]) * g.nz + cs[2];
#pragma unroll
        for (int q = 0; q < 19; q++) f[(size_t)q * nc + co] = f[(size_t)q * nc + ci];
        const bool solid = g.flags[ci] == SOLID;
        cu[co] = solid ? 0.0 : cu[ci];
        cu[nc + co] = solid ? 0.0 : cu[nc + ci];
        cu[2ull * nc + co] = solid ? 0.0 : cu[2ull * nc + ci];
        p[co] = solid ? 0.0 : p[ci];
    }
}

// MAC faces from the cell velocities (:328-376): the mean of the two cells,
// the adjacent cell on the domain's faces; the largest |new - old| of a face
// goes to rec->vmax_bits (the velocity change of :423-431).
__global__ void __launch_bounds__(kBlock) k_lbm_faces(Geo g, const double *cu, double *u, double *v,
                                                       double *w, StepRec *rec)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t nc = g.nc, total = g.nu + g.nv + g.nw;
    double dv = 0.0;
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < total; t += gridDim.x * kBlock) {
        int axis, i, j, k;
        uint32_t e;
        double *arr;
        if (t < g.nu) {
            axis = 0;
            e = t;
            arr = u;
            const uint32_t ii = fdiv(e, g.f_nynz), rem = e - ii * (uint32_t)(g.ny * g.nz);
            const uint32_t jj = fdiv(rem, g.f_nz);
            i = ii; j = jj; k = rem - jj * g.nz;
        } else if (t < g.nu + g.nv) {
            axis = 1;
            e = t - g.nu;
            arr = v;
            const uint32_t ii = fdiv(e, g.f_ny1nz), rem = e - ii * (uint32_t)((g.ny + 1) * g.nz);
            const uint32_t jj = fdiv(rem, g.f_nz);
            i = ii; j = jj; k = rem - jj * g.nz;
        } else {
            axis = 2;
            e = t - g.nu - g.nv;
            arr = w;
            const uint32_t ii = fdiv(e, g.f_nynz1), rem = e - ii * (uint32_t)(g.ny * (g.nz + 1));
            const uint32_t jj = fdiv(rem, g.f_nz1);
            i = ii; j = jj; k = rem - jj * (g.nz + 1);
        }
        const int n_ax = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
        const int pos = axis == 0 ? i : (axis == 1 ? j : k);
        const double *comp = cu + (size_t)axis * nc;
        auto cell = [&](int s) {  // the cell at position s along the axis, same other indices
            const int ci = axis == 0 ? s : i, cj = axis == 1 ? s : j, ck = axis == 2 ? s : k;
            return comp[((uint32_t)ci * g.ny + cj) * g.nz + ck];
        };
        double val;
        if (pos == 0) val = cell(0);
        else if (pos == n_ax) val = cell(n_ax - 1);
        else val = 0.5 * (cell(pos - 1) + cell(pos));
        dv = fmax(dv, fabs(val - arr[e]));
        arr[e] = val;
    }
    const double m = block_max(dv, lds);
    if (threadIdx.x == 0 && m > 0.0) atomic_max_nonneg(&rec->vmax_bits, m);
}

// ---- launchers -----------------------------------------------------------
void lbm_launch_rest(hipStream_t s, uint32_t nc, double *f)
{
    FFR_LAUNCH(k_lbm_rest, dim3(grid_for(nc)), dim3(kBlock), 0, s, nc, f);
}
void lbm_launch_relax(hipStream_t s, const Geo &g, double *f, double s2l, double inv_tau,
                      bool collide, StepRec *rec)
{
    FFR_LAUNCH(k_lbm_relax, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, f, s2l, inv_tau,
               collide ? 1 : 0, rec);
}
void lbm_launch_mask(hipStream_t s, const Geo &g, uint32_t *mask)
{
    FFR_LAUNCH(k_lbm_mask, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, mask);
}
void lbm_launch_stream(hipStream_t s, const Geo &g, const uint32_t *mask, const double *f,
                       double *f2, double *cu, double *p, double l2s)
{
    FFR_LAUNCH(k_lbm_stream, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, mask, f, f2, cu, p, l2s);
}
void lbm_launch_stream_collide(hipStream_t s, const Geo &g, const uint32_t *mask, const double *f,
                               double *f2, double *cu, double *p, double l2s, double s2l,
                               double inv_tau, StepRec *rec)
{
    FFR_LAUNCH(k_lbm_stream_collide, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, mask, f, f2, cu, p,
               l2s, s2l, inv_tau, rec);
}
void lbm_launch_patch_post(hipStream_t s, const Geo &g, double *f, const LbmPatch &pt, double *cu,
                           double *p)
{
    const uint64_t n = (uint64_t)(pt.a1 - pt.a0 + 1) * (pt.b1 - pt.b0 + 1);
    FFR_LAUNCH(k_lbm_patch_post, dim3(grid_for(n)), dim3(kBlock), 0, s, g, f, pt, cu, p);
}
void lbm_launch_patch(hipStream_t s, const Geo &g, double *f, const LbmPatch &pt, double *cu,
                      double *p, double l2s)
{
    const uint64_t n = (uint64_t)(pt.a1 - pt.a0 + 1) * (pt.b1 - pt.b0 + 1);
    FFR_LAUNCH(k_lbm_patch, dim3(grid_for(n)), dim3(kBlock), 0, s, g, f, pt, cu, p, l2s);
}
void lbm_launch_faces(hipStream_t s, const Geo &g, const double *cu, double *u, double *v,
                      double *w, StepRec *rec)
{
    const uint64_t n = (uint64_t)g.nu + g.nv + g.nw;
    FFR_LAUNCH(k_lbm_faces, dim3(grid_for(n)), dim3(kBlock), 0, s, g, cu, u, v, w, rec);
}

}  // namespace ffr
