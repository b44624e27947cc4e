// step_kernels.hip - the velocity/temperature stencils of one projection step.
//
// K1  predict:    u*, v*, w* (advection + diffusion, FluidSolver.cpp:83-309)
//                 + buoyancy on w* (:373-386) + no-slip and fan masks
//                 (BoundaryManager.cpp:60-130) from the scene's per-face mask
//                 codes, one face per thread, each block on one face array.
// K4  rhs:        b = -(rho/dt) div(u*) (:311-330).
// K6  correct:    u = u* - (dt/rho) grad p (:332-371) + no-slip, fan and
//                 outflow masks (BoundaryManager.cpp:60-130, 180-220)
//                 + max |u^{n+1}-u^n| (:527-534), in place over u^n.
// K8  temperature: T transport (:388-456) + max |div u| over FLUID (:470-484)
//                 + sum of T over FLUID (:458-468).
//     opening_T:  BoundaryManager.cpp:137-178.
//
// All kernels are templated on the field type R: double reproduces the
// reference's arithmetic operation by operation (contraction off, same
// operand order, true divisions), float is the mixed-precision form.
#include "../ffr_gang.hpp"
#include "../ffr_kernels.hpp"

#pragma clang fp contract(off)

namespace ffr {

template <typename R>
struct Axes {
    const R *hx, *hy, *hz, *i2x, *i2y, *i2z, *ihx, *ihy, *ihz;
};
__device__ __forceinline__ Axes<double> axes_of(const Geo &g, double)
{
    return {g.hx, g.hy, g.hz, g.i2x, g.i2y, g.i2z, g.ihx, g.ihy, g.ihz};
}
__device__ __forceinline__ Axes<float> axes_of(const Geo &g, float)
{
    return {g.hxf, g.hyf, g.hzf, g.i2xf, g.i2yf, g.i2zf, g.ihxf, g.ihyf, g.ihzf};
}

// Divisions by a spacing h: the FP64 form divides exactly as the reference
// does; the FP32 form multiplies by the reciprocal (rounded once on the host),
// which spares the IEEE division sequence. by_h2 divides by h*h.
__device__ __forceinline__ double by_h(double x, double h, double) { return x / h; }
__device__ __forceinline__ float by_h(float x, float, float ih) { return x * ih; }
__device__ __forceinline__ double by_h2(double x, double h, double) { return x / (h * h); }
__device__ __forceinline__ float by_h2(float x, float, float i2) { return x * i2; }
// an FP64 quantity of the step of field type R
template <typename R>
__device__ __forceinline__ double by_h64(double x, double h, double ih)
{
    return sizeof(R) == 8 ? x / h : x * ih;
}

// ---- one-direction advection terms, FluidSolver.cpp:47-81 --------------
// x / 8 is computed as x * 0.125: the same correctly rounded value (a power
// of two), so the FP64 form keeps the reference's bits.
template <typename R>
__device__ __forceinline__ R adv_upwind(R m, R c, R p, R vel, R h, R ih)
{
    if (vel > R(0)) return by_h(vel * (c - m), h, ih);
    return by_h(vel * (p - c), h, ih);
}

template <typename R>
__device__ __forceinline__ R adv_quick(R mm, R m, R c, R p, R vel, R h, R ih)
{
    R right, left;
    if (vel >= R(0)) {
        right = (R(6) * c + R(3) * p - m) * R(0.125);
        left = (R(6) * m + R(3) * c - mm) * R(0.125);
    } else {  // the reference's u<0 branch (:64-69)
        right = (R(6) * p + R(3) * c - mm) * R(0.125);
        left = (R(6) * c + R(3) * p - m) * R(0.125);
    }
    return by_h(vel * (right - left), h, ih);
}

template <typename R>
__device__ __forceinline__ R adv_lw(R m, R c, R p, R vel, R dt, R h, R ih)
{
    R cr = by_h(vel * dt, h, ih);
    R fr = vel * (R(0.5) * (c + p) - R(0.5) * cr * (p - c));
    R fl = vel * (R(0.5) * (m + c) - R(0.5) * cr * (c - m));
    return by_h(fr - fl, h, ih);
}

// Scheme switch and diffusion for one face (:132-162). nb = {-x,+x,-y,+y,-z,+z}.
template <typename R>
__device__ __forceinline__ R face_update(int scheme, bool qok, const R nb[6], const R q[3], R c,
                                         const R vel[3], const R h[3], const R ih[3],
                                         const R i2[3], R dt, R nu)
{
    R adv;
    if (scheme == QUICK && qok) {
        adv = adv_quick(q[0], nb[0], c, nb[1], vel[0], h[0], ih[0]) +
              adv_quick(q[1], nb[2], c, nb[3], vel[1], h[1], ih[1]) +
              adv_quick(q[2], nb[4], c, nb[5], vel[2], h[2], ih[2]);
    } else if (scheme == LAX_WENDROFF) {
        adv = adv_lw(nb[0], c, nb[1], vel[0], dt, h[0], ih[0]) +
              adv_lw(nb[2], c, nb[3], vel[1], dt, h[1], ih[1]) +
              adv_lw(nb[4], c, nb[5], vel[2], dt, h[2], ih[2]);
    } else {
        adv = adv_upwind(nb[0], c, nb[1], vel[0], h[0], ih[0]) +
              adv_upwind(nb[2], c, nb[3], vel[1], h[1], ih[1]) +
              adv_upwind(nb[4], c, nb[5], vel[2], h[2], ih[2]);
    }
    R dif = nu * (i2[0] * (nb[0] - R(2) * c + nb[1]) + i2[1] * (nb[2] - R(2) * c + nb[3]) +
                  i2[2] * (nb[4] - R(2) * c + nb[5]));
    return c + dt * (-adv + dif);
}

__device__ __forceinline__ double fan_v(const FanDev &f, int a, double) { return f.vel[a]; }
__device__ __forceinline__ float fan_v(const FanDev &f, int a, float) { return f.velf[a]; }

// Mask code of one face (run once per scene): no-slip zeroes wall faces
// unless the adjacent boundary cell is OUTFLOW (BoundaryManager.cpp:62-94)
// and all faces of SOLID cells (:96-107); fans then set their faces, in
// registration order, the last covering fan with a non-zero component along
// the face's axis winning (:111-130); outflow patches mark their boundary
// faces for the copy after the correction (:180-220). The face lies between
// cells lo = (i,j,k) - e_axis and hi = (i,j,k).
__device__ __forceinline__ uint16_t face_code(const Geo &g, int axis, int i, int j, int k)
{
    const uint8_t *f = g.flags;
    const int n_ax = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
    const int pos = axis == 0 ? i : (axis == 1 ? j : k);
    const int lo_i = i - (axis == 0), lo_j = j - (axis == 1), lo_k = k - (axis == 2);
    uint16_t code = 0;
    if (pos == 0 || pos == n_ax) code |= FC_BOUND;
    bool zero = false;
    if (pos == 0 && f[cidx(g, i, j, k)] != OUTFLOW) zero = true;
    if (pos == n_ax && f[cidx(g, lo_i, lo_j, lo_k)] != OUTFLOW) zero = true;
    if (pos > 0 && f[cidx(g, lo_i, lo_j, lo_k)] == SOLID) zero = true;
    if (pos < n_ax && f[cidx(g, i, j, k)] == SOLID) zero = true;
    int fan = -1;
    for (int n = 0; n < g.n_fans; n++) {
        const FanDev &fn = g.fans[n];
        if (fn.vel[axis] == 0.0) continue;
        bool in;
        if (axis == 0)
            in = i >= fn.i0 && i - 1 <= fn.i1 && j >= fn.j0 && j <= fn.j1 && k >= fn.k0 && k <= fn.k1;
        else if (axis == 1)
            in = i >= fn.i0 && i <= fn.i1 && j >= fn.j0 && j - 1 <= fn.j1 && k >= fn.k0 && k <= fn.k1;
        else
            in = i >= fn.i0 && i <= fn.i1 && j >= fn.j0 && j <= fn.j1 && k >= fn.k0 && k - 1 <= fn.k1;
        if (in) fan = n;
    }
    if (fan >= 0) code |= FC_FAN | (uint16_t)(fan << 8);
    else if (zero) code |= FC_ZERO;
    // outflow copy (BoundaryManager.cpp:180-220)
    if (axis == 0) {
        if (i == 0 && g.wall_of[wall_x(g, 0, j, k)]) code |= FC_OUT_LO;
        if (i == g.nx && g.wall_of[wall_x(g, 1, j, k)]) code |= FC_OUT_HI;
    } else if (axis == 1) {
        if (j == 0 && g.wall_of[wall_y(g, 0, i, k)]) code |= FC_OUT_LO;
        if (j == g.ny && g.wall_of[wall_y(g, 1, i, k)]) code |= FC_OUT_HI;
    } else {
        if (k == 0 && g.wall_of[wall_z(g, 0, i, j)]) code |= FC_OUT_LO;
        if (k == g.nz && g.wall_of[wall_z(g, 1, i, j)]) code |= FC_OUT_HI;
    }
    return code;
}

// (i, j, k) of face e of the array along `axis`
__device__ __forceinline__ void face_ijk(const Geo &g, int axis, uint32_t e, int &i, int &j, int &k)
{
    if (axis == 0) {
        const uint32_t ii = fdiv(e, g.f_nynz), rem = e - ii * (g.ny * g.nz);
        const uint32_t jj = fdiv(rem, g.f_nz);
        i = ii; j = jj; k = rem - jj * g.nz;
    } else if (axis == 1) {
        const uint32_t ii = fdiv(e, g.f_ny1nz), rem = e - ii * ((g.ny + 1) * g.nz);
        const uint32_t jj = fdiv(rem, g.f_nz);
        i = ii; j = jj; k = rem - jj * g.nz;
    } else {
        const uint32_t ii = fdiv(e, g.f_nynz1), rem = e - ii * (g.ny * (g.nz + 1));
        const uint32_t jj = fdiv(rem, g.f_nz1);
        i = ii; j = jj; k = rem - jj * (g.nz + 1);
    }
}

__global__ void __launch_bounds__(kBlock) k_face_codes(Geo g, uint16_t *fcode)
{
    const uint32_t n = g.nu + g.nv + g.nw;
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
        const int axis = t < g.nu ? 0 : (t < g.nu + g.nv ? 1 : 2);
        const uint32_t e = t - (axis >= 1 ? g.nu : 0u) - (axis == 2 ? g.nv : 0u);
        int i, j, k;
        face_ijk(g, axis, e, i, j, k);
        fcode[t] = face_code(g, axis, i, j, k);
    }
}

void launch_face_codes(hipStream_t s, const Geo &g, uint16_t *fcode)
{
    const uint64_t n = (uint64_t)g.nu + g.nv + g.nw;
    FFR_LAUNCH(k_face_codes, dim3(grid_for(n)), dim3(kBlock), 0, s, g, fcode);
}

// ---- scene construction on the device ------------------------------------
// Cell flags from marks applied in order, FLUID where none applies: a cell
// takes the type of the last mark whose box holds it, which is what marking
// the boxes one after the other gives (obstacles, fans, openings, automatic
// outflow: scene.py:89-106; boxes of BoundaryManager.cpp:6-44, 132-135,
// clamped to the grid - a box reaching past the grid holds only its cells).
__global__ void __launch_bounds__(kBlock) k_mark_cells(int nx, int ny, int nz, MarkList ml,
                                                        uint8_t *flags)
{
    const uint32_t nc = (uint32_t)nx * ny * nz, nynz = (uint32_t)ny * nz;
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
        const int i = (int)(c / nynz), rem = (int)(c - (uint32_t)i * nynz);
        const int j = rem / nz, k = rem - j * nz;
        uint8_t t = FLUID;
        for (int q = 0; q < ml.n; q++) {
            const int *m = ml.m[q];
            if (i >= m[1] && i <= m[2] && j >= m[3] && j <= m[4] && k >= m[5] && k <= m[6])
                t = (uint8_t)m[0];
        }
        flags[c] = t;
    }
}

// Neighbour codes of the pressure operator (NB_* bits) from the flags: the
// row structure of PoissonSolver.cpp:52-87 per cell.
__global__ void __launch_bounds__(kBlock) k_nb_codes(int nx, int ny, int nz, const uint8_t *flags,
                                                      uint16_t *nbm)
{
    const uint32_t nc = (uint32_t)nx * ny * nz, nynz = (uint32_t)ny * nz;
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
        const int i = (int)(c / nynz), rem = (int)(c - (uint32_t)i * nynz);
        const int j = rem / nz, k = rem - j * nz;
        const uint8_t f = flags[c];
        uint16_t m = 0;
        if (f == FLUID || f == INFLOW) m |= NB_SELF;
        if (f == OUTFLOW) m |= NB_SELF_OUT;
        if (f == SOLID) m |= NB_SELF_SOLID;
        if ((i + j + k) & 1) m |= NB_COLOR;
        const bool has[6] = {i > 0, i < nx - 1, j > 0, j < ny - 1, k > 0, k < nz - 1};
        const uint32_t nb[6] = {c - nynz, c + nynz, c - (uint32_t)nz, c + (uint32_t)nz, c - 1, c + 1};
#pragma unroll
        for (int q = 0; q < 6; q++) {
            if (!has[q]) continue;
            const uint8_t fn = flags[nb[q]];
            if (fn == FLUID || fn == INFLOW) m |= (uint16_t)(NB_UNK0 << q);
            else if (fn == OUTFLOW) m |= (uint16_t)(NB_OUT0 << q);
        }
        nbm[c] = m;
    }
}

void launch_mark_cells(hipStream_t s, int nx, int ny, int nz, const MarkList &ml, uint8_t *flags)
{
    FFR_LAUNCH(k_mark_cells, dim3(grid_for((uint64_t)nx * ny * nz)), dim3(kBlock), 0, s, nx, ny,
               nz, ml, flags);
}

void launch_nb_codes(hipStream_t s, int nx, int ny, int nz, const uint8_t *flags, uint16_t *nbm)
{
    FFR_LAUNCH(k_nb_codes, dim3(grid_for((uint64_t)nx * ny * nz)), dim3(kBlock), 0, s, nx, ny, nz,
               flags, nbm);
}

// the masks of a face from its code
template <typename R>
__device__ __forceinline__ R apply_mask(const Geo &g, uint16_t code, int axis, R val)
{
    if (code & FC_FAN) return fan_v(g.fans[code >> 8], axis, R(0));
    if (code & FC_ZERO) return R(0);
    return val;
}

// ---- K1: predictor -------------------------------------------------------
template <typename R, bool UNI>
__device__ __forceinline__ R predict_u(const Geo &g, const StepParams &sp, const Axes<R> &A,
                                       const R *u, const R *v, const R *w, int i, int j, int k,
                                       uint32_t e)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const uint32_t si = (uint32_t)ny * nz, sj = nz;
    const R c = u[e];
    if (i == 0 || i == nx) return c;  // boundary faces copied (:164-169)
    R nb[6], q[3] = {R(0), R(0), R(0)}, vel[3];
    const int ia = UNI ? 0 : i, ja = UNI ? 0 : j, ka = UNI ? 0 : k;  // uniform: one value per axis
    const R h[3] = {A.hx[ia], A.hy[ja], A.hz[ka]};
    const R ih[3] = {A.ihx[ia], A.ihy[ja], A.ihz[ka]};
    const R i2[3] = {A.i2x[ia], A.i2y[ja], A.i2z[ka]};
    nb[0] = i > 1 ? u[e - si] : R(0);
    nb[1] = i < nx - 1 ? u[e + si] : R(0);
    nb[2] = j > 0 ? u[e - sj] : R(0);
    nb[3] = j < ny - 1 ? u[e + sj] : R(0);
    nb[4] = k > 0 ? u[e - 1] : R(0);
    nb[5] = k < nz - 1 ? u[e + 1] : R(0);
    vel[0] = c;
    vel[1] = j > 0 ? R(0.5) * (v[vidx(g, i - 1, j, k)] + v[vidx(g, i, j, k)]) : R(0);
    vel[2] = k > 0 ? R(0.5) * (w[widx(g, i - 1, j, k)] + w[widx(g, i, j, k)]) : R(0);
    const bool qok = i >= 2 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2;
    if (sp.scheme == QUICK && qok) {
        q[0] = u[e - 2 * si];
        q[1] = u[e - 2 * sj];  // flat index, wraps into row j-1 at j=1 (:139)
        q[2] = u[e - 2];       // wraps at k=1 (:141)
    }
    return face_update<R>(sp.scheme, qok, nb, q, c, vel, h, ih, i2, (R)sp.dt, (R)sp.nu);
}

template <typename R, bool UNI>
__device__ __forceinline__ R predict_v(const Geo &g, const StepParams &sp, const Axes<R> &A,
                                       const R *u, const R *v, const R *w, int i, int j, int k,
                                       uint32_t e)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const uint32_t si = (uint32_t)(ny + 1) * nz, sj = nz;
    const R c = v[e];
    if (j == 0 || j == ny) return c;  // (:232-236)
    R nb[6], q[3] = {R(0), R(0), R(0)}, vel[3];
    const int ia = UNI ? 0 : i, ja = UNI ? 0 : j, ka = UNI ? 0 : k;  // uniform: one value per axis
    const R h[3] = {A.hx[ia], A.hy[ja], A.hz[ka]};
    const R ih[3] = {A.ihx[ia], A.ihy[ja], A.ihz[ka]};
    const R i2[3] = {A.i2x[ia], A.i2y[ja], A.i2z[ka]};
    vel[0] = i > 0 ? R(0.5) * (u[uidx(g, i, j - 1, k)] + u[uidx(g, i, j, k)]) : R(0);
    nb[0] = i > 0 ? v[e - si] : R(0);
    nb[1] = i < nx - 1 ? v[e + si] : R(0);
    nb[2] = j > 1 ? v[e - sj] : R(0);
    nb[3] = j < ny - 1 ? v[e + sj] : R(0);
    vel[1] = c;
    vel[2] = k > 0 ? R(0.5) * (w[widx(g, i, j - 1, k)] + w[widx(g, i, j, k)]) : R(0);
    nb[4] = k > 0 ? v[e - 1] : R(0);
    nb[5] = k < nz - 1 ? v[e + 1] : R(0);
    const bool qok = i >= 1 && i <= nx - 2 && j >= 2 && j <= ny - 2 && k >= 1 && k <= nz - 2;
    if (sp.scheme == QUICK && qok) {
        q[0] = v[e - si];      // adjacent value as the far upwind one (:208)
        q[1] = v[e - 2 * sj];
        q[2] = v[e - 1];       // adjacent (:211)
    }
    return face_update<R>(sp.scheme, qok, nb, q, c, vel, h, ih, i2, (R)sp.dt, (R)sp.nu);
}

template <typename R, bool UNI>
__device__ __forceinline__ R predict_w(const Geo &g, const StepParams &sp, const Axes<R> &A,
                                       const R *u, const R *v, const R *w, const R *T, int i,
                                       int j, int k, uint32_t e)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const uint32_t si = (uint32_t)ny * (nz + 1), sj = nz + 1;
    const R c = w[e];
    if (k == 0 || k == nz) return c;  // (:299-303)
    R nb[6], q[3] = {R(0), R(0), R(0)}, vel[3];
    const int ia = UNI ? 0 : i, ja = UNI ? 0 : j, ka = UNI ? 0 : k;  // uniform: one value per axis
    const R h[3] = {A.hx[ia], A.hy[ja], A.hz[ka]};
    const R ih[3] = {A.ihx[ia], A.ihy[ja], A.ihz[ka]};
    const R i2[3] = {A.i2x[ia], A.i2y[ja], A.i2z[ka]};
    vel[0] = i > 0 ? R(0.5) * (u[uidx(g, i, j, k - 1)] + u[uidx(g, i, j, k)]) : R(0);
    nb[0] = i > 0 ? w[e - si] : R(0);
    nb[1] = i < nx - 1 ? w[e + si] : R(0);
    vel[1] = j > 0 ? R(0.5) * (v[vidx(g, i, j, k - 1)] + v[vidx(g, i, j, k)]) : R(0);
    nb[2] = j > 0 ? w[e - sj] : R(0);
    nb[3] = j < ny - 1 ? w[e + sj] : R(0);
    nb[4] = k > 1 ? w[e - 1] : R(0);
    nb[5] = k < nz - 1 ? w[e + 1] : R(0);
    vel[2] = c;
    const bool qok = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 2 && k <= nz - 2;
    if (sp.scheme == QUICK && qok) {
        q[0] = w[e - si];  // adjacent (:275)
        q[1] = w[e - sj];  // adjacent (:276)
        q[2] = w[e - 2];
    }
    R val = face_update<R>(sp.scheme, qok, nb, q, c, vel, h, ih, i2, (R)sp.dt, (R)sp.nu);
    if (sp.buoyancy) {  // (:373-386), interior z-faces, before the BC masks
        const uint32_t cc = cidx(g, i, j, k);
        const R lo = T[cc - 1], hi = T[cc];
        val += (R)sp.buoy_coeff * (R(0.5) * (lo + hi) - (R)sp.T_ref);
    }
    return val;
}

// One thread per cell updates the u, v and w faces of that cell's low side
// (index (i,j,k) in each face array), so the three arrays' neighbourhoods of
// one region are fetched once into one cache; threads past the cells take the
// remaining boundary faces (u at i = nx, v at j = ny, w at k = nz).
template <typename R, bool UNI>
struct KB_predict {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const StepParams &sp,
                                               const R *u, const R *v, const R *w, const R *T,
                                               R *us, R *vs, R *ws, StepRec *rec)
{
    if (vb.raw == 0 && threadIdx.x == 0) {  // start-of-step diagnostics reset
        rec->vmax_bits = 0ull;
        rec->divmax_bits = 0ull;
        rec->tsum = 0.0;
        rec->nonfinite = 0;
    }
    const Axes<R> A = axes_of(g, R(0));
    const uint16_t *cu = g.fcode, *cv = g.fcode + g.nu, *cw = g.fcode + g.nu + g.nv;
    const uint32_t nxu = (uint32_t)g.ny * g.nz, nyv = (uint32_t)g.nx * g.nz;
    const uint32_t total = g.nc + nxu + nyv + (uint32_t)g.nx * g.ny;
    uint32_t b, end;
    block_span_of(total, vb.blk, vb.G, b, end);
    for (uint32_t t = b + threadIdx.x; t < end; t += kBlock) {
        if (t < g.nc) {
            const uint32_t i = fdiv(t, g.f_nynz), rem = t - i * nxu;
            const uint32_t j = fdiv(rem, g.f_nz), k = rem - j * g.nz;
            const uint32_t eu = t, ev = vidx(g, i, j, k), ew = widx(g, i, j, k);
            const R vu = predict_u<R, UNI>(g, sp, A, u, v, w, i, j, k, eu);
            const R vv = predict_v<R, UNI>(g, sp, A, u, v, w, i, j, k, ev);
            const R vw = predict_w<R, UNI>(g, sp, A, u, v, w, T, i, j, k, ew);
            us[eu] = apply_mask<R>(g, cu[eu], 0, vu);
            vs[ev] = apply_mask<R>(g, cv[ev], 1, vv);
            ws[ew] = apply_mask<R>(g, cw[ew], 2, vw);
        } else if (t < g.nc + nxu) {  // u at i = nx: copied, then masked
            const uint32_t e = g.nc + (t - g.nc);
            us[e] = apply_mask<R>(g, cu[e], 0, u[e]);
        } else if (t < g.nc + nxu + nyv) {  // v at j = ny
            const uint32_t q = t - g.nc - nxu, i = fdiv(q, g.f_nz), k = q - i * g.nz;
            const uint32_t e = vidx(g, i, g.ny, k);
            vs[e] = apply_mask<R>(g, cv[e], 1, v[e]);
        } else {  // w at k = nz
            const uint32_t q = t - g.nc - nxu - nyv, i = q / (uint32_t)g.ny, j = q - i * g.ny;
            const uint32_t e = widx(g, i, j, g.nz);
            ws[e] = apply_mask<R>(g, cw[e], 2, w[e]);
        }
    }
}
};
template <typename R, bool UNI>
__global__ void __launch_bounds__(kBlock) k_predict(Geo g, StepParams sp, const R *u,
                                                    const R *v, const R *w, const R *T, R *us,
                                                    R *vs, R *ws, StepRec *rec)
{
    KB_predict<R, UNI>::run(vb_here(), g, sp, u, v, w, T, us, vs, ws, rec);
}

template <typename R>
void launch_predict(hipStream_t s, const Geo &g, const StepParams &sp, const R *u, const R *v,
                    const R *w, const R *T, R *us, R *vs, R *ws, StepRec *rec)
{
    const uint64_t n = (uint64_t)g.nc + (uint64_t)g.ny * g.nz + (uint64_t)g.nx * g.nz +
                       (uint64_t)g.nx * g.ny;
    if (g.uniform)
        FFR_GANG((KB_predict<R, true>), GB_PLAIN, (k_predict<R, true>), dim3(grid_for(n)),
                 dim3(kBlock), 0, s, g, sp, u, v, w, T, us, vs, ws, rec);
    else
        FFR_GANG((KB_predict<R, false>), GB_PLAIN, (k_predict<R, false>), dim3(grid_for(n)),
                 dim3(kBlock), 0, s, g, sp, u, v, w, T, us, vs, ws, rec);
}

// ---- K4: Poisson right-hand side ----------------------------------------
template <typename R>
struct KB_rhs {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const StepParams &sp,
                                               const R *us, const R *vs, const R *ws, double *b)
{
    uint32_t sb, se;
    block_span_of(g.nc, vb.blk, vb.G, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        if (g.flags[c] == SOLID) {
            b[c] = 0.0;
            continue;
        }
        const uint32_t i = fdiv(c, g.f_nynz), rem = c - i * (g.ny * g.nz);
        const uint32_t j = fdiv(rem, g.f_nz), k = rem - j * g.nz;
        const double du = (double)us[c + g.ny * g.nz] - (double)us[c];
        const uint32_t ev = vidx(g, i, j, k), ew = widx(g, i, j, k);
        const double dv = (double)vs[ev + g.nz] - (double)vs[ev];
        const double dw = (double)ws[ew + 1] - (double)ws[ew];
        const double div = by_h64<R>(du, g.hx[i], g.ihx[i]) + by_h64<R>(dv, g.hy[j], g.ihy[j]) +
                           by_h64<R>(dw, g.hz[k], g.ihz[k]);
        b[c] = sp.rhs_factor * div;
    }
}
};
template <typename R>
__global__ void __launch_bounds__(kBlock) k_rhs(Geo g, StepParams sp, const R *us, const R *vs,
                                                const R *ws, double *b)
{
    KB_rhs<R>::run(vb_here(), g, sp, us, vs, ws, b);
}

template <typename R>
void launch_rhs(hipStream_t s, const Geo &g, const StepParams &sp, const R *us, const R *vs,
                const R *ws, double *b)
{
    FFR_GANG((KB_rhs<R>), GB_PLAIN, k_rhs<R>, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, sp, us,
             vs, ws, b);
}

// ---- K6: projection + post-projection masks + velocity change ----------
template <typename R>
__device__ __forceinline__ double pgrad(const Geo &g, const double *p, int axis, int i, int j,
                                        int k)
{
    const uint32_t c = cidx(g, i, j, k);
    if (axis == 0) return by_h64<R>(p[c] - p[c - g.ny * g.nz], g.cx[i], g.icx[i]);
    if (axis == 1) return by_h64<R>(p[c] - p[c - g.nz], g.cy[j], g.icy[j]);
    return by_h64<R>(p[c] - p[c - 1], g.cz[k], g.icz[k]);
}

// Value of a face after correction, no-slip and fans (before outflow copy).
template <typename R>
__device__ __forceinline__ R corrected(const Geo &g, const StepParams &sp, const R *fs,
                                       const double *p, int axis, int i, int j, int k, uint32_t e,
                                       uint16_t code)
{
    R x = fs[e];
    if (!(code & FC_BOUND)) {
        const double gp = pgrad<R>(g, p, axis, i, j, k);
        x = x - (R)(sp.dt_rho * gp);
    }
    return apply_mask<R>(g, code, axis, x);
}

// fp64: x -= dt_rho*gp exactly as the reference (:345); the generic version
// rounds dt_rho*gp to R first, which for R = double is the same operation.

// One block section per face array (u, v, w), one face per thread.
template <typename R>
struct KB_correct {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const StepParams &sp,
                                               const FaceSections &fsec, const R *us, const R *vs,
                                               const R *ws, const double *p, R *u, R *v, R *w,
                                               StepRec *rec)
{
    __shared__ double lds[kBlock / kWave];
    double dv = 0.0;
    bool bad = false;
    uint32_t lb, G, sb, se;
    const int axis = face_section_of(fsec, vb.blk, vb.G, lb, G);  // uniform per block
    // all three counts loaded, then combined: a select between the loads
    // would become one load at a selected offset into g, which keeps a copy
    // of g in scratch memory
    const uint32_t nu = g.nu, nv = g.nv, nw = g.nw;
    const uint32_t n = nu + (axis >= 1 ? nv - nu : 0u) + (axis == 2 ? nw - nv : 0u);
    const uint16_t *code = g.fcode + (axis >= 1 ? g.nu : 0u) + (axis == 2 ? g.nv : 0u);
    const R *fs = axis == 0 ? us : (axis == 1 ? vs : ws);
    R *out = axis == 0 ? u : (axis == 1 ? v : w);
    // index step to the next face along the axis
    const uint32_t step = axis == 0 ? (uint32_t)g.ny * g.nz : (axis == 1 ? (uint32_t)g.nz : 1u);
    block_span_of(n, lb, G, sb, se);
    for (uint32_t e = sb + threadIdx.x; e < se; e += kBlock) {
        const uint16_t c = code[e];
        int i, j, k;
        face_ijk(g, axis, e, i, j, k);
        R val;
        if (c & (FC_OUT_LO | FC_OUT_HI)) {  // outflow patch: the face one step inside
            const int d = (c & FC_OUT_LO) ? 1 : -1;
            const uint32_t e2 = d > 0 ? e + step : e - step;
            val = corrected<R>(g, sp, fs, p, axis, i + d * (axis == 0), j + d * (axis == 1),
                               k + d * (axis == 2), e2, code[e2]);
        } else {
            val = corrected<R>(g, sp, fs, p, axis, i, j, k, e, c);
        }
        const R old = out[e];
        out[e] = val;
        dv = fmax(dv, fabs((double)val - (double)old));  // fmax drops NaN as std::max does
        bad |= !isfinite((double)val);
    }
    if (bad) rec->nonfinite = 1;
    const double m = block_max(dv, lds);
    if (threadIdx.x == 0 && m > 0.0) atomic_max_nonneg(&rec->vmax_bits, m);
}
};
template <typename R>
__global__ void __launch_bounds__(kBlock) k_correct(Geo g, StepParams sp, FaceSections fsec,
                                                    const R *us, const R *vs, const R *ws,
                                                    const double *p, R *u, R *v, R *w,
                                                    StepRec *rec)
{
    KB_correct<R>::run(vb_here(), g, sp, fsec, us, vs, ws, p, u, v, w, rec);
}

template <typename R>
void launch_correct(hipStream_t s, const Geo &g, const StepParams &sp, const R *us, const R *vs,
                    const R *ws, const double *p, R *u, R *v, R *w, StepRec *rec)
{
    const uint64_t n = (uint64_t)g.nu + g.nv + g.nw;
    const uint32_t G = std::max(3u, grid_for(n));
    const FaceSections fs = face_sections(g.nu, g.nv, g.nw, G);
    FFR_GANG((KB_correct<R>), GB_PLAIN, k_correct<R>, dim3(G), dim3(kBlock), 0, s, g, sp, fs, us,
             vs, ws, p, u, v, w, rec);
}

// ---- divergence (FluidSolver.cpp:470-484) --------------------------------
template <typename R>
__device__ __forceinline__ double cell_div(const Geo &g, const R *u, const R *v, const R *w,
                                           int i, int j, int k, uint32_t c)
{
    const uint32_t ev = vidx(g, i, j, k), ew = widx(g, i, j, k);
    return fabs(by_h64<R>((double)u[c + g.ny * g.nz] - (double)u[c], g.hx[i], g.ihx[i]) +
                by_h64<R>((double)v[ev + g.nz] - (double)v[ev], g.hy[j], g.ihy[j]) +
                by_h64<R>((double)w[ew + 1] - (double)w[ew], g.hz[k], g.ihz[k]));
}

template <typename R>
struct KB_divmax {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const R *u, const R *v,
                                               const R *w, StepRec *rec)
{
    __shared__ double lds[kBlock / kWave];
    double d = 0.0;
    uint32_t sb, se;
    block_span_of(g.nc, vb.blk, vb.G, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        if (g.flags[c] != FLUID) continue;
        const uint32_t i = fdiv(c, g.f_nynz), rem = c - i * (g.ny * g.nz);
        const uint32_t j = fdiv(rem, g.f_nz), k = rem - j * g.nz;
        d = fmax(d, cell_div(g, u, v, w, i, j, k, c));  // drops NaN like std::max
    }
    const double m = block_max(d, lds);
    if (threadIdx.x == 0 && m > 0.0) atomic_max_nonneg(&rec->divmax_bits, m);
}
};
template <typename R>
__global__ void __launch_bounds__(kBlock) k_divmax(Geo g, const R *u, const R *v, const R *w,
                                                   StepRec *rec)
{
    KB_divmax<R>::run(vb_here(), g, u, v, w, rec);
}

template <typename R>
void launch_divmax(hipStream_t s, const Geo &g, const R *u, const R *v, const R *w, StepRec *rec)
{
    FFR_GANG((KB_divmax<R>), GB_PLAIN, k_divmax<R>, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, u,
             v, w, rec);
}

// ---- K8: temperature transport (FluidSolver.cpp:388-456) ---------------
template <typename R, bool UNI>
struct KB_temperature {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const StepParams &sp,
                                               const R *u, const R *v, const R *w, const R *T,
                                               R *Tn, StepRec *rec, const RedScratch &red)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t blk = vb.blk;
    double dmax = 0.0, tsum = 0.0;
    bool bad = false;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const Axes<R> A = axes_of(g, R(0));
    uint32_t sb, se;
    block_span_of(g.nc, blk, vb.G, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        if (g.flags[c] != FLUID) {  // non-FLUID cells are frozen (:402-405)
            Tn[c] = T[c];
            continue;
        }
        const uint32_t i = fdiv(c, g.f_nynz), rem = c - i * (ny * nz);
        const uint32_t j = fdiv(rem, g.f_nz), k = rem - j * nz;
        const int ia = UNI ? 0 : i, ja = UNI ? 0 : j, ka = UNI ? 0 : k;
        const R hx = A.hx[ia], hy = A.hy[ja], hz = A.hz[ka];
        const R ihx = A.ihx[ia], ihy = A.ihy[ja], ihz = A.ihz[ka];
        const R i2x = A.i2x[ia], i2y = A.i2y[ja], i2z = A.i2z[ka];
        const R Tc = T[c];
        const uint32_t ev = vidx(g, i, j, k), ew = widx(g, i, j, k);
        const R u0 = u[c], u1 = u[c + ny * nz], v0 = v[ev], v1 = v[ev + nz], w0 = w[ew],
                w1 = w[ew + 1];
        const R uc = R(0.5) * (u0 + u1), vc = R(0.5) * (v0 + v1), wc = R(0.5) * (w0 + w1);
        // neighbour T: Neumann (Tc) outside the domain and at SOLID (:417-422)
        auto tn = [&](int a, int bb, int d) -> R {
            if (a < 0 || a >= nx || bb < 0 || bb >= ny || d < 0 || d >= nz) return Tc;
            const uint32_t q = cidx(g, a, bb, d);
            return g.flags[q] == SOLID ? Tc : T[q];
        };
        const int ii = i, jj = j, kk = k;
        const R xm = tn(ii - 1, jj, kk), xp = tn(ii + 1, jj, kk);
        const R ym = tn(ii, jj - 1, kk), yp = tn(ii, jj + 1, kk);
        const R zm = tn(ii, jj, kk - 1), zp = tn(ii, jj, kk + 1);
        const R dt = (R)sp.dt;
        R adv;
        if (sp.scheme == QUICK && ii >= 1 && ii <= nx - 2 && jj >= 1 && jj <= ny - 2 && kk >= 1 &&
            kk <= nz - 2) {
            const R xmm = tn(ii - 2, jj, kk), ymm = tn(ii, jj - 2, kk), zmm = tn(ii, jj, kk - 2);
            adv = adv_quick(xmm, xm, Tc, xp, uc, hx, ihx) + adv_quick(ymm, ym, Tc, yp, vc, hy, ihy) +
                  adv_quick(zmm, zm, Tc, zp, wc, hz, ihz);
        } else if (sp.scheme == LAX_WENDROFF) {
            adv = adv_lw(xm, Tc, xp, uc, dt, hx, ihx) + adv_lw(ym, Tc, yp, vc, dt, hy, ihy) +
                  adv_lw(zm, Tc, zp, wc, dt, hz, ihz);
        } else {
            adv = adv_upwind(xm, Tc, xp, uc, hx, ihx) + adv_upwind(ym, Tc, yp, vc, hy, ihy) +
                  adv_upwind(zm, Tc, zp, wc, hz, ihz);
        }
        const R al = (R)sp.alpha_T;
        const R dif = al * (by_h2(xm - R(2) * Tc + xp, hx, i2x) + by_h2(ym - R(2) * Tc + yp, hy, i2y) +
                            by_h2(zm - R(2) * Tc + zp, hz, i2z));
        const R tnew = Tc + dt * (-adv + dif);
        Tn[c] = tnew;
        tsum += (double)tnew;
        bad |= !isfinite((double)tnew);
        dmax = fmax(dmax, fabs(by_h64<R>((double)u1 - (double)u0, g.hx[ia], g.ihx[ia]) +
                               by_h64<R>((double)v1 - (double)v0, g.hy[ja], g.ihy[ja]) +
                               by_h64<R>((double)w1 - (double)w0, g.hz[ka], g.ihz[ka])));
    }
    if (bad) rec->nonfinite = 1;
    const double m = block_max(dmax, lds);
    if (threadIdx.x == 0 && m > 0.0) atomic_max_nonneg(&rec->divmax_bits, m);
    block_partial(tsum, red.partials + SLOT_TSUM * kMaxBlocksCap, blk, lds);
}
};
template <typename R, bool UNI>
__global__ void __launch_bounds__(kBlock) k_temperature(Geo g, StepParams sp, const R *u,
                                                        const R *v, const R *w, const R *T,
                                                        R *Tn, StepRec *rec, RedScratch red)
{
    KB_temperature<R, UNI>::run(vb_here(), g, sp, u, v, w, T, Tn, rec, red);
}

template <typename R>
void launch_temperature(hipStream_t s, const Geo &g, const StepParams &sp, const R *u,
                        const R *v, const R *w, const R *T, R *Tn, StepRec *rec, RedScratch red)
{
    if (g.uniform)
        FFR_GANG((KB_temperature<R, true>), GB_PLAIN, (k_temperature<R, true>),
                 dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, sp, u, v, w, T, Tn, rec, red);
    else
        FFR_GANG((KB_temperature<R, false>), GB_PLAIN, (k_temperature<R, false>),
                 dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, sp, u, v, w, T, Tn, rec, red);
}

// Sum of T over FLUID cells (used when an opening overwrites FLUID cells).
template <typename R>
struct KB_tsum {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const R *T, StepRec *rec,
                                               const RedScratch &red)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t blk = vb.blk;
    double t = 0.0;
    uint32_t sb, se;
    block_span_of(g.nc, blk, vb.G, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock)
        if (g.flags[c] == FLUID) t += (double)T[c];
    block_partial(t, red.partials + SLOT_TSUM * kMaxBlocksCap, blk, lds);
}
};
template <typename R>
__global__ void __launch_bounds__(kBlock) k_tsum(Geo g, const R *T, StepRec *rec, RedScratch red)
{
    KB_tsum<R>::run(vb_here(), g, T, rec, red);
}

template <typename R>
void launch_tsum(hipStream_t s, const Geo &g, const R *T, StepRec *rec, RedScratch red)
{
    FFR_GANG((KB_tsum<R>), GB_PLAIN, k_tsum<R>, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, T, rec,
             red);
}

// rec->tsum from the per-block partials of temperature / tsum (one block).
struct KB_step_finalize {
    static __device__ __forceinline__ void run(const VB &vb, StepRec *rec, const RedScratch &red,
                                               uint32_t G)
{
    __shared__ double lds[kBlock / kWave];
    const double t = sum_partials(red.partials + SLOT_TSUM * kMaxBlocksCap, G, lds);
    if (threadIdx.x == 0) rec->tsum = t;
}
};
__global__ void __launch_bounds__(kBlock) k_step_finalize(StepRec *rec, RedScratch red, uint32_t G)
{
    KB_step_finalize::run(vb_here(), rec, red, G);
}

void launch_step_finalize(hipStream_t s, const Geo &g, StepRec *rec, RedScratch red)
{
    FFR_GANG((KB_step_finalize), GB_PLAIN, k_step_finalize, dim3(1), dim3(kBlock), 0, s, rec, red,
             grid_for(g.nc));
}

// ---- opening temperatures (BoundaryManager.cpp:137-178) -----------------
// One block walks the openings in registration order, so overlapping
// patches resolve exactly as the reference's sequential loop does.
template <typename R>
struct KB_opening_T {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const R *u, const R *v,
                                               const R *w, R *T)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    for (int n = 0; n < g.n_openings; n++) {
        const OpeningDev o = g.openings[n];
        const int na = o.a1 - o.a0 + 1, nb = o.b1 - o.b0 + 1;
        const int cnt = na * nb;
        for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
            const int a = o.a0 + t / nb, b = o.b0 + t % nb;
            double vf;
            uint32_t wall, inner;
            const int sd = o.side;
            if (o.axis == 0) {
                vf = (double)u[uidx(g, sd ? nx : 0, a, b)];
                wall = cidx(g, sd ? nx - 1 : 0, a, b);
                inner = cidx(g, sd ? nx - 2 : 1, a, b);
            } else if (o.axis == 1) {
                vf = (double)v[vidx(g, a, sd ? ny : 0, b)];
                wall = cidx(g, a, sd ? ny - 1 : 0, b);
                inner = cidx(g, a, sd ? ny - 2 : 1, b);
            } else {
                vf = (double)w[widx(g, a, b, sd ? nz : 0)];
                wall = cidx(g, a, b, sd ? nz - 1 : 0);
                inner = cidx(g, a, b, sd ? nz - 2 : 1);
            }
            const bool incoming = sd ? (vf <= 0.0) : (vf >= 0.0);
            T[wall] = incoming ? (R)o.T_out : T[inner];
        }
        __syncthreads();
    }
}
};
template <typename R>
__global__ void __launch_bounds__(kBlock) k_opening_T(Geo g, const R *u, const R *v, const R *w,
                                                      R *T)
{
    KB_opening_T<R>::run(vb_here(), g, u, v, w, T);
}

template <typename R>
void launch_opening_T(hipStream_t s, const Geo &g, const R *u, const R *v, const R *w, R *T)
{
    if (g.n_openings == 0) return;
    FFR_GANG((KB_opening_T<R>), GB_PLAIN, k_opening_T<R>, dim3(1), dim3(kBlock), 0, s, g, u, v, w, T);
}

// ---- precision conversion ------------------------------------------------
template <typename R>
__global__ void k_convert(const double *src, R *dst, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kBlock)
        dst[i] = (R)src[i];
}
template <typename R>
__global__ void k_convert_back(const R *src, double *dst, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kBlock)
        dst[i] = (double)src[i];
}
template <typename R>
void launch_convert(hipStream_t s, const double *src, R *dst, uint64_t n)
{
    if (n) FFR_LAUNCH(k_convert<R>, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, n);
}
template <typename R>
void launch_convert_back(hipStream_t s, const R *src, double *dst, uint64_t n)
{
    if (n) FFR_LAUNCH(k_convert_back<R>, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, n);
}

__global__ void __launch_bounds__(kBlock) k_copy64(const double *__restrict__ src,
                                                   double *__restrict__ dst, uint32_t n)
{
    uint32_t sb, se;
    block_span(n, xcd_block(), sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) dst[c] = src[c];
}

void launch_copy64(hipStream_t s, const double *src, double *dst, uint64_t n)
{
    FFR_LAUNCH(k_copy64, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, (uint32_t)n);
}

#define FFR_INSTANTIATE(R)                                                                      \
    template void launch_predict<R>(hipStream_t, const Geo &, const StepParams &, const R *,   \
                                    const R *, const R *, const R *, R *, R *, R *, StepRec *); \
    template void launch_correct<R>(hipStream_t, const Geo &, const StepParams &, const R *,   \
                                    const R *, const R *, const double *, R *, R *, R *,       \
                                    StepRec *);                                                \
    template void launch_divmax<R>(hipStream_t, const Geo &, const R *, const R *, const R *,  \
                                   StepRec *);                                                 \
    template void launch_temperature<R>(hipStream_t, const Geo &, const StepParams &,          \
                                        const R *, const R *, const R *, const R *, R *,       \
                                        StepRec *, RedScratch);                                \
    template void launch_tsum<R>(hipStream_t, const Geo &, const R *, StepRec *, RedScratch);  \
    template void launch_opening_T<R>(hipStream_t, const Geo &, const R *, const R *,          \
                                      const R *, R *);                                         \
    template void launch_rhs<R>(hipStream_t, const Geo &, const StepParams &, const R *,       \
                                const R *, const R *, double *);                               \
    template void launch_convert<R>(hipStream_t, const double *, R *, uint64_t);               \
    template void launch_convert_back<R>(hipStream_t, const R *, double *, uint64_t);

FFR_INSTANTIATE(double)
FFR_INSTANTIATE(float)

}  // namespace ffr
