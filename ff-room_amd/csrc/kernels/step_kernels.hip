// step_kernels.hip - the velocity/temperature stencils of one projection step.
//
// K1  predict:    u*, v*, w* (advection + diffusion, FluidSolver.cpp:83-309)
//                 + buoyancy on w* (:373-386) + no-slip and fan masks
//                 (BoundaryManager.cpp:60-130) from the scene's per-face mask
//                 codes; one thread per cell column marching x planes, the
//                 tile's y/z neighbourhood staged in LDS.
// K4  rhs:        b = -(rho/dt) div(u*) (:311-330).
// K6  correct:    u = u* - (dt/rho) grad p (:332-371) + no-slip, fan and
//                 outflow masks (BoundaryManager.cpp:60-130, 180-220)
//                 + max |u^{n+1}-u^n| (:527-534), in place over u^n; one
//                 thread per cell's three low faces.
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

// The FP32 form regroups the same terms for fused multiply-adds (it is
// checked against the FP64 form within FP32 rounding, not bit for bit):
// QUICK's 8 (right - left) is 3(c+p) - 7m + mm for vel >= 0 and, in the
// reference's other branch, 3(p-c) + m - mm; Lax-Wendroff's fr - fl is
// vel (p - m)/2 - vel cr (p - 2c + m)/2.
__device__ __forceinline__ float adv_upwind(float m, float c, float p, float vel, float, float ih)
{
    return vel * (vel > 0.0f ? c - m : p - c) * ih;
}
__device__ __forceinline__ float adv_quick(float mm, float m, float c, float p, float vel, float,
                                           float ih)
{
    const float pos = fmaf(3.0f, c + p, fmaf(-7.0f, m, mm));
    const float neg = fmaf(3.0f, p - c, m - mm);
    return vel * (vel >= 0.0f ? pos : neg) * (0.125f * ih);
}
__device__ __forceinline__ float adv_lw(float m, float c, float p, float vel, float dt, float,
                                        float ih)
{
    const float cr = vel * dt * ih;
    const float curv = fmaf(-2.0f, c, p + m);
    return 0.5f * vel * fmaf(-cr, curv, p - m) * ih;
}
// nu * sum over axes of (nb- - 2c + nb+) / h^2
template <typename R>
__device__ __forceinline__ R diffusion(const R nb[6], R c, const R i2[3], R nu)
{
    return nu * (i2[0] * (nb[0] - R(2) * c + nb[1]) + i2[1] * (nb[2] - R(2) * c + nb[3]) +
                 i2[2] * (nb[4] - R(2) * c + nb[5]));
}
__device__ __forceinline__ float diffusion(const float nb[6], float c, const float i2[3], float nu)
{
    float acc = i2[0] * fmaf(-2.0f, c, nb[0] + nb[1]);
    acc = fmaf(i2[1], fmaf(-2.0f, c, nb[2] + nb[3]), acc);
    acc = fmaf(i2[2], fmaf(-2.0f, c, nb[4] + nb[5]), acc);
    return nu * acc;
}
// c + dt (-adv + dif)
template <typename R>
__device__ __forceinline__ R advance(R c, R dt, R adv, R dif)
{
    return c + dt * (-adv + dif);
}
__device__ __forceinline__ float advance(float c, float dt, float adv, float dif)
{
    return fmaf(dt, dif - adv, c);
}

// Scheme and diffusion for one face (:132-162), the scheme fixed at compile
// time. nb = {-x,+x,-y,+y,-z,+z}; q = the far upwind values of QUICK.
template <int SCHEME, typename R>
__device__ __forceinline__ R face_update_s(bool qok, const R nb[6], const R q[3], R c,
                                           const R vel[3], const R h[3], const R ih[3],
                                           const R i2[3], R dt, R nu)
{
    R adv;
    if (SCHEME == QUICK && qok) {
        adv = adv_quick(q[0], nb[0], c, nb[1], vel[0], h[0], ih[0]) +
              adv_quick(q[1], nb[2], c, nb[3], vel[1], h[1], ih[1]) +
              adv_quick(q[2], nb[4], c, nb[5], vel[2], h[2], ih[2]);
    } else if (SCHEME == LAX_WENDROFF) {
        adv = adv_lw(nb[0], c, nb[1], vel[0], dt, h[0], ih[0]) +
              adv_lw(nb[2], c, nb[3], vel[1], dt, h[1], ih[1]) +
              adv_lw(nb[4], c, nb[5], vel[2], dt, h[2], ih[2]);
    } else {
        adv = adv_upwind(nb[0], c, nb[1], vel[0], h[0], ih[0]) +
              adv_upwind(nb[2], c, nb[3], vel[1], h[1], ih[1]) +
              adv_upwind(nb[4], c, nb[5], vel[2], h[2], ih[2]);
    }
    return advance(c, dt, adv, diffusion(nb, c, i2, nu));
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
// The predictor marches planes of cells along x (the slowest index). A block
// owns a tile of TJ x TK cells in (y, z) and a run of x planes; each thread
// owns one (j, k) column of the tile and updates the low u, v and w faces of
// its cell at every plane. The tile's plane of u, v, w (and T for buoyancy),
// with a halo of two rows/columns below and one above, is staged in LDS, so
// the y and z neighbours are LDS reads at fixed offsets; the x neighbours of
// the thread's own column stay in registers as the march advances. Two LDS
// buffers hold planes i and i+1, and the loads of plane i+2 are issued before
// plane i is computed. Halo entries outside the domain hold 0, the value the
// reference gives a missing neighbour (Q5); in-domain values the reference
// replaces by 0 (the boundary faces one step inside, FluidSolver.cpp:110-130,
// 192-202, 259-269) keep their selects. The QUICK reads that the reference
// takes through a wrapped flat index (u(i, j-2, k) at j = 1 and u(i, j, k-2)
// at k = 1, Q3) are global loads of that flat index.
namespace ptile {
constexpr int TK = 32, TJ = 8;              // cells of a tile in z and y
constexpr int RK = TK + 3, RJ = TJ + 3;     // with the halo: 2 below, 1 above
constexpr int PLANE = RK * RJ;              // LDS elements of one plane of one field
constexpr int SLOTS = (PLANE + kBlock - 1) / kBlock;
static_assert(TK * TJ == kBlock, "one cell column per thread");
}  // namespace ptile

// Tiling of the cell grid for the marching kernels: ntk x ntj tiles in
// (z, y), nic runs of ic planes in x; tile index k fastest, then y, then the
// run, so the contiguous tile ranges of one XCD share their halos in L2.
struct PlaneTiles {
    uint32_t ntk, ntj, nic, ic;
};

inline PlaneTiles plane_tiles(const Geo &g, uint32_t min_planes)
{
    PlaneTiles t;
    t.ntk = (uint32_t)(g.nz + ptile::TK - 1) / ptile::TK;
    t.ntj = (uint32_t)(g.ny + ptile::TJ - 1) / ptile::TJ;
    const uint32_t t2 = t.ntk * t.ntj;
    const uint32_t want = (4096 + t2 - 1) / t2;  // runs for about 4096 tiles
    const uint32_t most = std::max<uint32_t>(1, (uint32_t)g.nx / min_planes);
    const uint32_t runs = std::max<uint32_t>(1, std::min(want, most));
    t.ic = ((uint32_t)g.nx + runs - 1) / runs;
    t.nic = ((uint32_t)g.nx + t.ic - 1) / t.ic;
    return t;
}
inline uint32_t plane_tiles_grid(const PlaneTiles &t)
{
    return std::min<uint32_t>(t.ntk * t.ntj * t.nic, kMaxBlocksCap);
}

// Where one load slot of a thread lies in the tile's plane: its offsets in
// the u/T, v and w planes (k fastest), and whether that element exists in
// each array. Offsets of elements that do not exist are never used.
struct SlotMap {
    uint32_t ou, ov, ow;
    bool in_u, in_v, in_w;
};
__device__ __forceinline__ SlotMap slot_map(const Geo &g, int j0, int k0, int e)
{
    SlotMap m;
    const int jj = j0 - 2 + e / ptile::RK, kk = k0 - 2 + e % ptile::RK;
    const bool on = e < ptile::PLANE;  // the last slot of a thread may lie past the plane
    const bool jin = on && jj >= 0 && jj < g.ny, kin = on && kk >= 0 && kk < g.nz;
    m.in_u = jin && kin;
    m.in_v = on && jj >= 0 && jj <= g.ny && kin;
    m.in_w = jin && on && kk >= 0 && kk <= g.nz;
    const uint32_t jc = (uint32_t)std::max(jj, 0), kc = (uint32_t)std::max(kk, 0);
    m.ou = jc * g.nz + kc;
    m.ov = m.ou;
    m.ow = jc * (g.nz + 1) + kc;
    return m;
}

template <typename R>
struct PlaneRegs {
    R u[ptile::SLOTS], v[ptile::SLOTS], w[ptile::SLOTS], t[ptile::SLOTS];
};

// Raw buffer loads at a byte offset: an offset outside the array reads 0
// (the hardware range check), so a missing element costs no branch.
constexpr uint32_t kNoElement = 0xFFFFFFF0u;  // a byte offset past any array below 4 GiB
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off, float)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off, double)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

// Buffer resources of the arrays a march reads.
struct MarchBufs {
    __amdgpu_buffer_rsrc_t u, v, w, t, f;
};
template <typename R>
__device__ __forceinline__ MarchBufs march_bufs(const Geo &g, const R *u, const R *v, const R *w,
                                                const R *T)
{
    return {buf_rsrc(u, g.nu * sizeof(R)), buf_rsrc(v, g.nv * sizeof(R)),
            buf_rsrc(w, g.nw * sizeof(R)), buf_rsrc(T, T ? g.nc * sizeof(R) : 0u),
            buf_rsrc(g.flags, g.nc)};
}

// Issues the loads of plane p of u, v, w (and T) for this thread's slots; 0
// where the element does not exist.
template <typename R>
__device__ __forceinline__ void plane_load(const Geo &g, const SlotMap *sm, int p, bool with_t,
                                           const MarchBufs &b, PlaneRegs<R> &r)
{
    const bool pu = p >= 0 && p <= g.nx, pc = p >= 0 && p < g.nx;
    const uint32_t bu = (uint32_t)p * g.ny * g.nz, bv = (uint32_t)p * (g.ny + 1) * g.nz;
    const uint32_t bw = (uint32_t)p * g.ny * (g.nz + 1);
    constexpr uint32_t sz = sizeof(R);
#pragma unroll
    for (int s = 0; s < ptile::SLOTS; s++) {
        const SlotMap &m = sm[s];
        r.u[s] = buf_ld(b.u, (pu && m.in_u) ? (bu + m.ou) * sz : kNoElement, R(0));
        r.v[s] = buf_ld(b.v, (pc && m.in_v) ? (bv + m.ov) * sz : kNoElement, R(0));
        r.w[s] = buf_ld(b.w, (pc && m.in_w) ? (bw + m.ow) * sz : kNoElement, R(0));
        r.t[s] = with_t ? buf_ld(b.t, (pc && m.in_u) ? (bu + m.ou) * sz : kNoElement, R(0)) : R(0);
    }
}

template <typename R>
__device__ __forceinline__ void plane_store(const PlaneRegs<R> &r, bool with_t, R *su, R *sv, R *sw,
                                            R *st)
{
#pragma unroll
    for (int s = 0; s < ptile::SLOTS; s++) {
        const int e = threadIdx.x + s * kBlock;
        if (s + 1 < ptile::SLOTS || e < ptile::PLANE) {
            su[e] = r.u[s];
            sv[e] = r.v[s];
            sw[e] = r.w[s];
            if (with_t) st[e] = r.t[s];
        }
    }
}

// The per-face values of the thread's cell on plane p that do not come from
// the staged planes: the three face mask codes and, for QUICK, the two reads
// the reference takes through a wrapped flat index (Q3): u(p, j-1, nz-1) for
// k = 1 and u(p-1, ny-1, k) for j = 1. Loaded one plane ahead.
template <typename R>
struct FaceExtras {
    uint32_t cuv, cw;  // codes: u | v << 16, w
    R wk, wj;
};
template <typename R, int SCHEME>
__device__ __forceinline__ FaceExtras<R> face_extras(const Geo &g, const MarchBufs &b,
                                                     __amdgpu_buffer_rsrc_t codes, bool act, int p,
                                                     int j, int k)
{
    FaceExtras<R> x;
    const bool in = act && p < g.nx;
    const uint32_t eu = uidx(g, p, j, k), ev = vidx(g, p, j, k), ew = widx(g, p, j, k);
    const uint32_t cu = __builtin_amdgcn_raw_buffer_load_b16(codes, (int)(in ? eu * 2u : kNoElement), 0, 0);
    const uint32_t cv = __builtin_amdgcn_raw_buffer_load_b16(codes, (int)(in ? (g.nu + ev) * 2u : kNoElement), 0, 0);
    x.cw = __builtin_amdgcn_raw_buffer_load_b16(codes, (int)(in ? (g.nu + g.nv + ew) * 2u : kNoElement), 0, 0);
    x.cuv = cu | (cv << 16);
    constexpr uint32_t sz = sizeof(R);
    if (SCHEME == QUICK) {
        x.wk = buf_ld(b.u, (in && k == 1 && j >= 1) ? (eu - 2u) * sz : kNoElement, R(0));
        x.wj = buf_ld(b.u, (in && j == 1 && p >= 1) ? (eu - 2u * (uint32_t)g.nz) * sz : kNoElement,
                      R(0));
    } else {
        x.wk = x.wj = R(0);
    }
    return x;
}

template <typename R, bool UNI, int SCHEME>
struct KB_predict {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const StepParams &sp,
                                               const PlaneTiles pt, const R *u, const R *v,
                                               const R *w, const R *T, R *us, R *vs, R *ws,
                                               StepRec *rec)
{
    using namespace ptile;
    __shared__ R sU[2][PLANE], sV[2][PLANE], sW[2][PLANE], sT[2][PLANE];
    if (vb.raw == 0 && threadIdx.x == 0) {  // start-of-step diagnostics reset
        rec->vmax_bits = 0ull;
        rec->divmax_bits = 0ull;
        rec->tsum = 0.0;
        rec->nonfinite = 0;
    }
    const Axes<R> A = axes_of(g, R(0));
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const uint16_t *cu = g.fcode, *cv = g.fcode + g.nu, *cw = g.fcode + g.nu + g.nv;
    const __amdgpu_buffer_rsrc_t codes = buf_rsrc(g.fcode, (g.nu + g.nv + g.nw) * 2u);
    const bool with_t = sp.buoyancy != 0;
    const R dt = (R)sp.dt, nu = (R)sp.nu;
    const int tk = threadIdx.x % TK, tj = threadIdx.x / TK;
    const int own = (tj + 2) * RK + tk + 2;  // the thread's cell in the LDS plane
    const MarchBufs mb = march_bufs(g, u, v, w, with_t ? T : (const R *)nullptr);
    const uint32_t ntiles = pt.ntk * pt.ntj * pt.nic;
    for (uint32_t tile = vb.blk; tile < ntiles; tile += vb.G) {
        const uint32_t rest = tile / pt.ntk;
        const int k0 = (int)(tile - rest * pt.ntk) * TK;
        const int j0 = (int)(rest % pt.ntj) * TJ;
        const int i0 = (int)(rest / pt.ntj * pt.ic);
        const int i1 = std::min(nx, i0 + (int)pt.ic);
        const int j = j0 + tj, k = k0 + tk;
        const bool act = j < ny && k < nz;
        SlotMap sm[SLOTS];
#pragma unroll
        for (int s = 0; s < SLOTS; s++) sm[s] = slot_map(g, j0, k0, threadIdx.x + s * kBlock);
        // y and z spacings of the column, fixed along the march
        const int ja = UNI ? 0 : (act ? j : 0), ka = UNI ? 0 : (act ? k : 0);
        const R hy = A.hy[ja], hz = A.hz[ka], ihy = A.ihy[ja], ihz = A.ihz[ka];
        const R i2y = A.i2y[ja], i2z = A.i2z[ka];
        // the column's values on the planes before the run (x neighbours)
        R um2 = R(0), um1 = R(0), vm1 = R(0), wm1 = R(0);
        if (act && i0 >= 2) um2 = u[uidx(g, i0 - 2, j, k)];
        if (act && i0 >= 1) {
            um1 = u[uidx(g, i0 - 1, j, k)];
            vm1 = v[vidx(g, i0 - 1, j, k)];
            wm1 = w[widx(g, i0 - 1, j, k)];
        }
        FaceExtras<R> ex = face_extras<R, SCHEME>(g, mb, codes, act, i0, j, k);
        // planes i0 and i0+1 in LDS, plane i0+2 in flight; each plane step
        // issues the loads of the plane three ahead, so two planes' loads are
        // in flight while one is computed
        PlaneRegs<R> ra, rb;
        plane_load(g, sm, i0, with_t, mb, ra);
        plane_store(ra, with_t, sU[i0 & 1], sV[i0 & 1], sW[i0 & 1], sT[i0 & 1]);
        plane_load(g, sm, i0 + 1, with_t, mb, rb);
        plane_store(rb, with_t, sU[(i0 + 1) & 1], sV[(i0 + 1) & 1], sW[(i0 + 1) & 1],
                    sT[(i0 + 1) & 1]);
        if (i0 + 2 <= i1) plane_load(g, sm, i0 + 2, with_t, mb, ra);
        __syncthreads();
        auto plane = [&](int i, PlaneRegs<R> &ld, const PlaneRegs<R> &pend) {
            if (i + 3 <= i1) plane_load(g, sm, i + 3, with_t, mb, ld);
            const FaceExtras<R> exn = face_extras<R, SCHEME>(g, mb, codes, act, i + 1, j, k);
            const R *Uc = sU[i & 1], *Vc = sV[i & 1], *Wc = sW[i & 1], *Tc = sT[i & 1];
            const R *Un = sU[(i + 1) & 1], *Vn = sV[(i + 1) & 1], *Wn = sW[(i + 1) & 1];
            const R uc = Uc[own], vc = Vc[own], wc = Wc[own];
            if (act) {
                const int ia = UNI ? 0 : i;
                const R h[3] = {A.hx[ia], hy, hz};
                const R ih[3] = {A.ihx[ia], ihy, ihz};
                const R i2[3] = {A.i2x[ia], i2y, i2z};
                const uint32_t eu = uidx(g, i, j, k), ev = vidx(g, i, j, k), ew = widx(g, i, j, k);
                // u face (i, j, k), FluidSolver.cpp:83-162
                R ru = uc;
                if (i > 0) {
                    R nb[6], q[3] = {R(0), R(0), R(0)}, vel[3];
                    nb[0] = i > 1 ? um1 : R(0);
                    nb[1] = i < nx - 1 ? Un[own] : R(0);
                    nb[2] = Uc[own - RK];
                    nb[3] = Uc[own + RK];
                    nb[4] = Uc[own - 1];
                    nb[5] = Uc[own + 1];
                    vel[0] = uc;
                    vel[1] = j > 0 ? R(0.5) * (vm1 + vc) : R(0);
                    vel[2] = k > 0 ? R(0.5) * (wm1 + wc) : R(0);
                    const bool qok = i >= 2 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 &&
                                     k <= nz - 2;
                    if (SCHEME == QUICK && qok) {
                        q[0] = um2;
                        q[1] = j >= 2 ? Uc[own - 2 * RK] : ex.wj;  // Q3: wrapped flat index
                        q[2] = k >= 2 ? Uc[own - 2] : ex.wk;       // Q3
                    }
                    ru = face_update_s<SCHEME, R>(qok, nb, q, uc, vel, h, ih, i2, dt, nu);
                }
                us[eu] = apply_mask<R>(g, (uint16_t)ex.cuv, 0, ru);
                // v face (i, j, k), :164-230
                R rv = vc;
                if (j > 0) {
                    R nb[6], q[3] = {R(0), R(0), R(0)}, vel[3];
                    vel[0] = i > 0 ? R(0.5) * (Uc[own - RK] + uc) : R(0);
                    nb[0] = i > 0 ? vm1 : R(0);
                    nb[1] = i < nx - 1 ? Vn[own] : R(0);
                    nb[2] = j > 1 ? Vc[own - RK] : R(0);
                    nb[3] = j < ny - 1 ? Vc[own + RK] : R(0);
                    vel[1] = vc;
                    vel[2] = k > 0 ? R(0.5) * (Wc[own - RK] + wc) : R(0);
                    nb[4] = Vc[own - 1];
                    nb[5] = Vc[own + 1];
                    const bool qok = i >= 1 && i <= nx - 2 && j >= 2 && j <= ny - 2 && k >= 1 &&
                                     k <= nz - 2;
                    if (SCHEME == QUICK && qok) {
                        q[0] = vm1;             // adjacent (:208)
                        q[1] = Vc[own - 2 * RK];
                        q[2] = Vc[own - 1];     // adjacent (:211)
                    }
                    rv = face_update_s<SCHEME, R>(qok, nb, q, vc, vel, h, ih, i2, dt, nu);
                }
                vs[ev] = apply_mask<R>(g, (uint16_t)(ex.cuv >> 16), 1, rv);
                // w face (i, j, k), :232-297, and buoyancy (:373-386)
                R rw = wc;
                if (k > 0) {
                    R nb[6], q[3] = {R(0), R(0), R(0)}, vel[3];
                    vel[0] = i > 0 ? R(0.5) * (Uc[own - 1] + uc) : R(0);
                    nb[0] = i > 0 ? wm1 : R(0);
                    nb[1] = i < nx - 1 ? Wn[own] : R(0);
                    vel[1] = j > 0 ? R(0.5) * (Vc[own - 1] + vc) : R(0);
                    nb[2] = Wc[own - RK];
                    nb[3] = Wc[own + RK];
                    nb[4] = k > 1 ? Wc[own - 1] : R(0);
                    nb[5] = k < nz - 1 ? Wc[own + 1] : R(0);
                    vel[2] = wc;
                    const bool qok = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 2 &&
                                     k <= nz - 2;
                    if (SCHEME == QUICK && qok) {
                        q[0] = wm1;             // adjacent (:275)
                        q[1] = Wc[own - RK];    // adjacent (:276)
                        q[2] = Wc[own - 2];
                    }
                    rw = face_update_s<SCHEME, R>(qok, nb, q, wc, vel, h, ih, i2, dt, nu);
                    if (with_t)
                        rw += (R)sp.buoy_coeff * (R(0.5) * (Tc[own - 1] + Tc[own]) - (R)sp.T_ref);
                }
                ws[ew] = apply_mask<R>(g, (uint16_t)ex.cw, 2, rw);
                // the boundary faces past the last cells: copied, then masked
                if (j == ny - 1) {
                    const uint32_t e = ev + (uint32_t)nz;
                    vs[e] = apply_mask<R>(g, cv[e], 1, Vc[own + RK]);
                }
                if (k == nz - 1) {
                    const uint32_t e = ew + 1u;
                    ws[e] = apply_mask<R>(g, cw[e], 2, Wc[own + 1]);
                }
                if (i == nx - 1) {
                    const uint32_t e = eu + (uint32_t)ny * nz;
                    us[e] = apply_mask<R>(g, cu[e], 0, Un[own]);
                }
            }
            um2 = um1;
            um1 = uc;
            vm1 = vc;
            wm1 = wc;
            ex = exn;
            __syncthreads();  // plane i is read by every thread
            if (i + 2 <= i1) plane_store(pend, with_t, sU[i & 1], sV[i & 1], sW[i & 1], sT[i & 1]);
            __syncthreads();
        };
        for (int i = i0; i < i1; i += 2) {  // the two register sets alternate
            plane(i, rb, ra);
            if (i + 1 < i1) plane(i + 1, ra, rb);
        }
    }
}
};
template <typename R, bool UNI, int SCHEME>
__global__ void __launch_bounds__(kBlock) k_predict(Geo g, StepParams sp, PlaneTiles pt, const R *u,
                                                    const R *v, const R *w, const R *T, R *us,
                                                    R *vs, R *ws, StepRec *rec)
{
    KB_predict<R, UNI, SCHEME>::run(vb_here(), g, sp, pt, u, v, w, T, us, vs, ws, rec);
}

template <typename R, bool UNI, int SCHEME>
static void launch_predict_s(hipStream_t s, const Geo &g, const StepParams &sp, const R *u,
                             const R *v, const R *w, const R *T, R *us, R *vs, R *ws, StepRec *rec)
{
    const PlaneTiles pt = plane_tiles(g, 8);
    FFR_GANG((KB_predict<R, UNI, SCHEME>), GB_PLAIN, (k_predict<R, UNI, SCHEME>),
             dim3(plane_tiles_grid(pt)), dim3(kBlock), 0, s, g, sp, pt, u, v, w, T, us, vs, ws, rec);
}

template <typename R, bool UNI>
static void launch_predict_u(hipStream_t s, const Geo &g, const StepParams &sp, const R *u,
                             const R *v, const R *w, const R *T, R *us, R *vs, R *ws, StepRec *rec)
{
    if (sp.scheme == QUICK)
        launch_predict_s<R, UNI, QUICK>(s, g, sp, u, v, w, T, us, vs, ws, rec);
    else if (sp.scheme == LAX_WENDROFF)
        launch_predict_s<R, UNI, LAX_WENDROFF>(s, g, sp, u, v, w, T, us, vs, ws, rec);
    else
        launch_predict_s<R, UNI, UPWIND1>(s, g, sp, u, v, w, T, us, vs, ws, rec);
}

template <typename R>
void launch_predict(hipStream_t s, const Geo &g, const StepParams &sp, const R *u, const R *v,
                    const R *w, const R *T, R *us, R *vs, R *ws, StepRec *rec)
{
    if (g.uniform) launch_predict_u<R, true>(s, g, sp, u, v, w, T, us, vs, ws, rec);
    else launch_predict_u<R, false>(s, g, sp, u, v, w, T, us, vs, ws, rec);
}

// ---- K4: Poisson right-hand side ----------------------------------------
template <typename R>
struct KB_rhs {
    // two cells per thread from the block's chunks (xcd_chunks: the u faces
    // one plane up are then read again from L2), every load of both issued
    // before either is used; a SOLID cell's loads are made and dropped
    static constexpr int U = 2;
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const StepParams &sp,
                                               const R *__restrict__ us, const R *__restrict__ vs,
                                               const R *__restrict__ ws, double *__restrict__ b)
{
    const uint32_t si = (uint32_t)g.ny * g.nz, nz = g.nz;
    uint32_t ch0, cst, cend;
    xcd_chunks(g.nc, vb.blk, vb.G, ch0, cst, cend);
    for (uint32_t ch = ch0; ch < cend; ch += U * cst) {
        uint32_t cc[U], ii[U], jj[U], kk[U];
        bool in[U];
        uint8_t fl[U];
        R u0[U], u1[U], v0[U], v1[U], w0[U], w1[U];
#pragma unroll
        for (int t = 0; t < U; t++) {
            const uint32_t ct = (ch + t * cst) * kBlock + threadIdx.x;
            in[t] = ch + t * cst < cend && ct < g.nc;
            const uint32_t c = in[t] ? ct : 0u;
            const uint32_t i = fdiv(c, g.f_nynz), rem = c - i * si;
            const uint32_t j = fdiv(rem, g.f_nz), k = rem - j * nz;
            const uint32_t ev = c + i * nz, ew = c + i * (uint32_t)g.ny + j;
            cc[t] = c; ii[t] = i; jj[t] = j; kk[t] = k;
            fl[t] = g.flags[c];
            u0[t] = us[c]; u1[t] = us[c + si];
            v0[t] = vs[ev]; v1[t] = vs[ev + nz];
            w0[t] = ws[ew]; w1[t] = ws[ew + 1];
        }
#pragma unroll
        for (int t = 0; t < U; t++) {
            if (!in[t]) continue;
            if (fl[t] == SOLID) {
                b[cc[t]] = 0.0;
                continue;
            }
            const double du = (double)u1[t] - (double)u0[t];
            const double dv = (double)v1[t] - (double)v0[t];
            const double dw = (double)w1[t] - (double)w0[t];
            const double div = by_h64<R>(du, g.hx[ii[t]], g.ihx[ii[t]]) +
                               by_h64<R>(dv, g.hy[jj[t]], g.ihy[jj[t]]) +
                               by_h64<R>(dw, g.hz[kk[t]], g.ihz[kk[t]]);
            b[cc[t]] = sp.rhs_factor * div;
        }
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

// One thread per cell: the low u, v and w faces of cell (i, j, k) (flat
// indices c, c + i nz and c + i ny + j), so p[c] is read once for its three
// faces and the x, y, z neighbours below come from the same or a recent
// plane; the cells of the last plane, row and column also take the high
// boundary face of their axis. Each face is computed exactly as corrected()
// does; an outflow face copies the corrected face one step inside through
// the general path.
template <typename R>
struct KB_correct {
    static __device__ __forceinline__ R face(const Geo &g, const StepParams &sp, const R *fs,
                                             const double *p, int axis, int i, int j, int k,
                                             uint32_t e, uint32_t step, uint16_t code, R x,
                                             double pc, double pm)
    {
        if (code & (FC_OUT_LO | FC_OUT_HI)) {  // outflow patch: the face one step inside
            const int d = (code & FC_OUT_LO) ? 1 : -1;
            const uint32_t e2 = d > 0 ? e + step : e - step;
            const uint16_t *cod = g.fcode + (axis >= 1 ? g.nu : 0u) + (axis == 2 ? g.nv : 0u);
            return corrected<R>(g, sp, fs, p, axis, i + d * (axis == 0), j + d * (axis == 1),
                                k + d * (axis == 2), e2, cod[e2]);
        }
        if (!(code & FC_BOUND)) {
            const double gp = axis == 0   ? by_h64<R>(pc - pm, g.cx[i], g.icx[i])
                              : axis == 1 ? by_h64<R>(pc - pm, g.cy[j], g.icy[j])
                                          : by_h64<R>(pc - pm, g.cz[k], g.icz[k]);
            x = x - (R)(sp.dt_rho * gp);
        }
        return apply_mask<R>(g, code, axis, x);
    }

    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const StepParams &sp,
                                               const R *__restrict__ us, const R *__restrict__ vs,
                                               const R *__restrict__ ws, const double *__restrict__ p,
                                               R *__restrict__ u, R *__restrict__ v,
                                               R *__restrict__ w, StepRec *rec)
{
    __shared__ double lds[kBlock / kWave];
    double dv = 0.0;
    bool bad = false;
    const uint32_t si = (uint32_t)g.ny * g.nz, nz = (uint32_t)g.nz;
    const uint16_t *cu = g.fcode, *cv = g.fcode + g.nu, *cw = g.fcode + g.nu + g.nv;
    auto put = [&](R *out, uint32_t e, R val, R old) {
        out[e] = val;
        dv = fmax(dv, fabs((double)val - (double)old));  // fmax drops NaN as std::max does
        bad |= !isfinite((double)val);
    };
    uint32_t ch0, cstep, cend;
    xcd_chunks(g.nc, vb.blk, vb.G, ch0, cstep, cend);
    for (uint32_t ch = ch0; ch < cend; ch += cstep) {
        const uint32_t c = ch * kBlock + threadIdx.x;
        if (c >= g.nc) break;
        const uint32_t i = fdiv(c, g.f_nynz), rem = c - i * si;
        const uint32_t j = fdiv(rem, g.f_nz), k = rem - j * nz;
        const uint32_t ev = c + i * nz, ew = c + i * (uint32_t)g.ny + j;
        // every load of the cell's three faces before any use (neighbours
        // below a wall read the cell itself: those faces are boundary faces)
        const uint16_t ku = cu[c], kv = cv[ev], kw = cw[ew];
        const double pc = p[c], pxm = p[i ? c - si : c], pym = p[j ? c - nz : c],
                     pzm = p[k ? c - 1 : c];
        const R xu = us[c], xv = vs[ev], xw = ws[ew];
        const R ou = u[c], ov = v[ev], ow = w[ew];
        // the high w face of the column's last cell (a row end falls in most
        // waves, so its loads go out with the others)
        const bool top = k + 1 == nz;
        const uint32_t eh = top ? ew + 1 : ew;
        const uint16_t kh = cw[eh];
        const R xh = ws[eh], oh = w[eh];
        const R fu = face(g, sp, us, p, 0, i, j, k, c, si, ku, xu, pc, pxm);
        const R fv = face(g, sp, vs, p, 1, i, j, k, ev, nz, kv, xv, pc, pym);
        const R fw = face(g, sp, ws, p, 2, i, j, k, ew, 1u, kw, xw, pc, pzm);
        put(u, c, fu, ou);
        put(v, ev, fv, ov);
        put(w, ew, fw, ow);
        // the high boundary faces (no gradient; masks or an outflow copy)
        if (top) put(w, eh, face(g, sp, ws, p, 2, i, j, k + 1, eh, 1u, kh, xh, 0.0, 0.0), oh);
        if (i + 1 == (uint32_t)g.nx) {
            const uint32_t e = c + si;
            put(u, e, face(g, sp, us, p, 0, i + 1, j, k, e, si, cu[e], us[e], 0.0, 0.0), u[e]);
        }
        if (j + 1 == (uint32_t)g.ny) {
            const uint32_t e = ev + nz;
            put(v, e, face(g, sp, vs, p, 1, i, j + 1, k, e, nz, cv[e], vs[e], 0.0, 0.0), v[e]);
        }
    }
    if (bad) rec->nonfinite = 1;
    const double m = block_max(dv, lds);
    if (threadIdx.x == 0 && m > 0.0) atomic_max_nonneg(&rec->vmax_bits, m);
}
};
template <typename R>
__global__ void __launch_bounds__(kBlock) k_correct(Geo g, StepParams sp, const R *us,
                                                    const R *vs, const R *ws, const double *p,
                                                    R *u, R *v, R *w, StepRec *rec)
{
    KB_correct<R>::run(vb_here(), g, sp, us, vs, ws, p, u, v, w, rec);
}

template <typename R>
void launch_correct(hipStream_t s, const Geo &g, const StepParams &sp, const R *us, const R *vs,
                    const R *ws, const double *p, R *u, R *v, R *w, StepRec *rec)
{
    FFR_GANG((KB_correct<R>), GB_PLAIN, k_correct<R>, dim3(grid_for(g.nc)), dim3(kBlock), 0, s,
             g, sp, us, vs, ws, p, u, v, w, rec);
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
// The same march as the predictor: the tile's plane of T and of the cell
// flags, halo included, in LDS; the column's T and flags of the two planes
// behind and the u face ahead in registers. A neighbour outside the domain
// is staged as a SOLID cell, so the reference's Neumann rule - the cell's own
// T for an outside or SOLID neighbour (:417-422) - is one test. The cell's
// v and w faces are read directly (the pair of a face and the next one is
// one coalesced run of the array).
template <typename R>
__device__ __forceinline__ R t_diffusion(R xm, R xp, R ym, R yp, R zm, R zp, R c, const R h[3],
                                         const R i2[3], R al)
{
    return al * (by_h2(xm - R(2) * c + xp, h[0], i2[0]) + by_h2(ym - R(2) * c + yp, h[1], i2[1]) +
                 by_h2(zm - R(2) * c + zp, h[2], i2[2]));
}
__device__ __forceinline__ float t_diffusion(float xm, float xp, float ym, float yp, float zm,
                                             float zp, float c, const float *, const float i2[3],
                                             float al)
{
    float acc = i2[0] * fmaf(-2.0f, c, xm + xp);
    acc = fmaf(i2[1], fmaf(-2.0f, c, ym + yp), acc);
    acc = fmaf(i2[2], fmaf(-2.0f, c, zm + zp), acc);
    return al * acc;
}

template <typename R>
__device__ __forceinline__ void cell_plane_load(const Geo &g, const SlotMap *sm, int p,
                                                const MarchBufs &b, R *t, uint8_t *f)
{
    const bool pc = p >= 0 && p < g.nx;
    const uint32_t base = (uint32_t)p * g.ny * g.nz;
#pragma unroll
    for (int s = 0; s < ptile::SLOTS; s++) {
        const bool in = pc && sm[s].in_u;
        const uint32_t c = base + sm[s].ou;
        t[s] = buf_ld(b.t, in ? c * (uint32_t)sizeof(R) : kNoElement, R(0));
        const uint8_t fl = __builtin_amdgcn_raw_buffer_load_b8(b.f, (int)(in ? c : kNoElement), 0, 0);
        f[s] = in ? fl : (uint8_t)SOLID;
    }
}

template <typename R>
__device__ __forceinline__ void cell_plane_store(const R *t, const uint8_t *f, R *st, uint8_t *sf)
{
#pragma unroll
    for (int s = 0; s < ptile::SLOTS; s++) {
        const int e = threadIdx.x + s * kBlock;
        if (s + 1 < ptile::SLOTS || e < ptile::PLANE) {
            st[e] = t[s];
            sf[e] = f[s];
        }
    }
}

// The faces of the thread's cell on plane p that transport T: u on its high
// side (the low one is the previous plane's), v and w on both sides. Loaded
// one plane ahead.
template <typename R>
struct CellFaces {
    R uh, v0, v1, w0, w1;
};
template <typename R>
__device__ __forceinline__ CellFaces<R> cell_faces(const Geo &g, const MarchBufs &b, bool act,
                                                   int p, int j, int k)
{
    CellFaces<R> f;
    const bool in = act && p < g.nx;
    constexpr uint32_t sz = sizeof(R);
    const uint32_t eu = uidx(g, p + 1, j, k), ev = vidx(g, p, j, k), ew = widx(g, p, j, k);
    f.uh = buf_ld(b.u, in ? eu * sz : kNoElement, R(0));
    f.v0 = buf_ld(b.v, in ? ev * sz : kNoElement, R(0));
    f.v1 = buf_ld(b.v, in ? (ev + (uint32_t)g.nz) * sz : kNoElement, R(0));
    f.w0 = buf_ld(b.w, in ? ew * sz : kNoElement, R(0));
    f.w1 = buf_ld(b.w, in ? (ew + 1u) * sz : kNoElement, R(0));
    return f;
}

template <typename R, bool UNI, int SCHEME>
struct KB_temperature {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const StepParams &sp,
                                               const PlaneTiles pt, const R *u, const R *v,
                                               const R *w, const R *T, R *Tn, StepRec *rec,
                                               const RedScratch &red)
{
    using namespace ptile;
    __shared__ R sT[2][PLANE];
    __shared__ uint8_t sF[2][PLANE];
    __shared__ double lds[kBlock / kWave];
    double dmax = 0.0, tsum = 0.0;
    bool bad = false;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const Axes<R> A = axes_of(g, R(0));
    const R dt = (R)sp.dt, al = (R)sp.alpha_T;
    const int tk = threadIdx.x % TK, tj = threadIdx.x / TK;
    const int own = (tj + 2) * RK + tk + 2;
    const MarchBufs mb = march_bufs(g, u, v, w, T);
    const uint32_t ntiles = pt.ntk * pt.ntj * pt.nic;
    for (uint32_t tile = vb.blk; tile < ntiles; tile += vb.G) {
        const uint32_t rest = tile / pt.ntk;
        const int k0 = (int)(tile - rest * pt.ntk) * TK;
        const int j0 = (int)(rest % pt.ntj) * TJ;
        const int i0 = (int)(rest / pt.ntj * pt.ic);
        const int i1 = std::min(nx, i0 + (int)pt.ic);
        const int j = j0 + tj, k = k0 + tk;
        const bool act = j < ny && k < nz;
        SlotMap sm[SLOTS];
#pragma unroll
        for (int s = 0; s < SLOTS; s++) sm[s] = slot_map(g, j0, k0, threadIdx.x + s * kBlock);
        const int ja = UNI ? 0 : (act ? j : 0), ka = UNI ? 0 : (act ? k : 0);
        const R hy = A.hy[ja], hz = A.hz[ka], ihy = A.ihy[ja], ihz = A.ihz[ka];
        const R i2y = A.i2y[ja], i2z = A.i2z[ka];
        R tm2 = R(0), tm1 = R(0), ul = R(0);
        uint8_t fm2 = SOLID, fm1 = SOLID;
        if (act) {
            if (i0 >= 2) {
                tm2 = T[cidx(g, i0 - 2, j, k)];
                fm2 = g.flags[cidx(g, i0 - 2, j, k)];
            }
            if (i0 >= 1) {
                tm1 = T[cidx(g, i0 - 1, j, k)];
                fm1 = g.flags[cidx(g, i0 - 1, j, k)];
            }
            ul = u[uidx(g, i0, j, k)];
        }
        CellFaces<R> cf = cell_faces<R>(g, mb, act, i0, j, k);
        R nt[SLOTS];
        uint8_t nf[SLOTS];
        cell_plane_load(g, sm, i0, mb, nt, nf);
        cell_plane_store(nt, nf, sT[i0 & 1], sF[i0 & 1]);
        cell_plane_load(g, sm, i0 + 1, mb, nt, nf);
        cell_plane_store(nt, nf, sT[(i0 + 1) & 1], sF[(i0 + 1) & 1]);
        __syncthreads();
        for (int i = i0; i < i1; i++) {
            const bool more = i + 2 <= i1;
            if (more) cell_plane_load(g, sm, i + 2, mb, nt, nf);
            const CellFaces<R> cfn = cell_faces<R>(g, mb, act, i + 1, j, k);
            const R *Tc_ = sT[i & 1], *Tn_ = sT[(i + 1) & 1];
            const uint8_t *Fc = sF[i & 1], *Fn = sF[(i + 1) & 1];
            const R tc = Tc_[own];
            const uint8_t fc = Fc[own];
            const R uh = cf.uh;  // u(i+1, j, k)
            if (act) {
                const uint32_t c = cidx(g, i, j, k);
                if (fc != FLUID) {  // non-FLUID cells are frozen (:402-405)
                    Tn[c] = tc;
                } else {
                    // Neumann at SOLID and outside the domain (:417-422)
                    auto tn = [&](R t, uint8_t f) -> R { return f == SOLID ? tc : t; };
                    const R xm = tn(tm1, fm1), xp = tn(Tn_[own], Fn[own]);
                    const R ym = tn(Tc_[own - RK], Fc[own - RK]), yp = tn(Tc_[own + RK], Fc[own + RK]);
                    const R zm = tn(Tc_[own - 1], Fc[own - 1]), zp = tn(Tc_[own + 1], Fc[own + 1]);
                    const int ia = UNI ? 0 : i;
                    const R h[3] = {A.hx[ia], hy, hz};
                    const R ih[3] = {A.ihx[ia], ihy, ihz};
                    const R i2[3] = {A.i2x[ia], i2y, i2z};
                    const R u0 = ul, u1 = uh, v0 = cf.v0, v1 = cf.v1, w0 = cf.w0, w1 = cf.w1;
                    const R uc = R(0.5) * (u0 + u1), vc = R(0.5) * (v0 + v1), wc = R(0.5) * (w0 + w1);
                    R adv;
                    if (SCHEME == QUICK && i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 &&
                        k >= 1 && k <= nz - 2) {
                        const R xmm = tn(tm2, fm2), ymm = tn(Tc_[own - 2 * RK], Fc[own - 2 * RK]);
                        const R zmm = tn(Tc_[own - 2], Fc[own - 2]);
                        adv = adv_quick(xmm, xm, tc, xp, uc, h[0], ih[0]) +
                              adv_quick(ymm, ym, tc, yp, vc, h[1], ih[1]) +
                              adv_quick(zmm, zm, tc, zp, wc, h[2], ih[2]);
                    } else if (SCHEME == LAX_WENDROFF) {
                        adv = adv_lw(xm, tc, xp, uc, dt, h[0], ih[0]) +
                              adv_lw(ym, tc, yp, vc, dt, h[1], ih[1]) +
                              adv_lw(zm, tc, zp, wc, dt, h[2], ih[2]);
                    } else {
                        adv = adv_upwind(xm, tc, xp, uc, h[0], ih[0]) +
                              adv_upwind(ym, tc, yp, vc, h[1], ih[1]) +
                              adv_upwind(zm, tc, zp, wc, h[2], ih[2]);
                    }
                    const R dif = t_diffusion(xm, xp, ym, yp, zm, zp, tc, h, i2, al);
                    const R tnew = advance(tc, dt, adv, dif);
                    Tn[c] = tnew;
                    tsum += (double)tnew;
                    bad |= !isfinite((double)tnew);
                    const int iy = UNI ? 0 : j, iz = UNI ? 0 : k;
                    dmax = fmax(dmax, fabs(by_h64<R>((double)u1 - (double)u0, g.hx[ia], g.ihx[ia]) +
                                           by_h64<R>((double)v1 - (double)v0, g.hy[iy], g.ihy[iy]) +
                                           by_h64<R>((double)w1 - (double)w0, g.hz[iz], g.ihz[iz])));
                }
            }
            tm2 = tm1;
            tm1 = tc;
            fm2 = fm1;
            fm1 = fc;
            ul = uh;
            cf = cfn;
            __syncthreads();
            if (more) cell_plane_store(nt, nf, sT[i & 1], sF[i & 1]);
            __syncthreads();
        }
    }
    if (bad) rec->nonfinite = 1;
    const double m = block_max(dmax, lds);
    if (threadIdx.x == 0 && m > 0.0) atomic_max_nonneg(&rec->divmax_bits, m);
    block_partial(tsum, red.partials + SLOT_TSUM * kMaxBlocksCap, vb.blk, lds);
}
};
template <typename R, bool UNI, int SCHEME>
__global__ void __launch_bounds__(kBlock) k_temperature(Geo g, StepParams sp, PlaneTiles pt,
                                                        const R *u, const R *v, const R *w,
                                                        const R *T, R *Tn, StepRec *rec,
                                                        RedScratch red)
{
    KB_temperature<R, UNI, SCHEME>::run(vb_here(), g, sp, pt, u, v, w, T, Tn, rec, red);
}

template <typename R, bool UNI, int SCHEME>
static unsigned launch_temperature_s(hipStream_t s, const Geo &g, const StepParams &sp,
                                     const R *u, const R *v, const R *w, const R *T, R *Tn,
                                     StepRec *rec, RedScratch red)
{
    const PlaneTiles pt = plane_tiles(g, 8);
    const unsigned G = plane_tiles_grid(pt);
    FFR_GANG((KB_temperature<R, UNI, SCHEME>), GB_PLAIN, (k_temperature<R, UNI, SCHEME>), dim3(G),
             dim3(kBlock), 0, s, g, sp, pt, u, v, w, T, Tn, rec, red);
    return G;
}

template <typename R, bool UNI>
static unsigned launch_temperature_u(hipStream_t s, const Geo &g, const StepParams &sp,
                                     const R *u, const R *v, const R *w, const R *T, R *Tn,
                                     StepRec *rec, RedScratch red)
{
    if (sp.scheme == QUICK)
        return launch_temperature_s<R, UNI, QUICK>(s, g, sp, u, v, w, T, Tn, rec, red);
    if (sp.scheme == LAX_WENDROFF)
        return launch_temperature_s<R, UNI, LAX_WENDROFF>(s, g, sp, u, v, w, T, Tn, rec, red);
    return launch_temperature_s<R, UNI, UPWIND1>(s, g, sp, u, v, w, T, Tn, rec, red);
}

template <typename R>
unsigned launch_temperature(hipStream_t s, const Geo &g, const StepParams &sp, const R *u,
                            const R *v, const R *w, const R *T, R *Tn, StepRec *rec,
                            RedScratch red)
{
    if (g.uniform) return launch_temperature_u<R, true>(s, g, sp, u, v, w, T, Tn, rec, red);
    return launch_temperature_u<R, false>(s, g, sp, u, v, w, T, Tn, rec, red);
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
unsigned launch_tsum(hipStream_t s, const Geo &g, const R *T, StepRec *rec, RedScratch red)
{
    const unsigned G = grid_for(g.nc);
    FFR_GANG((KB_tsum<R>), GB_PLAIN, k_tsum<R>, dim3(G), dim3(kBlock), 0, s, g, T, rec, red);
    return G;
}

// rec->tsum from the per-block partials of temperature / tsum (one block);
// G is the block count of the launch that wrote them.
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

void launch_step_finalize(hipStream_t s, StepRec *rec, RedScratch red, unsigned G)
{
    FFR_GANG((KB_step_finalize), GB_PLAIN, k_step_finalize, dim3(1), dim3(kBlock), 0, s, rec, red,
             G);
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
    template unsigned launch_temperature<R>(hipStream_t, const Geo &, const StepParams &,      \
                                            const R *, const R *, const R *, const R *, R *,   \
                                            StepRec *, RedScratch);                            \
    template unsigned launch_tsum<R>(hipStream_t, const Geo &, const R *, StepRec *,           \
                                     RedScratch);                                              \
    template void launch_opening_T<R>(hipStream_t, const Geo &, const R *, const R *,          \
                                      const R *, R *);                                         \
    template void launch_rhs<R>(hipStream_t, const Geo &, const StepParams &, const R *,       \
                                const R *, const R *, double *);                               \
    template void launch_convert<R>(hipStream_t, const double *, R *, uint64_t);               \
    template void launch_convert_back<R>(hipStream_t, const R *, double *, uint64_t);

FFR_INSTANTIATE(double)
FFR_INSTANTIATE(float)

}  // namespace ffr
