// mg_kernels.hip - kernels of the multigrid-preconditioned pressure solve
// (see ffr_mg.hpp for the method). Level-0 vectors are FP32; coarser levels
// store z and r as TC (float or bfloat16) and compute in FP32. Dot products
// accumulate in FP64 and are reduced deterministically (block partials, fixed
// order). Every kernel of the inner iteration returns at once when the
// solve or the inner iteration has stopped, so the host may enqueue ahead.
#include "../ffr_kernels.hpp"
#include "../ffr_gang.hpp"
#include "../ffr_mg.hpp"
#include "../ffr_mg_launch.hpp"

namespace ffr {

__device__ __forceinline__ float ldv(const float *p, uint32_t i) { return p[i]; }
__device__ __forceinline__ float ldv(const __hip_bfloat16 *p, uint32_t i) { return __bfloat162float(p[i]); }
__device__ __forceinline__ void stv(float *p, uint32_t i, float v) { p[i] = v; }
__device__ __forceinline__ void stv(__hip_bfloat16 *p, uint32_t i, float v) { p[i] = __float2bfloat16(v); }

// vector element access through a buffer resource (float or bfloat16)
__device__ __forceinline__ float bldv(__amdgpu_buffer_rsrc_t r, uint32_t i, const float *)
{
    return buf_ldf(r, i);
}
__device__ __forceinline__ float bldv(__amdgpu_buffer_rsrc_t r, uint32_t i, const __hip_bfloat16 *)
{
    return __bfloat162float(__builtin_bit_cast(__hip_bfloat16, buf_ldu16(r, i)));
}
__device__ __forceinline__ void bstv(__amdgpu_buffer_rsrc_t r, uint32_t i, float v, const float *)
{
    buf_stf(r, i, v);
}
__device__ __forceinline__ void bstv(__amdgpu_buffer_rsrc_t r, uint32_t i, float v, const __hip_bfloat16 *)
{
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, __float2bfloat16(v)), r,
                                          (int)(i * 2u), 0, 0);
}

// x rounded to the storage type T and back (what a store and a load give)
__device__ __forceinline__ float rnd(float x, const float *) { return x; }
__device__ __forceinline__ float rnd(float x, const __hip_bfloat16 *)
{
    return __bfloat162float(__float2bfloat16(x));
}

__device__ __forceinline__ bool unknown0(uint8_t f) { return f == FLUID || f == INFLOW; }
__device__ __forceinline__ double *pslot(RedScratch red, int k) { return red.partials + k * kMaxBlocksCap; }

__device__ __forceinline__ bool stopped(const MgState *m) { return m->solve_done || m->inner_done; }

// Deferred stop test: the flags are read at kernel entry with vector loads
// (which complete in order, unlike scalar loads that the kernel-argument wait
// would also cover), and the empty asm makes the value opaque where it is
// tested, so the compiler cannot hoist the branch (or unswitch a loop on it)
// above the data loads issued before that point: the flag's memory latency
// overlaps the data's instead of preceding it. The value is uniform.
__device__ __forceinline__ uint32_t stop_flags(const MgState *m)
{
    const auto r = buf_rsrc(m, sizeof(MgState));
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)offsetof(MgState, solve_done), 0, 0) |
           __builtin_amdgcn_raw_buffer_load_b32(r, (int)offsetof(MgState, inner_done), 0, 0);
}
__device__ __forceinline__ bool late(uint32_t v)
{
    asm volatile("" : "+v"(v));
    return __builtin_amdgcn_readfirstlane(v) != 0;
}

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

// ---- level-0 operator of a stretched grid (ST) ------------------------------
// The reference row (PoissonSolver.cpp:34-49) scales face f of cell c by
// 1/(h_f d_c), h_f the centre distance across the face and d_c the cell's
// width along its axis, so the two cells of a face see different scales and
// the operator is not symmetric. Multiplied by the cell's volume relative to
// the mean cell's (w_x w_y w_z, w = d / mean width of the axis), a row's face
// coefficient becomes w_b w_c / (h_f m_a) (the other two axes' relative
// widths, m_a the face axis' mean width): the same from both sides. The inner
// solve runs on that symmetric, volume-weighted form with the outer residual
// weighted alike, which has the same solution; Geo.mgf holds, per index of
// each axis, {w, t below, t above} with t = 1/(h_f m_a) (0 at the walls), so
// a cell's couplings are t w w from three 16-byte loads.
template <bool ST>
struct Cof {
    float f[6];  // -x +x -y +y -z +z
};
template <>
struct Cof<false> {};  // uniform spacing: the per-axis constants of Geo

template <bool ST>
__device__ __forceinline__ Cof<ST> cof_at(const Geo &g, uint32_t c)
{
    Cof<ST> r;
    if constexpr (ST) {
        c = c < g.nc ? c : g.nc - 1;  // slots past the end: any in-range cell
        const uint32_t i = fdiv(c, g.f_nynz), rem = c - i * (uint32_t)(g.ny * g.nz);
        const uint32_t j = fdiv(rem, g.f_nz), k = rem - j * (uint32_t)g.nz;
        const float4 X = g.mgf[i], Y = g.mgf[g.nx + j], Z = g.mgf[g.nx + g.ny + k];
        const float yz = Y.x * Z.x, xz = X.x * Z.x, xy = X.x * Y.x;
        r.f[0] = X.y * yz;
        r.f[1] = X.z * yz;
        r.f[2] = Y.y * xz;
        r.f[3] = Y.z * xz;
        r.f[4] = Z.y * xy;
        r.f[5] = Z.z * xy;
    }
    return r;
}
// weight of face q in the diagonal: 1 to an unknown neighbour, 2 to an
// OUTFLOW one (PoissonSolver.cpp:52-87), 0 across a wall or to SOLID
__device__ __forceinline__ float face_cnt(uint16_t m, int q)
{
    return (m & (NB_UNK0 << q)) ? 1.f : ((m & (NB_OUT0 << q)) ? 2.f : 0.f);
}

// ---- level-0 split layout (see ffr_mg.hpp) -----------------------------------
// position of natural cell c (natural code m) in its own colour's array / the
// base of the other colour's array
__device__ __forceinline__ uint32_t sown(uint32_t c, uint16_t m, uint32_t H)
{
    return ((m & NB_COLOR) ? H : 0u) + (c >> 1);
}
__device__ __forceinline__ uint32_t sother(uint16_t m, uint32_t H) { return (m & NB_COLOR) ? 0u : H; }

// neighbour sum of natural cell c from the other colour's array xo
__device__ __forceinline__ float off0s(const Geo &g, const float *__restrict__ xo, uint32_t c,
                                       uint16_t m)
{
    const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz;
    float ox = 0.f, oy = 0.f, oz = 0.f;
    if (m & (NB_UNK0 << 0)) ox += xo[(c - si) >> 1];
    if (m & (NB_UNK0 << 1)) ox += xo[(c + si) >> 1];
    if (m & (NB_UNK0 << 2)) oy += xo[(c - sj) >> 1];
    if (m & (NB_UNK0 << 3)) oy += xo[(c + sj) >> 1];
    if (m & (NB_UNK0 << 4)) oz += xo[(c - 1) >> 1];
    if (m & (NB_UNK0 << 5)) oz += xo[(c + 1) >> 1];
    return g.sx * ox + g.sy * oy + g.sz * oz;
}

// split codes: nbs[colour*H + (c>>1)] = code of c with NB_COLOR replaced by
// c & 1 (so the natural index is recovered as 2h + bit); unused slot stays 0
__global__ void __launch_bounds__(kBlock) k_mg_split_codes(Geo g, uint16_t *nbs, uint32_t H)
{
    uint32_t sb, se;
    block_span(g.nc, xcd_block(), sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        const uint16_t m = g.nbm[c];
        nbs[sown(c, m, H)] = (uint16_t)((m & ~NB_COLOR) | ((c & 1u) ? NB_COLOR : 0u));
    }
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

// ---- hierarchy construction: A_{l+1} = P^T A_l P, per-axis scaled -----------
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
// diagonal part along `axis` of an unknown fine cell: its couplings plus
// 2 * scale per OUTFLOW neighbour (level 0, PoissonSolver.cpp:52-87), or the
// stored part (coarser levels)
__device__ __forceinline__ float dpart0(const Geo &g, uint16_t m, int axis)
{
    const uint32_t two = 3u << (2 * axis);
    const float cnt = (float)(__popc(m & (two << 1)) + 2 * __popc(m & (two << 7)));
    return (axis == 0 ? g.sx : (axis == 1 ? g.sy : g.sz)) * cnt;
}

// One thread per coarse cell: sums over the aggregate's children of the
// diagonal parts (minus twice the couplings inside the aggregate) and of the
// couplings across its lower faces, per axis, then scales each axis by 1/2
// when the coarse level halved it.
template <bool FINE0, bool ST>
__global__ void __launch_bounds__(kBlock) k_mg_build(Geo g, MgLevel F, MgLevel C)
{
    uint32_t sb, se;
    block_span(C.n, xcd_block(), sb, se);
    for (uint32_t I = sb + threadIdx.x; I < se; I += kBlock) {
        int ic, jc, kc;
        decode(I, C.f_nynz, C.f_nz, C.ny, C.nz, ic, jc, kc);
        const int i0 = ic << C.shx, i1 = min(i0 + (1 << C.shx) - 1, F.nx - 1);
        const int j0 = jc << C.shy, j1 = min(j0 + (1 << C.shy) - 1, F.ny - 1);
        const int k0 = kc << C.shz, k1 = min(k0 + (1 << C.shz) - 1, F.nz - 1);
        float D[3] = {0.f, 0.f, 0.f}, inner[3] = {0.f, 0.f, 0.f}, low[3] = {0.f, 0.f, 0.f};
        bool any = false;
        for (int a = i0; a <= i1; a++)
            for (int b = j0; b <= j1; b++)
                for (int d = k0; d <= k1; d++) {
                    bool act;
                    float dp[3];
                    if constexpr (FINE0 && ST) {
                        const uint32_t c = cidx(g, a, b, d);
                        const uint16_t m = g.nbm[c];
                        act = (m & NB_SELF) != 0;
                        const Cof<ST> cf = cof_at<ST>(g, c);
#pragma unroll
                        for (int ax = 0; ax < 3; ax++)
                            dp[ax] = act ? cf.f[2 * ax] * face_cnt(m, 2 * ax) +
                                               cf.f[2 * ax + 1] * face_cnt(m, 2 * ax + 1)
                                         : 0.f;
                    } else if (FINE0) {
                        const uint16_t m = g.nbm[cidx(g, a, b, d)];
                        act = (m & NB_SELF) != 0;
                        for (int ax = 0; ax < 3; ax++) dp[ax] = act ? dpart0(g, m, ax) : 0.f;
                    } else {
                        const uint32_t c = ((uint32_t)a * F.ny + b) * F.nz + d;
                        act = F.active[c] != 0;
                        dp[0] = act ? F.dgx[c] : 0.f;
                        dp[1] = act ? F.dgy[c] : 0.f;
                        dp[2] = act ? F.dgz[c] : 0.f;
                    }
                    if (!act) continue;
                    any = true;
                    auto cp = [&](int axis, int x, int y, int z) {
                        if constexpr (FINE0 && ST) {  // the lower face's coupling, both cells unknown
                            const uint32_t c = cidx(g, x, y, z);
                            const uint16_t m = g.nbm[c];
                            if (!(m & NB_SELF) || !(m & (NB_UNK0 << (2 * axis)))) return 0.f;
                            const Cof<ST> cf = cof_at<ST>(g, c);
                            return axis == 0 ? cf.f[0] : (axis == 1 ? cf.f[2] : cf.f[4]);
                        }
                        return FINE0 ? coupling0(g, axis, x, y, z) : couplingc(F, axis, x, y, z);
                    };
                    for (int ax = 0; ax < 3; ax++) D[ax] += dp[ax];
                    // couplings to the lower neighbour: inside the aggregate, or
                    // across its lower face
                    if (a > i0) inner[0] += cp(0, a, b, d);
                    else if (ic > 0) low[0] += cp(0, a, b, d);
                    if (b > j0) inner[1] += cp(1, a, b, d);
                    else if (jc > 0) low[1] += cp(1, a, b, d);
                    if (d > k0) inner[2] += cp(2, a, b, d);
                    else if (kc > 0) low[2] += cp(2, a, b, d);
                }
        const float sc[3] = {C.shx ? 0.5f : 1.f, C.shy ? 0.5f : 1.f, C.shz ? 0.5f : 1.f};
        if (any) {
            float dg3[3];
            for (int ax = 0; ax < 3; ax++) dg3[ax] = sc[ax] * (D[ax] - 2.f * inner[ax]);
            float dg = dg3[0] + dg3[1] + dg3[2];
            if (!(dg > 1e-30f)) dg = 1.f;  // a pocket with no Dirichlet cell: keep finite
            C.active[I] = 1;
            C.diag[I] = dg;
            C.rdiag[I] = 1.f / dg;
            C.dgx[I] = dg3[0];
            C.dgy[I] = dg3[1];
            C.dgz[I] = dg3[2];
            C.cx[I] = sc[0] * low[0];
            C.cy[I] = sc[1] * low[1];
            C.cz[I] = sc[2] * low[2];
        } else {
            C.active[I] = 0;
            C.diag[I] = 1.f;
            C.rdiag[I] = 1.f;
            C.dgx[I] = C.dgy[I] = C.dgz[I] = 0.f;
            C.cx[I] = C.cy[I] = C.cz[I] = 0.f;
        }
    }
}

// ---- red-black Gauss-Seidel pass of one colour ------------------------------
// z := (r + sum coupling * z_nb) / diag on unknowns of `color`; a pass reads
// only the other colour, so with zero_start (z = 0 before the pass) the
// neighbour sum is skipped.
//
// Level 0 (split layout): r, the code and z of the colour are streamed
// contiguously and the six neighbours are read from the other colour's
// array. The natural index of a slot comes from its parity formula, so every
// load of a cell is issued before any returns; the gathers go through buffer
// resources, so neighbours beyond either end of the other array read 0 and
// values of faces that do not couple are dropped by the code. Each thread
// takes kU cells per round from its block's chunks (xcd_chunks: the blocks
// of an XCD walk one range together, so the other colour's values a pass
// gathers one plane later are still in that XCD's L2). Interior cells (six unknown
// neighbours, none OUTFLOW) skip the per-face selects and the diagonal
// count; their arithmetic is the general formula's with every face present.
constexpr int kU = 4;
// cells per thread of the gather kernels with the most values in flight
// (the visiting order of a thread's cells, and so its partial sums, is the
// same for any unroll: its chunks in order)
constexpr int kU2 = 2;
constexpr uint16_t kInterior = 0x007F;  // NB_SELF + six NB_UNK0 bits, no NB_OUT0 bit
constexpr uint16_t kFaceBits = 0x1FFF;

__device__ __forceinline__ uint32_t split_parity(const MgLevel &L, uint32_t h, uint32_t col)
{
    uint32_t b = col;
    if (L.par_nz) b ^= fdiv(h, L.f_hz) & 1u;
    if (L.par_ny) b ^= fdiv(h, L.f_hyz) & 1u;
    return b;
}

// the six neighbours of natural cell c from the other colour's buffer
__device__ __forceinline__ void nb_gather(__amdgpu_buffer_rsrc_t xo, uint32_t c, uint32_t si,
                                          uint32_t sj, float v[6])
{
    v[0] = buf_ldf(xo, (c - si) >> 1);
    v[1] = buf_ldf(xo, (c + si) >> 1);
    v[2] = buf_ldf(xo, (c - sj) >> 1);
    v[3] = buf_ldf(xo, (c + sj) >> 1);
    v[4] = buf_ldf(xo, (c - 1u) >> 1);
    v[5] = buf_ldf(xo, (c + 1u) >> 1);
}

// neighbour sum from gathered values (same association as a sequential sum)
__device__ __forceinline__ float nb_sum(const Geo &g, uint16_t nb, const float v[6])
{
    if ((nb & kFaceBits) == kInterior) return g.sx * (v[0] + v[1]) + g.sy * (v[2] + v[3]) + g.sz * (v[4] + v[5]);
    const float ox = ((nb & (NB_UNK0 << 0)) ? v[0] : 0.f) + ((nb & (NB_UNK0 << 1)) ? v[1] : 0.f);
    const float oy = ((nb & (NB_UNK0 << 2)) ? v[2] : 0.f) + ((nb & (NB_UNK0 << 3)) ? v[3] : 0.f);
    const float oz = ((nb & (NB_UNK0 << 4)) ? v[4] : 0.f) + ((nb & (NB_UNK0 << 5)) ? v[5] : 0.f);
    return g.sx * ox + g.sy * oy + g.sz * oz;
}

__device__ __forceinline__ float diag_of(const Geo &g, uint16_t nb, float dint)
{
    return (nb & kFaceBits) == kInterior ? dint : diag0m(g, nb);
}

template <bool ST>
__device__ __forceinline__ float nb_sum_t(const Geo &g, uint16_t nb, const float v[6],
                                          const Cof<ST> &cf)
{
    if constexpr (!ST) {
        return nb_sum(g, nb, v);
    } else {
        const float ox = ((nb & (NB_UNK0 << 0)) ? cf.f[0] * v[0] : 0.f) +
                         ((nb & (NB_UNK0 << 1)) ? cf.f[1] * v[1] : 0.f);
        const float oy = ((nb & (NB_UNK0 << 2)) ? cf.f[2] * v[2] : 0.f) +
                         ((nb & (NB_UNK0 << 3)) ? cf.f[3] * v[3] : 0.f);
        const float oz = ((nb & (NB_UNK0 << 4)) ? cf.f[4] * v[4] : 0.f) +
                         ((nb & (NB_UNK0 << 5)) ? cf.f[5] * v[5] : 0.f);
        return ox + oy + oz;
    }
}
template <bool ST>
__device__ __forceinline__ float diag_t(const Geo &g, uint16_t nb, float dint, const Cof<ST> &cf)
{
    if constexpr (!ST) {
        return diag_of(g, nb, dint);
    } else {
        float d = 0.f;
#pragma unroll
        for (int q = 0; q < 6; q++) d += cf.f[q] * face_cnt(nb, q);
        return d < 1e-30f ? 1.f : d;
    }
}

// With DOT (the V-cycle's last pass on level 0, a red pass), the kernel also
// writes the partials of r.z over both colours for the PCG: the thread of
// own slot h adds r z of that slot after its update and r z of the other
// colour's slot h (the black array is never longer than the red one).
template <bool DOT, bool ST>
struct KB_mg_smooth0 {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const MgLevel &L,
                                               const uint16_t *__restrict__ code,
                                               const float *__restrict__ zo, float *__restrict__ zc,
                                               const float *__restrict__ rc,
                                               const float *__restrict__ ro, uint32_t cnt,
                                               uint32_t cnt_o, uint32_t col, int zero_start,
                                               const MgState *m, const RedScratch &red)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t stop = stop_flags(m);  // tested after the first loads
    const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz;
    const float dint = diag0m(g, kInterior);
    const auto bcode = buf_rsrc(code, 2u * cnt), bz = buf_rsrc(zc, 4u * cnt),
               br = buf_rsrc(rc, 4u * cnt), bo = buf_rsrc(zo, 4u * cnt_o),
               bro = buf_rsrc(ro, 4u * cnt_o);
    const uint32_t blk = vb.blk;
    double acc = 0.0;
    uint32_t ch0, cst, cend;
    xcd_chunks(cnt, blk, vb.G, ch0, cst, cend);
    for (uint32_t ch = ch0; ch < cend; ch += kU * cst) {
        uint16_t nb[kU];
        float rv[kU], zv[kU][6], rov[kU], zov[kU];
        uint32_t cnat[kU];  // natural cell index (stretched grids)
#pragma unroll
        for (int t = 0; t < kU; t++) {
            const uint32_t h = (ch + t * cst) * kBlock + threadIdx.x;  // reads past cnt return 0
            const uint32_t cn = 2 * h + split_parity(L, h, col);
            nb[t] = buf_ldu16(bcode, h);
            rv[t] = buf_ldf(br, h);
            cnat[t] = cn;
            if (!zero_start) nb_gather(bo, cn, si, sj, zv[t]);
            if (DOT) {  // past cnt_o both read 0
                rov[t] = buf_ldf(bro, h);
                zov[t] = buf_ldf(bo, h);
            }
        }
        if (late(stop)) return;
#pragma unroll
        for (int t = 0; t < kU; t++) {
            const uint32_t h = (ch + t * cst) * kBlock + threadIdx.x;
            if (ch + t * cst >= cend) break;  // h past cnt: the store is dropped, rv and zov are 0
            float v = 0.f;
            if (nb[t] & NB_SELF) {
                const Cof<ST> cf = cof_at<ST>(g, cnat[t]);
                const float of = zero_start ? 0.f : nb_sum_t<ST>(g, nb[t], zv[t], cf);
                v = (rv[t] + of) / diag_t<ST>(g, nb[t], dint, cf);
            }
            buf_stf(bz, h, v);
            if (DOT) acc += (double)rv[t] * (double)v + (double)rov[t] * (double)zov[t];
        }
    }
    if (DOT) block_partial(acc, pslot(red, SLOT_MG_A), blk, lds);
}
};
template <bool DOT, bool ST>
__global__ void __launch_bounds__(kBlock) FFR_RESIDENT k_mg_smooth0(Geo g, MgLevel L,
                                                       const uint16_t *__restrict__ code,
                                                       const float *__restrict__ zo,
                                                       float *__restrict__ zc,
                                                       const float *__restrict__ rc,
                                                       const float *__restrict__ ro, uint32_t cnt,
                                                       uint32_t cnt_o, uint32_t col,
                                                       int zero_start, const MgState *m,
                                                       RedScratch red)
{
    KB_mg_smooth0<DOT, ST>::run(vb_here(), g, L, code, zo, zc, rc, ro, cnt, cnt_o, col, zero_start, m, red);
}

// Coarse levels (natural layout, stored operator). A coupling is 0 across a
// wall (k_mg_build stores none there, and the upper-side coupling of the last
// cell lies past the array), so all twelve operator and neighbour values are
// gathered unconditionally through buffer resources (out-of-range reads give
// 0) and the stop flag is waited for only after they are issued.
template <typename T>
struct CoarseOp {  // buffer views of one coarse level
    __amdgpu_buffer_rsrc_t cx, cy, cz, dg, rd, act;
};
template <typename T>
__device__ __forceinline__ CoarseOp<T> coarse_op(const MgLevel &L)
{
    const uint32_t b = 4u * L.n;
    return {buf_rsrc(L.cx, b), buf_rsrc(L.cy, b), buf_rsrc(L.cz, b), buf_rsrc(L.diag, b),
            buf_rsrc(L.rdiag, b), buf_rsrc(L.active, L.n)};
}
// coupled neighbour sum of cell c (same order as rowc)
template <typename T>
__device__ __forceinline__ float coarse_row(const CoarseOp<T> &o, __amdgpu_buffer_rsrc_t x,
                                            uint32_t c, uint32_t si, uint32_t sj)
{
    const T *tag = nullptr;
    float of = 0.f;
    of += buf_ldf(o.cx, c) * bldv(x, c - si, tag);
    of += buf_ldf(o.cx, c + si) * bldv(x, c + si, tag);
    of += buf_ldf(o.cy, c) * bldv(x, c - sj, tag);
    of += buf_ldf(o.cy, c + sj) * bldv(x, c + sj, tag);
    of += buf_ldf(o.cz, c) * bldv(x, c - 1u, tag);
    of += buf_ldf(o.cz, c + 1u) * bldv(x, c + 1u, tag);
    return of;
}

template <typename T>
struct KB_mg_smooth {
    static __device__ __forceinline__ void run(const VB &vb, const MgLevel &L, T *__restrict__ z,
                                               const T *__restrict__ r, int color, int zero_start,
                                               const MgState *m)
{
    const uint32_t stop = stop_flags(m);
    const T *tag = nullptr;
    const uint32_t si = (uint32_t)L.ny * L.nz, sj = L.nz;
    const CoarseOp<T> o = coarse_op<T>(L);
    const auto bz = buf_rsrc(z, (uint32_t)sizeof(T) * L.n), br = buf_rsrc(r, (uint32_t)sizeof(T) * L.n);
    uint32_t sb, se;
    block_span_of(L.n, vb.blk, vb.G, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        int i, j, k;
        decode(c, L.f_nynz, L.f_nz, L.ny, L.nz, i, j, k);
        const bool mine = ((i + j + k) & 1) == color;
        const uint8_t a = buf_ldu8(o.act, c);
        const float rv = bldv(br, c, tag), rd = buf_ldf(o.rd, c);
        const float of = zero_start ? 0.f : coarse_row<T>(o, bz, c, si, sj);
        if (late(stop)) return;
        if (!mine) {
            if (zero_start) bstv(bz, c, 0.f, tag);
            continue;
        }
        bstv(bz, c, a ? (rv + of) * rd : 0.f, tag);
    }
}
};
template <typename T>
__global__ void __launch_bounds__(kBlock) k_mg_smooth(MgLevel L, T *__restrict__ z,
                                                      const T *__restrict__ r, int color,
                                                      int zero_start, const MgState *m)
{
    KB_mg_smooth<T>::run(vb_here(), L, z, r, color, zero_start, m);
}

// ---- fused residual + restriction: rc = P^T (r - A z) ------------------------
// One thread per (coarse cell, child) pair: lanes 8I..8I+7 of a wave compute
// the residuals of the up-to-8 fine cells of aggregate I (children beyond an
// odd fine extent give 0) and sum them with an xor-shuffle tree, which every
// lane evaluates in the same order (deterministic).
// With red0, the coarse level's first pass (the red sweep from z = 0) is done
// here as well: zc = rc / diag on red unknowns, 0 elsewhere (rc rounded to its
// storage type first, as that pass would read it).
template <typename TC>
__device__ __forceinline__ void coarse_first_red(const MgLevel &C, TC *zc, uint32_t I, int ic,
                                                 int jc, int kc, float v)
{
    const bool red = ((ic + jc + kc) & 1) == 0;
    stv(zc, I, (red && C.active[I]) ? rnd(v, zc) * C.rdiag[I] : 0.f);
}

template <typename TF, typename TC, bool L0, bool ST>
struct KB_mg_restrict {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const MgLevel &F,
                                               const MgLevel &C, const TF *z, const TF *r, TC *rc,
                                               TC *zc, int red0, const MgState *m)
{
    const uint32_t stop = stop_flags(m);  // tested after the first loads
    uint32_t sb, se;
    block_span_of(C.n * 8u, vb.blk, vb.G, sb, se);  // spans are whole multiples of 8
    if (L0) {
        // level 0 (split layout): the colour of a child is (a+b+d)&1, so all
        // loads of kU children are issued before any is used
        const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz, H = F.H;
        const float dint = diag0m(g, kInterior);
        const float *zs = (const float *)z, *rs = (const float *)r;
        const auto bred = buf_rsrc(zs, 4u * H), bblack = buf_rsrc(zs + H, 4u * (g.nc - H));
        for (uint32_t t0 = sb + threadIdx.x; t0 < se; t0 += kU * kBlock) {
            uint16_t nb[kU];
            float rv[kU], zown[kU], zv[kU][6];
            bool in[kU];
            uint32_t cnat[kU];  // natural cell index (stretched grids)
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const uint32_t t = t0 + u * kBlock < se ? t0 + u * kBlock : t0;
                const uint32_t I = t >> 3, q = t & 7u;
                int ic, jc, kc;
                decode(I, C.f_nynz, C.f_nz, C.ny, C.nz, ic, jc, kc);
                const int qx = (int)(q >> 2), qy = (int)((q >> 1) & 1u), qz = (int)(q & 1u);
                const int a = (ic << C.shx) + qx, b = (jc << C.shy) + qy, d = (kc << C.shz) + qz;
                in[u] = qx <= C.shx && qy <= C.shy && qz <= C.shz && a < F.nx && b < F.ny &&
                        d < F.nz;
                const uint32_t c = in[u] ? ((uint32_t)a * F.ny + b) * F.nz + d : 0u;
                const uint32_t col = in[u] ? (uint32_t)(a + b + d) & 1u : 0u;
                const uint32_t own = (col ? H : 0u) + (c >> 1);
                nb[u] = g.nbm[c];
                rv[u] = rs[own];
                zown[u] = zs[own];
                cnat[u] = c;
                nb_gather(col ? bred : bblack, c, si, sj, zv[u]);
            }
            if (late(stop)) return;
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const uint32_t t = t0 + u * kBlock;
                if (t >= se) break;
                float v = 0.f;
                if (in[u] && (nb[u] & NB_SELF)) {
                    const Cof<ST> cf = cof_at<ST>(g, cnat[u]);
                    v = rv[u] - (diag_t<ST>(g, nb[u], dint, cf) * zown[u] -
                                 nb_sum_t<ST>(g, nb[u], zv[u], cf));
                }
                v += __shfl_xor(v, 1);
                v += __shfl_xor(v, 2);
                v += __shfl_xor(v, 4);
                if ((t & 7u) == 0) {
                    const uint32_t I = t >> 3;
                    stv(rc, I, v);
                    if (red0) {
                        int ic, jc, kc;
                        decode(I, C.f_nynz, C.f_nz, C.ny, C.nz, ic, jc, kc);
                        coarse_first_red<TC>(C, zc, I, ic, jc, kc, v);
                    }
                }
            }
        }
        return;
    }
    const TF *tag = nullptr;
    const uint32_t si = (uint32_t)F.ny * F.nz, sj = F.nz;
    const CoarseOp<TF> o = coarse_op<TF>(F);
    const auto bz = buf_rsrc(z, (uint32_t)sizeof(TF) * F.n), br = buf_rsrc(r, (uint32_t)sizeof(TF) * F.n);
    for (uint32_t t = sb + threadIdx.x; t < se; t += kBlock) {
        const uint32_t I = t >> 3, q = t & 7u;
        int ic, jc, kc;
        decode(I, C.f_nynz, C.f_nz, C.ny, C.nz, ic, jc, kc);
        const int qx = (int)(q >> 2), qy = (int)((q >> 1) & 1u), qz = (int)(q & 1u);
        const int a = (ic << C.shx) + qx, b = (jc << C.shy) + qy, d = (kc << C.shz) + qz;
        const bool in = qx <= C.shx && qy <= C.shy && qz <= C.shz && a < F.nx && b < F.ny &&
                        d < F.nz;
        const uint32_t c = in ? ((uint32_t)a * F.ny + b) * F.nz + d : 0u;
        const uint8_t act = buf_ldu8(o.act, c);
        const float rv = bldv(br, c, tag), zv = bldv(bz, c, tag), dg = buf_ldf(o.dg, c);
        const float of = coarse_row<TF>(o, bz, c, si, sj);
        if (late(stop)) return;
        float v = (in && act) ? rv - (dg * zv - of) : 0.f;
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        if (q == 0) {
            stv(rc, I, v);
            if (red0) coarse_first_red<TC>(C, zc, I, ic, jc, kc, v);
        }
    }
}
};
template <typename TF, typename TC, bool L0, bool ST>
__global__ void __launch_bounds__(kBlock) k_mg_restrict(Geo g, MgLevel F, MgLevel C, const TF *z,
                                                        const TF *r, TC *rc, TC *zc, int red0,
                                                        const MgState *m)
{
    KB_mg_restrict<TF, TC, L0, ST>::run(vb_here(), g, F, C, z, r, rc, zc, red0, m);
}

// ---- coarse-level restriction in two passes (large levels) ------------------
// The fused kernel above gives each (coarse cell, child) pair a lane, so a
// wave reads its fine level in short runs from two planes and two rows; on a
// large level the residual is instead formed one thread per fine cell in the
// natural order (coalesced) into an FP32 scratch, then summed per coarse cell
// in the fused kernel's lane order (the same residual formula and sum order).
template <typename T>
struct KB_mg_cresid {
    static __device__ __forceinline__ void run(const VB &vb, const MgLevel &F, const T *__restrict__ z,
                                               const T *__restrict__ r, float *__restrict__ res,
                                               const MgState *m)
{
    const uint32_t stop = stop_flags(m);
    const T *tag = nullptr;
    const uint32_t si = (uint32_t)F.ny * F.nz, sj = F.nz;
    const CoarseOp<T> o = coarse_op<T>(F);
    const auto bz = buf_rsrc(z, (uint32_t)sizeof(T) * F.n), br = buf_rsrc(r, (uint32_t)sizeof(T) * F.n),
               bres = buf_rsrc(res, 4u * F.n);
    uint32_t ch0, cst, cend;
    xcd_chunks(F.n, vb.blk, vb.G, ch0, cst, cend);
    for (uint32_t ch = ch0; ch < cend; ch += kU2 * cst) {
        float v[kU2];
        uint8_t a[kU2];
#pragma unroll
        for (int t = 0; t < kU2; t++) {  // past n: zero reads, dropped stores
            const uint32_t c = (ch + t * cst) * kBlock + threadIdx.x;
            a[t] = buf_ldu8(o.act, c);
            v[t] = bldv(br, c, tag) - (buf_ldf(o.dg, c) * bldv(bz, c, tag) - coarse_row<T>(o, bz, c, si, sj));
        }
        if (late(stop)) return;
#pragma unroll
        for (int t = 0; t < kU2; t++) {
            if (ch + t * cst >= cend) break;
            buf_stf(bres, (ch + t * cst) * kBlock + threadIdx.x, a[t] ? v[t] : 0.f);
        }
    }
}
};
template <typename T>
__global__ void __launch_bounds__(kBlock) k_mg_cresid(MgLevel F, const T *__restrict__ z,
                                                      const T *__restrict__ r, float *__restrict__ res,
                                                      const MgState *m)
{
    KB_mg_cresid<T>::run(vb_here(), F, z, r, res, m);
}

template <typename TC>
struct KB_mg_csum {
    static __device__ __forceinline__ void run(const VB &vb, const MgLevel &F, const MgLevel &C,
                                               const float *__restrict__ res, TC *__restrict__ rc,
                                               TC *__restrict__ zc, int red0, const MgState *m)
{
    const uint32_t stop = stop_flags(m);
    const auto bres = buf_rsrc(res, 4u * F.n);
    uint32_t sb, se;
    block_span_of(C.n, vb.blk, vb.G, sb, se);
    for (uint32_t I = sb + threadIdx.x; I < se; I += kBlock) {
        int ic, jc, kc;
        decode(I, C.f_nynz, C.f_nz, C.ny, C.nz, ic, jc, kc);
        float part[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {  // child (qx, qy, qz): the fused kernel's lane q
            const int qx = q >> 2, qy = (q >> 1) & 1, qz = q & 1;
            const int a = (ic << C.shx) + qx, b = (jc << C.shy) + qy, d = (kc << C.shz) + qz;
            const bool in = qx <= C.shx && qy <= C.shy && qz <= C.shz && a < F.nx && b < F.ny &&
                            d < F.nz;
            part[q] = buf_ldf(bres, in ? ((uint32_t)a * F.ny + b) * F.nz + d : 0xFFFFFFFFu >> 2);
        }
        if (late(stop)) return;
        // the xor-tree order of the fused kernel's lane sum (lanes q ^ 1, q ^ 2, q ^ 4)
        const float s01 = part[0] + part[1], s23 = part[2] + part[3];
        const float s45 = part[4] + part[5], s67 = part[6] + part[7];
        const float v = (s01 + s23) + (s45 + s67);
        stv(rc, I, v);
        if (red0) coarse_first_red<TC>(C, zc, I, ic, jc, kc, v);
    }
}
};
template <typename TC>
__global__ void __launch_bounds__(kBlock) k_mg_csum(MgLevel F, MgLevel C, const float *__restrict__ res,
                                                    TC *__restrict__ rc, TC *__restrict__ zc, int red0,
                                                    const MgState *m)
{
    KB_mg_csum<TC>::run(vb_here(), F, C, res, rc, zc, red0, m);
}

// ---- level-0 restriction in two passes --------------------------------------
// (1) res = r - A z on the red slots of the split layout (one thread per slot,
//     the smoother's gather pattern). Pre-smoothing ends with a black pass,
//     which leaves r - A z = 0 on every black cell (up to that pass's
//     rounding: z_black = (r + offdiag z_red) / diag), so only the red half is
//     formed; (2) rc = P^T res, one thread per coarse cell summing its
//     up-to-8 children (black children count 0), whose split slots are
//     contiguous across consecutive coarse cells. res lives in the level-0
//     scratch, which is not in use while the preconditioner runs.
template <bool ST>
struct KB_mg_resid0 {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const MgLevel &L,
                                               const float *__restrict__ z,
                                               const float *__restrict__ r, float *__restrict__ res,
                                               const MgState *m)
{
    const uint32_t stop = stop_flags(m);
    const uint32_t H = L.H, si = (uint32_t)g.ny * g.nz, sj = g.nz;
    const float dint = diag0m(g, kInterior);
    const auto bcode = buf_rsrc(L.nbs, 2u * H), bz = buf_rsrc(z, 4u * H),
               br = buf_rsrc(r, 4u * H), bblack = buf_rsrc(z + H, 4u * (g.nc - H)),
               bres = buf_rsrc(res, 4u * H);
    uint32_t ch0, cst, cend;
    xcd_chunks(H, vb.blk, vb.G, ch0, cst, cend);
    for (uint32_t ch = ch0; ch < cend; ch += kU2 * cst) {
        uint16_t nb[kU2];
        float rv[kU2], zs[kU2], zv[kU2][6];
        uint32_t cnat[kU2];  // natural cell index (stretched grids)
#pragma unroll
        for (int t = 0; t < kU2; t++) {
            const uint32_t h = (ch + t * cst) * kBlock + threadIdx.x;  // reads past H return 0
            const uint32_t cn = 2 * h + split_parity(L, h, 0u);
            nb[t] = buf_ldu16(bcode, h);
            rv[t] = buf_ldf(br, h);
            zs[t] = buf_ldf(bz, h);
            cnat[t] = cn;
            nb_gather(bblack, cn, si, sj, zv[t]);
        }
        if (late(stop)) return;
#pragma unroll
        for (int t = 0; t < kU2; t++) {
            const uint32_t h = (ch + t * cst) * kBlock + threadIdx.x;
            if (ch + t * cst >= cend) break;  // stores past H are dropped
            float v = 0.f;
            if (nb[t] & NB_SELF) {
                const Cof<ST> cf = cof_at<ST>(g, cnat[t]);
                v = rv[t] - (diag_t<ST>(g, nb[t], dint, cf) * zs[t] - nb_sum_t<ST>(g, nb[t], zv[t], cf));
            }
            buf_stf(bres, h, v);
        }
    }
}
};
template <bool ST>
__global__ void __launch_bounds__(kBlock) FFR_RESIDENT k_mg_resid0(Geo g, MgLevel L, const float *__restrict__ z,
                                                      const float *__restrict__ r,
                                                      float *__restrict__ res, const MgState *m)
{
    KB_mg_resid0<ST>::run(vb_here(), g, L, z, r, res, m);
}

template <typename TC>
struct KB_mg_restrict_sum0 {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const MgLevel &F,
                                               const MgLevel &C, const float *__restrict__ res,
                                               TC *__restrict__ rc, TC *__restrict__ zc, int red0,
                                               const MgState *m)
{
    const uint32_t stop = stop_flags(m);
    const uint32_t H = F.H;
    const auto bred = buf_rsrc(res, 4u * H);
    uint32_t sb, se;
    block_span_of(C.n, vb.blk, vb.G, sb, se);
    for (uint32_t I = sb + threadIdx.x; I < se; I += kBlock) {
        int ic, jc, kc;
        decode(I, C.f_nynz, C.f_nz, C.ny, C.nz, ic, jc, kc);
        float part[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {  // child (qx, qy, qz), in the same order as the 8-lane sum
            const int qx = q >> 2, qy = (q >> 1) & 1, qz = q & 1;
            const int a = (ic << C.shx) + qx, b = (jc << C.shy) + qy, d = (kc << C.shz) + qz;
            const bool in = qx <= C.shx && qy <= C.shy && qz <= C.shz && a < F.nx && b < F.ny &&
                            d < F.nz;
            const uint32_t c = ((uint32_t)a * F.ny + b) * F.nz + d;
            // out-of-range children read past the buffer (0); black children
            // have no residual after the pre-smoothing's last (black) pass
            const uint32_t h = in && !((a + b + d) & 1) ? (c >> 1) : 0xFFFFFFFFu >> 3;
            part[q] = buf_ldf(bred, h);
        }
        if (late(stop)) return;
        // the xor-tree order of k_mg_restrict's lane sum
        const float s01 = part[0] + part[1], s23 = part[2] + part[3];
        const float s45 = part[4] + part[5], s67 = part[6] + part[7];
        const float v = (s01 + s23) + (s45 + s67);
        stv(rc, I, v);
        if (red0) coarse_first_red<TC>(C, zc, I, ic, jc, kc, v);
    }
}
};
template <typename TC>
__global__ void __launch_bounds__(kBlock) k_mg_restrict_sum0(Geo g, MgLevel F, MgLevel C,
                                                             const float *__restrict__ res,
                                                             TC *__restrict__ rc,
                                                             TC *__restrict__ zc, int red0,
                                                             const MgState *m)
{
    KB_mg_restrict_sum0<TC>::run(vb_here(), g, F, C, res, rc, zc, red0, m);
}

// ---- prolongation fused into the first post-smoothing sweep ----------------
// After the coarse correction the first post pass is the black sweep; it reads
// only red neighbours, so the prolonged red values z_nb + zc[parent(nb)] are
// formed where they are read (rounded to the level's storage type, as a stored
// prolongation would give) and never stored: the red sweep that follows
// overwrites every red value. Black unknowns get (r + sum coupling *
// (z_nb + zc)) / diag, black non-unknowns 0, red slots are not touched.

// parent of fine cell (i, j, k) on the coarse level C (an out-of-range fine
// index gives an index whose read returns 0)
__device__ __forceinline__ uint32_t parent_of(const MgLevel &C, int i, int j, int k)
{
    return ((uint32_t)(i >> C.shx) * C.ny + (uint32_t)(j >> C.shy)) * C.nz + (uint32_t)(k >> C.shz);
}

// level 0 (split layout): one thread per black slot, kU slots per thread
template <typename TC, bool ST>
struct KB_mg_post_black0 {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const MgLevel &L,
                                               const MgLevel &C, const TC *__restrict__ zc,
                                               const MgState *m)
{
    const uint32_t stop = stop_flags(m);
    const uint32_t H = L.H, cnt = g.nc - H;
    const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz;
    const float dint = diag0m(g, kInterior);
    const TC *tag = nullptr;
    float *zb = (float *)L.z + H;
    const auto bcode = buf_rsrc(L.nbs + H, 2u * cnt), br = buf_rsrc((const float *)L.r + H, 4u * cnt),
               bred = buf_rsrc(L.z, 4u * H), bzc = buf_rsrc(zc, (uint32_t)sizeof(TC) * C.n),
               bz = buf_rsrc(zb, 4u * cnt);
    uint32_t ch0, cst, cend;
    xcd_chunks(cnt, vb.blk, vb.G, ch0, cst, cend);
    for (uint32_t ch = ch0; ch < cend; ch += kU2 * cst) {
        uint16_t nb[kU2];
        float rv[kU2], zv[kU2][6];
        uint32_t cnat[kU2];  // natural cell index (stretched grids)
#pragma unroll
        for (int t = 0; t < kU2; t++) {
            const uint32_t h = (ch + t * cst) * kBlock + threadIdx.x;  // reads past cnt return 0
            nb[t] = buf_ldu16(bcode, h);
            rv[t] = buf_ldf(br, h);
            const uint32_t c = 2 * h + split_parity(L, h, 1u);
            cnat[t] = c;
            const int i = (int)fdiv(c, g.f_nynz), rem = (int)(c - (uint32_t)i * si);
            const int j = (int)fdiv((uint32_t)rem, g.f_nz), k = rem - j * (int)sj;
            nb_gather(bred, c, si, sj, zv[t]);
            zv[t][0] += bldv(bzc, parent_of(C, i - 1, j, k), tag);
            zv[t][1] += bldv(bzc, parent_of(C, i + 1, j, k), tag);
            zv[t][2] += bldv(bzc, parent_of(C, i, j - 1, k), tag);
            zv[t][3] += bldv(bzc, parent_of(C, i, j + 1, k), tag);
            zv[t][4] += bldv(bzc, parent_of(C, i, j, k - 1), tag);
            zv[t][5] += bldv(bzc, parent_of(C, i, j, k + 1), tag);
        }
        if (late(stop)) return;
#pragma unroll
        for (int t = 0; t < kU2; t++) {
            const uint32_t h = (ch + t * cst) * kBlock + threadIdx.x;
            if (ch + t * cst >= cend) break;  // stores past cnt are dropped
            float v = 0.f;
            if (nb[t] & NB_SELF) {
                const Cof<ST> cf = cof_at<ST>(g, cnat[t]);
                v = (rv[t] + nb_sum_t<ST>(g, nb[t], zv[t], cf)) / diag_t<ST>(g, nb[t], dint, cf);
            }
            buf_stf(bz, h, v);
        }
    }
}
};
template <typename TC, bool ST>
__global__ void __launch_bounds__(kBlock) FFR_RESIDENT k_mg_post_black0(Geo g, MgLevel L, MgLevel C,
                                                           const TC *__restrict__ zc,
                                                           const MgState *m)
{
    KB_mg_post_black0<TC, ST>::run(vb_here(), g, L, C, zc, m);
}

// coarse levels (natural layout)
template <typename T, typename TC>
struct KB_mg_post_black {
    static __device__ __forceinline__ void run(const VB &vb, const MgLevel &L, const MgLevel &C,
                                               T *__restrict__ z, const T *__restrict__ r,
                                               const TC *__restrict__ zc, const MgState *m)
{
    const uint32_t stop = stop_flags(m);
    const T *tag = nullptr;
    const TC *ctag = nullptr;
    const uint32_t si = (uint32_t)L.ny * L.nz, sj = L.nz;
    const CoarseOp<T> o = coarse_op<T>(L);
    const auto bz = buf_rsrc(z, (uint32_t)sizeof(T) * L.n), br = buf_rsrc(r, (uint32_t)sizeof(T) * L.n),
               bzc = buf_rsrc(zc, (uint32_t)sizeof(TC) * C.n);
    uint32_t sb, se;
    block_span_of(L.n, vb.blk, vb.G, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        int i, j, k;
        decode(c, L.f_nynz, L.f_nz, L.ny, L.nz, i, j, k);
        const bool black = ((i + j + k) & 1) == 1;
        const uint8_t a = buf_ldu8(o.act, c);
        const float rv = bldv(br, c, tag), rd = buf_ldf(o.rd, c);
        auto nbv = [&](uint32_t q, int x, int y, int w) {  // prolonged neighbour value
            return rnd(bldv(bz, q, tag) + bldv(bzc, parent_of(C, x, y, w), ctag), tag);
        };
        float of = 0.f;
        of += buf_ldf(o.cx, c) * nbv(c - si, i - 1, j, k);
        of += buf_ldf(o.cx, c + si) * nbv(c + si, i + 1, j, k);
        of += buf_ldf(o.cy, c) * nbv(c - sj, i, j - 1, k);
        of += buf_ldf(o.cy, c + sj) * nbv(c + sj, i, j + 1, k);
        of += buf_ldf(o.cz, c) * nbv(c - 1u, i, j, k - 1);
        of += buf_ldf(o.cz, c + 1u) * nbv(c + 1u, i, j, k + 1);
        if (late(stop)) return;
        if (black) bstv(bz, c, a ? (rv + of) * rd : 0.f, tag);
    }
}
};
template <typename T, typename TC>
__global__ void __launch_bounds__(kBlock) k_mg_post_black(MgLevel L, MgLevel C, T *__restrict__ z,
                                                          const T *__restrict__ r,
                                                          const TC *__restrict__ zc,
                                                          const MgState *m)
{
    KB_mg_post_black<T, TC>::run(vb_here(), L, C, z, r, zc, m);
}

// ---- coarsest level: symmetric red-black sweeps from z = 0 (one block) -----
template <typename T, bool L0, bool ST>
struct KB_mg_coarse {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const MgLevel &L, T *z,
                                               const T *r, int sweeps, const MgState *m)
{
    if (stopped(m)) return;
    // level 0 (a single-level hierarchy) uses the split layout: slots [0, n)
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
                const float *zs = (const float *)z;
                const uint32_t sc = sown(c, nb, L.H);
                if constexpr (ST) {
                    const Cof<ST> cf = cof_at<ST>(g, c);
                    const float *xo = zs + sother(nb, L.H);
                    const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz;
                    float v[6];  // values of faces that do not couple are dropped
                    v[0] = (nb & (NB_UNK0 << 0)) ? xo[(c - si) >> 1] : 0.f;
                    v[1] = (nb & (NB_UNK0 << 1)) ? xo[(c + si) >> 1] : 0.f;
                    v[2] = (nb & (NB_UNK0 << 2)) ? xo[(c - sj) >> 1] : 0.f;
                    v[3] = (nb & (NB_UNK0 << 3)) ? xo[(c + sj) >> 1] : 0.f;
                    v[4] = (nb & (NB_UNK0 << 4)) ? xo[(c - 1) >> 1] : 0.f;
                    v[5] = (nb & (NB_UNK0 << 5)) ? xo[(c + 1) >> 1] : 0.f;
                    stv(z, sc, (ldv(r, sc) + nb_sum_t<ST>(g, nb, v, cf)) / diag_t<ST>(g, nb, 0.f, cf));
                } else {
                    stv(z, sc, (ldv(r, sc) + off0s(g, zs + sother(nb, L.H), c, nb)) / diag0m(g, nb));
                }
            } else {
                if (!L.active[c]) continue;
                stv(z, c, (ldv(r, c) + rowc<T>(L, z, i, j, k, c)) * L.rdiag[c]);
            }
        }
        __syncthreads();
    }
}
};
template <typename T, bool L0, bool ST>
__global__ void __launch_bounds__(1024) k_mg_coarse(Geo g, MgLevel L, T *z, const T *r, int sweeps,
                                                    const MgState *m)
{
    KB_mg_coarse<T, L0, ST>::run(vb_here(), g, L, z, r, sweeps, m);
}

// ---- the small levels in one workgroup (LDS-resident) --------------------------
// The V-cycle from level tl.first down to the coarsest and back, with the same
// passes as the per-level kernels (red-black sweeps, P^T restriction of the
// residual, piecewise-constant prolongation, symmetric sweeps on the coarsest
// level). Each level's operator is staged in LDS once and its z and r live in
// LDS as T; phases are separated by workgroup barriers. Input r and output z
// of level tl.first are exchanged with global memory. Levels before
// tl.lds_from (scenes served side by side: mg_tail_cells) run the same
// phases on their global arrays - one CU and one launch for what would be
// several launches per level - ordered by the same workgroup barriers.
// A view's type says where its vectors are (GL: global), so the LDS levels'
// accesses stay LDS instructions.
template <typename T, bool GL>
struct TailView {  // one level of the tail
    using elem = T;
    uint32_t n, H, si, sj;
    int nx, ny, nz;
    int shx, shy, shz;
    int par_nz, par_ny;
    FastDiv f_nz, f_nynz, f_hz, f_hyz;
    bool staged;             // operator in LDS at base + tc, else in global memory (G)
    unsigned char *base;     // the level's LDS region
    TailCarve tc;
    const MgLevel *G;
    T *z, *r;  // LDS (global when GL); z[-si .. n + si) readable, the pads 0
};

template <typename T, bool GL>
__device__ __forceinline__ TailView<T, GL> tail_view(const TailLevels &tl, unsigned char *smem, int l)
{
    TailView<T, GL> v;
    const MgLevel &L = tl.lv[l];
    v.n = L.n;
    v.H = (L.n + 1) / 2;
    v.nx = L.nx;
    v.ny = L.ny;
    v.nz = L.nz;
    v.si = (uint32_t)L.ny * L.nz;
    v.sj = (uint32_t)L.nz;
    v.shx = L.shx;
    v.shy = L.shy;
    v.shz = L.shz;
    v.par_nz = L.par_nz;
    v.par_ny = L.par_ny;
    v.f_nz = L.f_nz;
    v.f_nynz = L.f_nynz;
    v.f_hz = L.f_hz;
    v.f_hyz = L.f_hyz;
    v.G = &tl.lv[l];
    if constexpr (GL) {
        v.staged = false;
        v.base = nullptr;
        v.tc = TailCarve{};
        v.z = reinterpret_cast<T *>(L.z);
        v.r = reinterpret_cast<T *>(L.r);
    } else {
        v.staged = l >= tl.staged_from;
        v.tc = tail_carve(v.n, v.si, v.sj, (uint32_t)sizeof(T), v.staged);
        v.base = smem + tl.off[l];
        v.z = reinterpret_cast<T *>(v.base + v.tc.z);
        v.r = reinterpret_cast<T *>(v.base + v.tc.r);
    }
    return v;
}

// The operator of a tail level, from LDS (ST: pointers derived from the LDS
// region, so the compiler emits LDS loads) or from global memory
struct TailOp {
    const float *cx, *cy, *cz, *diag, *rdiag;
    const uint8_t *act;
};
template <bool ST, typename V>
__device__ __forceinline__ TailOp tail_op(const V &L)
{
    if (ST)
        return {reinterpret_cast<const float *>(L.base + L.tc.cx),
                reinterpret_cast<const float *>(L.base + L.tc.cy),
                reinterpret_cast<const float *>(L.base + L.tc.cz),
                reinterpret_cast<const float *>(L.base + L.tc.diag),
                reinterpret_cast<const float *>(L.base + L.tc.rdiag), L.base + L.tc.act};
    return {L.G->cx, L.G->cy, L.G->cz, L.G->diag, L.G->rdiag, L.G->active};
}

// natural index of slot m of colour col (cells 2m and 2m + 1 always differ
// in colour; the parity formula is level 0's split-layout one)
template <typename V>
__device__ __forceinline__ uint32_t tail_cell(const V &L, uint32_t m, uint32_t col)
{
    uint32_t b = col;
    if (L.par_nz) b ^= fdiv(m, L.f_hz) & 1u;
    if (L.par_ny) b ^= fdiv(m, L.f_hyz) & 1u;
    return 2 * m + b;
}

// coupled neighbour sum of cell c from x (padded): the couplings of the
// faces on walls are 0, and the upper neighbours' couplings past the end
// read staged zero pads or, for an operator in global memory, the next
// array of the arena (finite) times a zero pad of x
template <typename V, typename T>
__device__ __forceinline__ float tail_row(const V &L, const TailOp &o, const T *x, uint32_t c)
{
    const T *xc = x + c;
    float of = 0.f;
    of += o.cx[c] * ldv(xc - L.si, 0u);
    of += o.cx[c + L.si] * ldv(xc + L.si, 0u);
    of += o.cy[c] * ldv(xc - L.sj, 0u);
    of += o.cy[c + L.sj] * ldv(xc + L.sj, 0u);
    of += o.cz[c] * ldv(xc - 1, 0u);
    of += o.cz[c + 1] * ldv(xc + 1, 0u);
    return of;
}

// one colour pass: z = (r + offdiag z) / diag on cells of colour `color`
// (from z = 0 with zero_start), by threads [0, nthr) in steps of nthr. Cells
// with no unknown child have r = 0 and no couplings, so their z stays 0
// without a test. A zero_start pass also writes 0 to the other cell of each
// pair: a row's last cell and the next row's first cell can share a colour
// (even nz), and the zero coupling between them must multiply a finite
// value, not whatever LDS held. Every load is issued for all U cells of a
// thread before any is used (out-of-range slots read cell 0 and store
// nothing).
template <int U, bool ST, typename V>
__device__ __forceinline__ void tail_pass_u(const V &L, uint32_t color, int zero_start,
                                            uint32_t tid, uint32_t nthr)
{
    const TailOp o = tail_op<ST>(L);
    for (uint32_t m0 = tid; m0 < L.H; m0 += U * nthr) {
        uint32_t c[U];
        bool ok[U];
        float rv[U], rd[U], of[U];
#pragma unroll
        for (int t = 0; t < U; t++) {
            const uint32_t m = m0 + (uint32_t)t * nthr;
            c[t] = tail_cell(L, m < L.H ? m : 0u, color);
            ok[t] = m < L.H && c[t] < L.n;
            const uint32_t cc = ok[t] ? c[t] : 0u;
            rv[t] = ldv(L.r, cc);
            rd[t] = o.rdiag[cc];
            of[t] = zero_start ? 0.f : tail_row(L, o, L.z, cc);
        }
#pragma unroll
        for (int t = 0; t < U; t++) {
            const uint32_t m = m0 + (uint32_t)t * nthr;
            if (zero_start && m < L.H && (c[t] ^ 1u) < L.n) stv(L.z, c[t] ^ 1u, 0.f);
            if (ok[t]) stv(L.z, c[t], (rv[t] + of[t]) * rd[t]);
        }
    }
}

// The per-cell phases are instruction-bound: a thread takes 1, 2 or 4 cells
// per round, as few as cover the level, so no slot is computed for nothing
// while the loads of a thread's cells are still issued together.
template <bool ST, typename V>
__device__ __forceinline__ void tail_pass(const V &L, uint32_t color, int zero_start,
                                          uint32_t tid, uint32_t nthr)
{
    if (L.H <= nthr) tail_pass_u<1, ST>(L, color, zero_start, tid, nthr);
    else if (L.H <= 2 * nthr) tail_pass_u<2, ST>(L, color, zero_start, tid, nthr);
    else tail_pass_u<4, ST>(L, color, zero_start, tid, nthr);
}

// Who runs a tail phase, and the barrier that ends it: the threads of one
// workgroup (LDS levels, or a tail of one workgroup) or of the tail's
// workgroups on one XCD (global levels, TailLevels::groups > 1). The group
// barrier needs no device-scope fence: the workgroups share the XCD's L2, a
// wave's stores are in it once its vmcnt has drained, the arrival counter is
// an L2 atomic, and the CU's L1 is invalidated after the wait so that no line
// another workgroup has since written is read stale (an agent-scope release
// would also write the L2 back to HBM: measured 2.4-9.2 us per barrier of 1-8
// workgroups on MI355X, against 0.6-1.0 us this way).
struct TailSync {
    uint32_t tid, nthr;  // this thread's index among the phase's threads, their count
    unsigned *ctr;       // group arrival counter (null: one workgroup)
    int *fault;
    unsigned G, k;       // workgroups; barriers passed (thread 0)
    bool broken;         // a wait timed out (thread 0): later barriers only arrive
    __device__ __forceinline__ void sync()
    {
        if (!ctr) {
            __syncthreads();
            return;
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores are in the L2
        __syncthreads();
        if (threadIdx.x == 0) {
            k++;
            __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!broken) {
                const unsigned target = G * k;
                unsigned polls = 0;
                while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                    __builtin_amdgcn_s_sleep(2);
                    if (++polls == kTailSpinMax) {
                        broken = true;
                        __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the CU's L1 invalidated
        }
        __syncthreads();
    }
};
__device__ __forceinline__ TailSync tail_solo()
{
    return TailSync{threadIdx.x, (uint32_t)kTailThreads, nullptr, nullptr, 1u, 0u, false};
}

template <typename V>
__device__ __forceinline__ void tail_smooth(const V &L, uint32_t color, int zero_start, TailSync &ts)
{
    if (L.staged) tail_pass<true>(L, color, zero_start, ts.tid, ts.nthr);
    else tail_pass<false>(L, color, zero_start, ts.tid, ts.nthr);
    ts.sync();
}

// residuals r - A z of level F into res (FP32), all cells
template <int U, bool ST, typename V>
__device__ __forceinline__ void tail_residual_u(const V &F, float *res, uint32_t tid, uint32_t nthr)
{
    const TailOp o = tail_op<ST>(F);
    for (uint32_t c0 = tid; c0 < F.n; c0 += U * nthr) {
        float v[U];
        uint8_t a[U];
#pragma unroll
        for (int t = 0; t < U; t++) {
            const uint32_t c = c0 + (uint32_t)t * nthr;
            const uint32_t cc = c < F.n ? c : 0u;
            a[t] = o.act[cc];
            v[t] = ldv(F.r, cc) - (o.diag[cc] * ldv(F.z, cc) - tail_row(F, o, F.z, cc));
        }
#pragma unroll
        for (int t = 0; t < U; t++) {
            const uint32_t c = c0 + (uint32_t)t * nthr;
            if (c < F.n) res[c] = a[t] ? v[t] : 0.f;
        }
    }
}
template <bool ST, typename V>
__device__ __forceinline__ void tail_residual(const V &F, float *res, uint32_t tid, uint32_t nthr)
{
    if (F.n <= nthr) tail_residual_u<1, ST>(F, res, tid, nthr);
    else if (F.n <= 2 * nthr) tail_residual_u<2, ST>(F, res, tid, nthr);
    else tail_residual_u<4, ST>(F, res, tid, nthr);
}

// F.z += prolonged C.z on unknown cells
template <int U, bool ST, typename VF, typename VC>
__device__ __forceinline__ void tail_prolong_u(const VF &F, const VC &C, uint32_t tid, uint32_t nthr)
{
    const TailOp o = tail_op<ST>(F);
    for (uint32_t c0 = tid; c0 < F.n; c0 += U * nthr) {
        float v[U];
        uint8_t a[U];
#pragma unroll
        for (int t = 0; t < U; t++) {
            const uint32_t c = c0 + (uint32_t)t * nthr;
            const uint32_t cc = c < F.n ? c : 0u;
            int i, j, k;
            decode(cc, F.f_nynz, F.f_nz, F.ny, F.nz, i, j, k);
            const uint32_t P =
                ((uint32_t)(i >> C.shx) * C.ny + (j >> C.shy)) * C.nz + (k >> C.shz);
            a[t] = o.act[cc];
            v[t] = ldv(F.z, cc) + ldv(C.z, P);
        }
#pragma unroll
        for (int t = 0; t < U; t++) {
            const uint32_t c = c0 + (uint32_t)t * nthr;
            if (c < F.n && a[t]) stv(F.z, c, v[t]);
        }
    }
}
template <bool ST, typename VF, typename VC>
__device__ __forceinline__ void tail_prolong(const VF &F, const VC &C, uint32_t tid, uint32_t nthr)
{
    if (F.n <= nthr) tail_prolong_u<1, ST>(F, C, tid, nthr);
    else if (F.n <= 2 * nthr) tail_prolong_u<2, ST>(F, C, tid, nthr);
    else tail_prolong_u<4, ST>(F, C, tid, nthr);
}

// rc = P^T res: each coarse cell of C sums its fine children
template <typename VC>
__device__ __forceinline__ void tail_restrict_sum(const MgLevel &F, const VC &C, const float *res,
                                                  uint32_t tid, uint32_t nthr)
{
    for (uint32_t I = tid; I < C.n; I += nthr) {
        int ic, jc, kc;
        decode(I, C.f_nynz, C.f_nz, C.ny, C.nz, ic, jc, kc);
        const int i0 = ic << C.shx, j0 = jc << C.shy, k0 = kc << C.shz;
        const int i1 = min(i0 + (1 << C.shx) - 1, F.nx - 1),
                  j1 = min(j0 + (1 << C.shy) - 1, F.ny - 1),
                  k1 = min(k0 + (1 << C.shz) - 1, F.nz - 1);
        float sum = 0.f;
        for (int a = i0; a <= i1; a++)
            for (int b = j0; b <= j1; b++)
                for (int d = k0; d <= k1; d++) sum += res[((uint32_t)a * F.ny + b) * F.nz + d];
        stv(C.r, I, sum);
    }
}

// one level of the way down: pre-smoothing, the residual into res (loads
// issued together), then each coarse cell sums its children
template <typename VF, typename VC>
__device__ __forceinline__ void tail_down(const TailLevels &tl, const VF &F, const VC &C, float *res,
                                          TailSync &ts)
{
    for (int sw = 0; sw < tl.sweeps; sw++) {
        tail_smooth(F, 0, sw == 0, ts);
        tail_smooth(F, 1, 0, ts);
    }
    if (F.staged) tail_residual<true>(F, res, ts.tid, ts.nthr);
    else tail_residual<false>(F, res, ts.tid, ts.nthr);
    ts.sync();
    tail_restrict_sum(*F.G, C, res, ts.tid, ts.nthr);
    ts.sync();
}

// one level of the way up: prolongation, then post-smoothing
template <typename VF, typename VC>
__device__ __forceinline__ void tail_up(const TailLevels &tl, const VF &F, const VC &C, TailSync &ts)
{
    if (F.staged) tail_prolong<true>(F, C, ts.tid, ts.nthr);
    else tail_prolong<false>(F, C, ts.tid, ts.nthr);
    ts.sync();
    for (int sw = 0; sw < tl.sweeps; sw++) {
        tail_smooth(F, 1, 0, ts);
        tail_smooth(F, 0, 0, ts);
    }
}

template <typename T>
struct KB_mg_tail {
    static __device__ __forceinline__ void run(const VB &vb, const TailLevels &tl, const MgState *m)
{
    extern __shared__ __align__(16) unsigned char tail_smem[];
    if (stopped(m)) return;
    // a group (global levels shared by G workgroups on XCD tl.xcd): this
    // block's member index w; member 0 also runs the LDS levels
    const unsigned G = tl.groups > 1 && tl.first < tl.lds_from ? (unsigned)tl.groups : 1u;
    uint32_t w = 0;
    if (G > 1) {
        if ((int)(blockIdx.x % kXcds) != tl.xcd) return;
        w = blockIdx.x / kXcds;
    }
    const bool lead = w == 0;
    TailSync solo = tail_solo();
    TailSync grp = solo;
    if (G > 1) {
        grp.tid = w * kTailThreads + threadIdx.x;
        grp.nthr = G * kTailThreads;
        grp.ctr = tl.gsync;
        grp.fault = tl.fault;
        grp.G = G;
    }
    const int lf = max(tl.first, tl.lds_from);  // the first level in LDS
    if (lead) {
        for (int l = lf; l <= tl.last; l++) {  // stage the operators, clear the pads
            const MgLevel &Lv = tl.lv[l];
            const TailView<T, false> L = tail_view<T, false>(tl, tail_smem, l);
            if (L.staged) {
                float *cx = reinterpret_cast<float *>(L.base + L.tc.cx),
                      *cy = reinterpret_cast<float *>(L.base + L.tc.cy),
                      *cz = reinterpret_cast<float *>(L.base + L.tc.cz),
                      *dg = reinterpret_cast<float *>(L.base + L.tc.diag),
                      *rd = reinterpret_cast<float *>(L.base + L.tc.rdiag);
                uint8_t *act = L.base + L.tc.act;
                for (uint32_t c = threadIdx.x; c < L.n; c += kTailThreads) {
                    cx[c] = Lv.cx[c];
                    cy[c] = Lv.cy[c];
                    cz[c] = Lv.cz[c];
                    dg[c] = Lv.diag[c];
                    rd[c] = Lv.rdiag[c];
                    act[c] = Lv.active[c];
                }
                for (uint32_t c = L.n + threadIdx.x; c < L.n + L.si; c += kTailThreads) cx[c] = 0.f;
                for (uint32_t c = L.n + threadIdx.x; c < L.n + L.sj; c += kTailThreads) cy[c] = 0.f;
                if (threadIdx.x == 0) cz[L.n] = 0.f;
            }
            for (uint32_t c = threadIdx.x; c < L.si; c += kTailThreads) {
                stv(L.z - L.si, c, 0.f);  // the pads below and above the level
                stv(L.z + L.n, c, 0.f);
            }
        }
        if (tl.first >= tl.lds_from) {  // the first level's r into LDS
            const TailView<T, false> F = tail_view<T, false>(tl, tail_smem, tl.first);
            const T *rg = (const T *)tl.lv[tl.first].r;
            for (uint32_t c = threadIdx.x; c < F.n; c += kTailThreads) F.r[c] = rg[c];
        }
    }
    __syncthreads();
    float *const lres = reinterpret_cast<float *>(tail_smem + tl.scratch_off);
    // down: pre-smoothing and restriction of the global levels (residuals to
    // the global scratch); a group restricts the last of them into the
    // global r of the first LDS level, which its lead then loads into LDS
    for (int l = tl.first; l < tl.lds_from && l < tl.last; l++) {
        if (l + 1 < tl.lds_from || G > 1)
            tail_down(tl, tail_view<T, true>(tl, tail_smem, l), tail_view<T, true>(tl, tail_smem, l + 1),
                      tl.gscratch, grp);
        else
            tail_down(tl, tail_view<T, true>(tl, tail_smem, l), tail_view<T, false>(tl, tail_smem, l + 1),
                      tl.gscratch, grp);
    }
    if (G > 1 && lead) {  // the group's restriction into the first LDS level
        const TailView<T, false> C = tail_view<T, false>(tl, tail_smem, lf);
        const T *rg = (const T *)tl.lv[lf].r;
        for (uint32_t c = threadIdx.x; c < C.n; c += kTailThreads) C.r[c] = rg[c];
        __syncthreads();
    }
    if (lead) {
        // the LDS levels (LDS residual scratch)
        for (int l = lf; l < tl.last; l++)
            tail_down(tl, tail_view<T, false>(tl, tail_smem, l), tail_view<T, false>(tl, tail_smem, l + 1),
                      lres, solo);
        // coarsest: z = G r with the dense inverse, or (red, black)^s then
        // (black, red)^s from z = 0
        {
            const TailView<T, false> L = tail_view<T, false>(tl, tail_smem, tl.last);
            if (tl.ginv) {
                // thread i forms row i; G is symmetric, so reading G[j][i] for
                // consecutive i is a coalesced read of row j; r_j is broadcast
                const uint32_t i = threadIdx.x;
                float acc = 0.f;
                if (i < L.n) {
                    const float *gcol = tl.ginv + i;
                    uint32_t j = 0;
                    for (; j + 8 <= L.n; j += 8) {
                        float gv[8];
#pragma unroll
                        for (int t = 0; t < 8; t++) gv[t] = gcol[(size_t)(j + t) * L.n];
#pragma unroll
                        for (int t = 0; t < 8; t++) acc += gv[t] * ldv(L.r, j + t);
                    }
                    for (; j < L.n; j++) acc += gcol[(size_t)j * L.n] * ldv(L.r, j);
                    stv(L.z, i, acc);
                }
                __syncthreads();
            } else if (threadIdx.x < kWave) {
                // one wave: its LDS accesses are ordered, so passes need no
                // workgroup barrier (a wavefront fence orders them for the compiler)
                for (int pass = 0; pass < 4 * tl.coarse_sweeps; pass++) {
                    const int half = pass / (2 * tl.coarse_sweeps);
                    const uint32_t col = (uint32_t)(((pass & 1) ^ half) & 1);
                    if (L.staged) tail_pass<true>(L, col, pass == 0, threadIdx.x, kWave);
                    else tail_pass<false>(L, col, pass == 0, threadIdx.x, kWave);
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                }
            }
            if (!tl.ginv) __syncthreads();
        }
        for (int l = tl.last - 1; l >= lf; l--)
            tail_up(tl, tail_view<T, false>(tl, tail_smem, l), tail_view<T, false>(tl, tail_smem, l + 1),
                    solo);
        if (G > 1) {  // the first LDS level's z to its global array, for the group's prolongation
            const TailView<T, false> C = tail_view<T, false>(tl, tail_smem, lf);
            T *zg = (T *)tl.lv[lf].z;
            for (uint32_t c = threadIdx.x; c < C.n; c += kTailThreads) zg[c] = C.z[c];
        }
    }
    // up: prolongation and post-smoothing of the global levels
    if (G > 1) grp.sync();  // the lead's z of the first LDS level is in global memory
    for (int l = min(tl.lds_from, tl.last) - 1; l >= tl.first; l--) {
        if (l + 1 < tl.lds_from || G > 1)
            tail_up(tl, tail_view<T, true>(tl, tail_smem, l), tail_view<T, true>(tl, tail_smem, l + 1), grp);
        else
            tail_up(tl, tail_view<T, true>(tl, tail_smem, l), tail_view<T, false>(tl, tail_smem, l + 1), grp);
    }
    if (tl.first >= tl.lds_from && lead) {  // the first level's z back to global memory
        const TailView<T, false> F = tail_view<T, false>(tl, tail_smem, tl.first);
        T *zg = (T *)tl.lv[tl.first].z;
        for (uint32_t c = threadIdx.x; c < F.n; c += kTailThreads) zg[c] = F.z[c];
    }
    if (G > 1 && threadIdx.x == 0) {
        // every member has left its last wait once it departs: the last one
        // to depart resets the counters for the next launch
        const unsigned d = __hip_atomic_fetch_add(tl.gsync + 1, 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        if (d == G - 1) {
            __hip_atomic_store(tl.gsync, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(tl.gsync + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}
};
template <typename T>
__global__ void __launch_bounds__(kTailThreads) k_mg_tail(TailLevels tl, const MgState *m)
{
    KB_mg_tail<T>::run(vb_here(), tl, m);
}

// ---- inner PCG on level 0 (FP32 vectors, FP64 dots) -------------------------
// Iteration it (m->inner_it) of an inner solve starts after the V-cycle of
// the previous one left z = M r and the partials of r.z (gdot of them):
//   beta = r.z / (r.z before); p = z + beta p (p = z on the first iteration,
//   whose r.z is recorded as rz[0]); q = A p; partials of p.q.
//
// Main form: the V-cycle ended with a red pass on level 0, which set
// z_red = (r_red + offdiag(z_black)) / diag. So (A z) on red cells is r_red
// up to that pass's rounding, and q comes from the recurrence
// q = A z + beta q: red slots gather nothing, black slots gather the six red
// z; p and q are updated in place. The recurrence drifts from A p only at
// FP32 rounding level over an inner solve's few iterations, and the outer
// FP64 residual is recomputed from scratch every pass. Thread work pairs red
// slot h with black slot H + h, so every block streams both colours.
template <bool ST>
struct KB_pcg_dirq {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const MgLevel &L,
                                               const float *__restrict__ z,
                                               const float *__restrict__ r, float *p, float *q,
                                               MgState *m, const RedScratch &red, uint32_t gdot)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t stop = stop_flags(m);  // tested after the reduction's loads
    const int it = m->inner_it;
    const double rz_prev = m->rz[(it + 1) & 1];  // rz of iteration it - 1 (unused at 0)
    const double rz_new = sum_partials(pslot(red, SLOT_MG_A), gdot, lds);
    if (late(stop)) return;  // uniform: every thread of the grid returns
    const bool first = it == 0;
    const float beta = first ? 0.f : (float)(rz_new / rz_prev);
    const uint32_t blk = vb.blk, H = L.H, nb_black = g.nc - H;
    const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz;
    const float dint = diag0m(g, kInterior);
    // red views [0, H), black views [H, nc) (reads past an end return 0 and
    // stores there are dropped)
    const auto cr = buf_rsrc(L.nbs, 2u * H), cb = buf_rsrc(L.nbs + H, 2u * nb_black),
               zr = buf_rsrc(z, 4u * H), zb = buf_rsrc(z + H, 4u * nb_black),
               rr = buf_rsrc(r, 4u * H), pr = buf_rsrc(p, 4u * H),
               pb = buf_rsrc(p + H, 4u * nb_black), qr = buf_rsrc(q, 4u * H),
               qb = buf_rsrc(q + H, 4u * nb_black);
    double acc = 0.0;
    uint32_t ch0, cst, cend;
    xcd_chunks(H, blk, vb.G, ch0, cst, cend);
    for (uint32_t ch = ch0; ch < cend; ch += kU2 * cst) {
        uint16_t nr[kU2], nk[kU2];
        float zrv[kU2], rv[kU2], prv[kU2], qrv[kU2], zbv[kU2], pbv[kU2], qbv[kU2], zn[kU2][6];
        uint32_t cnat[kU2];  // natural cell index (stretched grids)
#pragma unroll
        for (int t = 0; t < kU2; t++) {
            const uint32_t h = (ch + t * cst) * kBlock + threadIdx.x;  // reads past an end return 0
            const uint32_t cn = 2 * h + split_parity(L, h, 1u);
            nr[t] = buf_ldu16(cr, h);
            zrv[t] = buf_ldf(zr, h);
            rv[t] = buf_ldf(rr, h);
            nk[t] = buf_ldu16(cb, h);
            zbv[t] = buf_ldf(zb, h);
            cnat[t] = cn;
            nb_gather(zr, cn, si, sj, zn[t]);
            if (!first) {
                prv[t] = buf_ldf(pr, h);
                qrv[t] = buf_ldf(qr, h);
                pbv[t] = buf_ldf(pb, h);
                qbv[t] = buf_ldf(qb, h);
            }
        }
#pragma unroll
        for (int t = 0; t < kU2; t++) {
            const uint32_t h = (ch + t * cst) * kBlock + threadIdx.x;
            if (ch + t * cst >= cend) break;  // slots past H: zero reads, dropped stores
            const float azr = (nr[t] & NB_SELF) ? rv[t] : 0.f;
            float azb = 0.f;
            if (nk[t] & NB_SELF) {
                const Cof<ST> cf = cof_at<ST>(g, cnat[t]);
                azb = diag_t<ST>(g, nk[t], dint, cf) * zbv[t] - nb_sum_t<ST>(g, nk[t], zn[t], cf);
            }
            float p0 = zrv[t], q0 = azr, p1 = zbv[t], q1 = azb;
            if (!first) {
                p0 = __builtin_fmaf(beta, prv[t], p0);
                q0 = __builtin_fmaf(beta, qrv[t], q0);
                p1 = __builtin_fmaf(beta, pbv[t], p1);
                q1 = __builtin_fmaf(beta, qbv[t], q1);
            }
            if (nr[t] & NB_SELF) acc += (double)p0 * (double)q0;
            if (nk[t] & NB_SELF) acc += (double)p1 * (double)q1;
            buf_stf(pr, h, p0);
            buf_stf(qr, h, q0);
            buf_stf(pb, h, p1);
            buf_stf(qb, h, q1);
        }
    }
    block_partial(acc, pslot(red, SLOT_SPMV), blk, lds);
    if (blk == 0 && threadIdx.x == 0) m->rz[it & 1] = rz_new;
}
};
template <bool ST>
__global__ void __launch_bounds__(kBlock) k_pcg_dirq(Geo g, MgLevel L, const float *__restrict__ z,
                                                     const float *__restrict__ r, float *p,
                                                     float *q, MgState *m, RedScratch red,
                                                     uint32_t gdot)
{
    KB_pcg_dirq<ST>::run(vb_here(), g, L, z, r, p, q, m, red, gdot);
}

// General form (level 0 is the tail or the coarsest level, small grids):
// p = z + beta p in place, then q = A p with gathers (k_pcg_spmv).
struct KB_pcg_dir {
    static __device__ __forceinline__ void run(const VB &vb, uint32_t n, const float *z, float *p,
                                               MgState *m, const RedScratch &red, uint32_t gdot)
{
    __shared__ double lds[kBlock / kWave];
    if (stopped(m)) return;
    const int it = m->inner_it;
    const bool first = it == 0;
    const double rz_new = sum_partials(pslot(red, SLOT_MG_A), gdot, lds);
    const float beta = first ? 0.f : (float)(rz_new / m->rz[(it + 1) & 1]);
    const uint32_t blk = vb.blk;
    uint32_t sb, se;
    block_span_of(n, blk, vb.G, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock)
        p[c] = first ? z[c] : __builtin_fmaf(beta, p[c], z[c]);
    if (blk == 0 && threadIdx.x == 0) m->rz[it & 1] = rz_new;
}
};
__global__ void __launch_bounds__(kBlock) k_pcg_dir(uint32_t n, const float *z, float *p, MgState *m,
                                                    RedScratch red, uint32_t gdot)
{
    KB_pcg_dir::run(vb_here(), n, z, p, m, red, gdot);
}

// q = A p, partials of p.q (split layout: slot s of colour s >= H); loads as
// in k_mg_smooth0
template <bool ST>
struct KB_pcg_spmv {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const MgLevel &L,
                                               const float *__restrict__ p, float *__restrict__ q,
                                               const MgState *m, const RedScratch &red)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t stop = stop_flags(m);  // tested after the first loads
    const uint32_t blk = vb.blk, H = L.H;
    const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz;
    const float dint = diag0m(g, kInterior);
    const auto bcode = buf_rsrc(L.nbs, 2u * g.nc), bp = buf_rsrc(p, 4u * g.nc),
               bred = buf_rsrc(p, 4u * H), bblack = buf_rsrc(p + H, 4u * (g.nc - H)),
               bq = buf_rsrc(q, 4u * g.nc);
    double acc = 0.0;
    uint32_t sb, se;
    block_span_of(g.nc, blk, vb.G, sb, se);
    for (uint32_t s0 = sb + threadIdx.x; s0 < se; s0 += kU * kBlock) {
        uint16_t nb[kU];
        float pv[kU], zv[kU][6];
        uint32_t cnat[kU];  // natural cell index (stretched grids)
#pragma unroll
        for (int t = 0; t < kU; t++) {
            const uint32_t s = s0 + t * kBlock;
            const uint32_t col = s >= H ? 1u : 0u;
            const uint32_t h = s - (col ? H : 0u);
            const uint32_t cn = 2 * h + split_parity(L, h, col);
            nb[t] = buf_ldu16(bcode, s);
            pv[t] = buf_ldf(bp, s);
            cnat[t] = cn;
            nb_gather(col ? bred : bblack, cn, si, sj, zv[t]);
        }
        if (late(stop)) return;  // uniform: every thread of the grid returns
#pragma unroll
        for (int t = 0; t < kU; t++) {
            const uint32_t s = s0 + t * kBlock;
            if (s >= se) break;
            float qv = 0.f;
            if (nb[t] & NB_SELF) {
                const Cof<ST> cf = cof_at<ST>(g, cnat[t]);
                qv = diag_t<ST>(g, nb[t], dint, cf) * pv[t] - nb_sum_t<ST>(g, nb[t], zv[t], cf);
                acc += (double)pv[t] * (double)qv;
            }
            buf_stf(bq, s, qv);
        }
    }
    block_partial(acc, pslot(red, SLOT_SPMV), blk, lds);
}
};
template <bool ST>
__global__ void __launch_bounds__(kBlock) k_pcg_spmv(Geo g, MgLevel L, const float *__restrict__ p,
                                                     float *__restrict__ q, const MgState *m,
                                                     RedScratch red)
{
    KB_pcg_spmv<ST>::run(vb_here(), g, L, p, q, m, red);
}

// alpha = rz / p.q; e += alpha p (e = alpha p on an inner solve's first
// iteration); r -= alpha q; partials of r.r
//
// The V-cycle's first pass, the red sweep from z = 0 on level 0, is done here:
// z = r / diag on red unknowns (red slots [0, H) of the split layout), 0 on
// other red slots; the V-cycle then starts with the black sweep.
template <bool ST>
struct KB_pcg_update {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const MgLevel &L,
                                               float *__restrict__ e, float *__restrict__ r,
                                               const float *__restrict__ p,
                                               const float *__restrict__ q, float *__restrict__ z,
                                               MgState *m, const RedScratch &red, uint32_t gq)
{
    __shared__ double lds[kBlock / kWave];
    if (stopped(m)) return;
    const uint32_t blk = vb.blk;
    const int it = m->inner_it;
    const bool first = it == 0;
    uint32_t sb, se;
    block_span_of(g.nc, blk, vb.G, sb, se);
    const double pq = sum_partials(pslot(red, SLOT_SPMV), gq, lds);  // gq p.q partials
    if (!(pq > 0.0)) {  // breakdown or a non-positive operator: end the inner solve
        if (first)      // the correction of an inner solve that made no step is 0
            for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) e[c] = 0.f;
        if (blk == 0 && threadIdx.x == 0) m->inner_done = 1;
        return;
    }
    const float alpha = (float)(m->rz[it & 1] / pq);
    const float dint = diag0m(g, kInterior);
    double acc = 0.0;
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        // the four distinct vectors (restrict) and a red slot's code: every
        // load goes out before a store (black slots read slot 0's code, unused)
        const float pv = p[c], qv = q[c], r0 = r[c], e0 = first ? 0.f : e[c];
        const uint16_t nb = L.nbs[c < L.H ? c : 0u];
        e[c] = e0 + alpha * pv;
        const float rv = r0 - alpha * qv;
        r[c] = rv;
        acc += (double)rv * (double)rv;
        if (c < L.H) {  // red slot
            const Cof<ST> cf = cof_at<ST>(g, 2 * c + split_parity(L, c, 0u));
            z[c] = (nb & NB_SELF) ? rv / diag_t<ST>(g, nb, dint, cf) : 0.f;
        }
    }
    block_partial(acc, pslot(red, SLOT_UPDATE), blk, lds);
}
};
template <bool ST>
__global__ void __launch_bounds__(kBlock) k_pcg_update(Geo g, MgLevel L, float *e, float *r,
                                                       const float *p, const float *q, float *z,
                                                       MgState *m, RedScratch red, uint32_t gq)
{
    KB_pcg_update<ST>::run(vb_here(), g, L, e, r, p, q, z, m, red, gq);
}

// stopping test of the inner solve (one block): its relative target, or the
// outer target reached already
struct KB_pcg_check {
    static __device__ __forceinline__ void run(const VB &vb, MgState *m, const RedScratch &red,
                                               uint32_t G)
{
    __shared__ double lds[kBlock / kWave];
    if (stopped(m)) return;
    const double rr = sum_partials(pslot(red, SLOT_UPDATE), G, lds);
    if (threadIdx.x != 0) return;
    const int it = m->inner_it + 1;
    m->inner_it = it;
    m->total_iters += 1;
    // the inner residual is the next outer residual up to FP32 rounding (of
    // order 1e-7 of rnorm0): once it is below a quarter of the outer target
    // (measured: a half lets some steps miss it and take one more pass), the
    // inner solve has done what the outer test needs
    const double rn = sqrt(rr);
    if (rn <= m->inner_tol * m->rnorm0 || rn <= 0.25 * m->outer_tol * m->bnorm ||
        it >= m->max_inner || !(rr == rr) || m->total_iters >= m->total_cap)
        m->inner_done = 1;
}
};
__global__ void __launch_bounds__(kBlock) k_pcg_check(MgState *m, RedScratch red, uint32_t G)
{
    KB_pcg_check::run(vb_here(), m, red, G);
}

// partials of r.z (after a V-cycle whose level 0 is the tail or the coarsest
// level; otherwise its last red pass writes them)
struct KB_pcg_dot_rz {
    static __device__ __forceinline__ void run(const VB &vb, uint32_t n, const float *r,
                                               const float *z, const MgState *m,
                                               const RedScratch &red)
{
    __shared__ double lds[kBlock / kWave];
    if (stopped(m)) return;
    const uint32_t blk = vb.blk;
    double acc = 0.0;
    uint32_t sb, se;
    block_span_of(n, blk, vb.G, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) acc += (double)r[c] * (double)z[c];
    block_partial(acc, pslot(red, SLOT_MG_A), blk, lds);
}
};
__global__ void __launch_bounds__(kBlock) k_pcg_dot_rz(uint32_t n, const float *r, const float *z,
                                                       const MgState *m, RedScratch red)
{
    KB_pcg_dot_rz::run(vb_here(), n, r, z, m, red);
}

// ---- outer FP64 defect correction -------------------------------------------
// x = 0 (PoissonSolver.cpp:96) or, warm, the pressure x already holds (the
// previous step's: SOLID cells 0, OUTFLOW values never read as unknowns);
// partials of ||b||^2 over non-SOLID cells (PoissonSolver.cpp:101-113); the
// outer correction e needs no clearing: the first outer pass does not read it
// and each inner solve's first update writes it
// The solve's start, with x holding p_n, the last step's pressure:
// warm 0: x = 0; 1: x keeps p_n. With `out` (a spare buffer of the
// scene's, which the host then makes the solve's x, while x's buffer keeps
// p_n as history): warm 1: out = p_n; 2: out = 2 p_n - p_{n-1}; 3: out =
// 3 p_n - 3 p_{n-1} + p_{n-2} (linear and quadratic extrapolation in time,
// p_{n-1} in h1; p_{n-2} is out's own old content). Nothing is copied to
// shift the history: the host rotates the three buffers' roles.
struct KB_dc_init {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g, const double *b,
                                               double *x, int warm, const double *h1,
                                               double *out, const RedScratch &red)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t blk = vb.blk;
    double acc = 0.0;
    uint32_t sb, se;
    block_span_of(g.nc, blk, vb.G, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        if (!warm) {
            x[c] = 0.0;
        } else if (out) {
            const double pn = x[c];
            if (warm == 1) out[c] = pn;
            else if (warm == 2) out[c] = 2.0 * pn - h1[c];
            else out[c] = 3.0 * (pn - h1[c]) + out[c];
        }
        if (g.flags[c] != SOLID) acc += b[c] * b[c];
    }
    block_partial(acc, pslot(red, SLOT_INIT), blk, lds);
}
};
__global__ void __launch_bounds__(kBlock) k_dc_init(Geo g, const double *b, double *x, int warm,
                                                    const double *h1, double *out, RedScratch red)
{
    KB_dc_init::run(vb_here(), g, b, x, warm, h1, out, red);
}

struct KB_dc_start {
    static __device__ __forceinline__ void run(const VB &vb, MgState *m, const RedScratch &red,
                                               uint32_t G, double outer_tol, int max_outer,
                                               int max_inner, double inner_tol, int total_cap)
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
    m->total_cap = total_cap;
    m->passes = 0;
    for (int k = 0; k < kMgHist; k++) m->hist[k] = 0;
    m->inner_it = 0;
    m->inner_done = 1;
    m->solve_done = m->bnorm < 1e-15 ? 1 : 0;
    m->res = 0.0;
}
};
__global__ void __launch_bounds__(kBlock) k_dc_start(MgState *m, RedScratch red, uint32_t G,
                                                     double outer_tol, int max_outer, int max_inner,
                                                     double inner_tol, int total_cap)
{
    KB_dc_start::run(vb_here(), m, red, G, outer_tol, max_outer, max_inner, inner_tol, total_cap);
}

// FP64 reference row (PoissonSolver.cpp:52-87) from the neighbour code, with
// one scale per axis on a uniform grid or the reference's per-face scales
// (PoissonSolver.cpp:34-49, Geo.lap) on a stretched one (ST); neighbour
// values come from `val` (x + e for the outer residual, x for the OUTFLOW
// back substitution).
template <bool ST, typename F>
__device__ __forceinline__ void row64(const Geo &g, uint32_t c, uint16_t m, F val, double &diag,
                                      double &off)
{
    const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz;
    double sxl, sxr, syl, syr, szl, szr;
    if constexpr (ST) {
        const uint32_t i = fdiv(c, g.f_nynz), rem = c - i * si;
        const uint32_t j = fdiv(rem, g.f_nz), k = rem - j * sj;
        const double *L = g.lap;
        sxl = L[i]; sxr = L[g.nx + i];
        syl = L[2 * g.nx + j]; syr = L[2 * g.nx + g.ny + j];
        szl = L[2 * g.nx + 2 * g.ny + k]; szr = L[2 * g.nx + 2 * g.ny + g.nz + k];
    } else {
        sxl = sxr = g.sxd;
        syl = syr = g.syd;
        szl = szr = g.szd;
    }
    double dg = 0.0, of = 0.0;
#define FFR_FACE(q, nc, sc)                                                   \
    if (m & (NB_UNK0 << q)) { dg += (sc); of += (sc) * val(nc); }            \
    else if (m & (NB_OUT0 << q)) dg += 2.0 * (sc);
    FFR_FACE(0, c - si, sxl)
    FFR_FACE(1, c + si, sxr)
    FFR_FACE(2, c - sj, syl)
    FFR_FACE(3, c + sj, syr)
    FFR_FACE(4, c - 1, szl)
    FFR_FACE(5, c + 1, szr)
#undef FFR_FACE
    diag = dg < 1e-30 ? 1.0 : dg;
    off = of;
}

// The same row from the six neighbour values already gathered (faces in the
// order -x, +x, -y, +y, -z, +z; a value whose face does not couple is not
// used), (i, j, k) = ijk for the per-face scales of a stretched grid.
template <bool ST>
__device__ __forceinline__ void row64v(const Geo &g, const uint32_t ijk[3], uint16_t m,
                                       const double val[6], double &diag, double &off)
{
    double sc[6];
    if constexpr (ST) {
        const double *L = g.lap;
        sc[0] = L[ijk[0]]; sc[1] = L[g.nx + ijk[0]];
        sc[2] = L[2 * g.nx + ijk[1]]; sc[3] = L[2 * g.nx + g.ny + ijk[1]];
        sc[4] = L[2 * g.nx + 2 * g.ny + ijk[2]]; sc[5] = L[2 * g.nx + 2 * g.ny + g.nz + ijk[2]];
    } else {
        sc[0] = sc[1] = g.sxd;
        sc[2] = sc[3] = g.syd;
        sc[4] = sc[5] = g.szd;
    }
    double dg = 0.0, of = 0.0;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        if (m & (NB_UNK0 << q)) { dg += sc[q]; of += sc[q] * val[q]; }
        else if (m & (NB_OUT0 << q)) dg += 2.0 * sc[q];
    }
    diag = dg < 1e-30 ? 1.0 : dg;
    off = of;
}

// x_out = x_in + e on unknowns; r = b - A x_out (FP64); r32 = (float) r,
// weighted by the cell's relative volume on a stretched grid (the inner
// solve's symmetric form); partials of ||r||^2.
// Each thread takes kU2 cells of its block's chunks (xcd_chunks) and issues every load of
// both (code, b, x, own e, the six x and e neighbours) before using any: the
// colour of a cell, which places its e values in the split layout, comes
// from its coordinates rather than from its code, and the neighbour gathers
// go through buffer resources (out of range reads give 0; values of faces
// that do not couple are dropped by the code, in the row's face order). An
// interior row (six unknown neighbours, uniform grid) sums its neighbours per
// axis before scaling, fewer FP64 operations than the general row.
//
// P0: the solve's first pass, which the host knows it launches: e is 0 then,
// so x stays where it is (xa), e is not read and x not written. Pass k >= 1
// reads x from xa for odd k and from xb for even k and writes the other, so
// after P passes x is in xb exactly when P is even.
template <bool ST, bool P0>
struct KB_dc_residual {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g,
                                               const double *__restrict__ b, double *xa, double *xb,
                                               const float *__restrict__ e, float *__restrict__ r32,
                                               float *__restrict__ z0, const MgState *m,
                                               const RedScratch &red)
{
    __shared__ double lds[kBlock / kWave];
    // the solve's stop flag, tested after the first loads (see stop_flags)
    const uint32_t stop = __builtin_amdgcn_raw_buffer_load_b32(
        buf_rsrc(m, sizeof(MgState)), (int)offsetof(MgState, solve_done), 0, 0);
    // x_in is the buffer of the current pass parity, x_out the other one
    const bool odd = P0 || (m->passes & 1);
    const double *__restrict__ xin = odd ? xa : xb;
    double *__restrict__ xout = odd ? xb : xa;  // not written by P0
    const uint32_t blk = vb.blk;
    const uint32_t si = (uint32_t)g.ny * g.nz, sj = g.nz, H = (g.nc + 1) / 2;
    // the diagonal of an interior row, summed in row64v's face order
    const double dint64 = ((((g.sxd + g.sxd) + g.syd) + g.syd) + g.szd) + g.szd;
    const auto bx = buf_rsrc(xin, 8u * g.nc), be = buf_rsrc(e, 4u * g.nc);
    const auto bred = buf_rsrc(e, 4u * H), bblack = buf_rsrc(e + H, 4u * (g.nc - H));
    double acc = 0.0;
    uint32_t ch0, cstep, cend;
    xcd_chunks(g.nc, blk, vb.G, ch0, cstep, cend);
    for (uint32_t ch = ch0; ch < cend; ch += kU2 * cstep) {
        const uint32_t c0 = ch * kBlock + threadIdx.x;
        uint16_t nb[kU2];
        double bv[kU2], xv[kU2], xn[kU2][6];
        float ev[kU2], en[kU2][6];
        uint32_t cc[kU2], col[kU2], ijk[kU2][3];
        bool in[kU2];
#pragma unroll
        for (int t = 0; t < kU2; t++) {
            const uint32_t ct = (ch + t * cstep) * kBlock + threadIdx.x;
            in[t] = ch + t * cstep < cend && ct < g.nc;
            const uint32_t c = in[t] ? ct : (c0 < g.nc ? c0 : g.nc - 1);
            const uint32_t i = fdiv(c, g.f_nynz), rem = c - i * si;
            const uint32_t j = fdiv(rem, g.f_nz), k = rem - j * sj;
            cc[t] = c;
            col[t] = (i + j + k) & 1u;
            ijk[t][0] = i; ijk[t][1] = j; ijk[t][2] = k;
            nb[t] = g.nbm[c];
            bv[t] = b[c];
            xv[t] = xin[c];
            ev[t] = P0 ? 0.f : e[(col[t] ? H : 0u) + (c >> 1)];
            const auto bo = col[t] ? bred : bblack;  // the other colour's e
            const uint32_t q[6] = {c - si, c + si, c - sj, c + sj, c - 1u, c + 1u};
#pragma unroll
            for (int f = 0; f < 6; f++) {
                xn[t][f] = buf_ldd(bx, q[f]);
                en[t][f] = P0 ? 0.f : buf_ldf(bo, q[f] >> 1);
            }
        }
        if (late(stop)) return;
#pragma unroll
        for (int t = 0; t < kU2; t++) {
            if (!in[t]) continue;
            const uint32_t c = cc[t];
            const uint32_t sc = (col[t] ? H : 0u) + (c >> 1);  // e and r32: split layout
            const bool red_slot = col[t] == 0u;
            if (!(nb[t] & NB_SELF)) {
                if (!P0) xout[c] = 0.0;
                r32[sc] = 0.f;
                if (red_slot) z0[sc] = 0.f;
                continue;
            }
            double vals[6];
#pragma unroll
            for (int f = 0; f < 6; f++) vals[f] = xn[t][f] + (double)en[t][f];
            double diag, off;
            if (!ST && (nb[t] & kFaceBits) == kInterior) {  // six unknown neighbours
                diag = dint64;
                off = g.sxd * (vals[0] + vals[1]) + g.syd * (vals[2] + vals[3]) + g.szd * (vals[4] + vals[5]);
            } else {
                row64v<ST>(g, ijk[t], nb[t], vals, diag, off);
            }
            const double xc = xv[t] + (double)ev[t];
            const double rv = bv[t] - (diag * xc - off);
            if (!P0) xout[c] = xc;
            if constexpr (ST) {
                const double vol = (double)g.mgf[ijk[t][0]].x * (double)g.mgf[g.nx + ijk[t][1]].x *
                                   (double)g.mgf[g.nx + g.ny + ijk[t][2]].x;
                const float rf = (float)(rv * vol);
                r32[sc] = rf;
                if (red_slot) z0[sc] = rf / diag_t<ST>(g, nb[t], 0.f, cof_at<ST>(g, c));
            } else {
                const float rf = (float)rv;
                r32[sc] = rf;
                if (red_slot) z0[sc] = rf / diag0m(g, nb[t]);  // the V-cycle's first (red, z = 0) sweep
            }
            acc += rv * rv;
        }
    }
    block_partial(acc, pslot(red, SLOT_UPDATE), blk, lds);
}
};
template <bool ST, bool P0>
__global__ void __launch_bounds__(kBlock) k_dc_residual(Geo g, const double *__restrict__ b,
                                                        double *xa, double *xb,
                                                        const float *__restrict__ e,
                                                        float *__restrict__ r32,
                                                        float *__restrict__ z0, const MgState *m,
                                                        RedScratch red)
{
    KB_dc_residual<ST, P0>::run(vb_here(), g, b, xa, xb, e, r32, z0, m, red);
}

// outer stopping test: ||r|| / ||b|| < tol (one block)
struct KB_dc_check {
    static __device__ __forceinline__ void run(const VB &vb, MgState *m, const RedScratch &red,
                                               uint32_t G)
{
    __shared__ double lds[kBlock / kWave];
    if (m->solve_done) return;
    const double rr = sum_partials(pslot(red, SLOT_UPDATE), G, lds);
    if (threadIdx.x != 0) return;
    const double rn = sqrt(rr);
    const int pass = m->passes + 1;  // the residual kernel before this one ran
    m->passes = pass;
    if (pass >= 2 && pass - 2 < kMgHist) m->hist[pass - 2] = m->inner_it;
    m->res = rn / m->bnorm;
    if (m->res < m->outer_tol || m->outer >= m->max_outer || !(rr == rr) ||
        m->total_iters >= m->total_cap) {
        m->solve_done = 1;
        return;
    }
    m->outer += 1;
    m->rnorm0 = rn;
    m->inner_it = 0;
    m->inner_done = 0;
}
};
__global__ void __launch_bounds__(kBlock) k_dc_check(MgState *m, RedScratch red, uint32_t G)
{
    KB_dc_check::run(vb_here(), m, red, G);
}

// p: unknowns from x; OUTFLOW rows solved from their neighbours (or pinned
// to 0 for the reference multigrid's convention); SOLID cells 0. In place
// (xa == p holds the last pass's output, which the host knows from the pass
// count): the unknowns are not touched, only SOLID and OUTFLOW cells written.
template <bool ST>
struct KB_dc_finalize {
    static __device__ __forceinline__ void run(const VB &vb, const Geo &g,
                                               const double *__restrict__ b, const double *xa,
                                               const double *xb, double *p, int pin_outflow,
                                               int inplace, const MgState *m)
{
    // the last pass's output: xb after an even, non-zero number of passes
    const double *x = inplace ? p : ((m->passes > 0 && !(m->passes & 1)) ? xb : xa);
    // no pass: the right-hand side was 0 (||b|| < 1e-15), and the solution
    // is 0, as the reference CG returns from its zero start, whatever start
    // the solve was given
    const bool zero = m->passes == 0;
    uint32_t sb, se;
    block_span_of(g.nc, vb.blk, vb.G, sb, se);
    for (uint32_t c = sb + threadIdx.x; c < se; c += kBlock) {
        if (zero) {
            p[c] = 0.0;
            continue;
        }
        const uint16_t nb = g.nbm[c];
        if (nb & NB_SELF) {
            if (!inplace) p[c] = x[c];
            continue;
        }
        if ((nb & NB_SELF_SOLID) || pin_outflow) {
            p[c] = 0.0;
            continue;
        }
        double diag, off;
        row64<ST>(g, c, nb, [&](uint32_t q) { return x[q]; }, diag, off);
        p[c] = (b[c] + off) / diag;
    }
}
};
template <bool ST>
__global__ void __launch_bounds__(kBlock) k_dc_finalize(Geo g, const double *__restrict__ b,
                                                        const double *xa, const double *xb,
                                                        double *p, int pin_outflow, int inplace,
                                                        const MgState *m)
{
    KB_dc_finalize<ST>::run(vb_here(), g, b, xa, xb, p, pin_outflow, inplace, m);
}

// ---- launchers ------------------------------------------------------------------
void mg_launch_build(hipStream_t s, const Geo &g, const MgLevel &F, const MgLevel &C, bool fine0)
{
    if (fine0 && !g.uniform)
        FFR_LAUNCH((k_mg_build<true, true>), dim3(grid_for(C.n)), dim3(kBlock), 0, s, g, F, C);
    else if (fine0)
        FFR_LAUNCH((k_mg_build<true, false>), dim3(grid_for(C.n)), dim3(kBlock), 0, s, g, F, C);
    else
        FFR_LAUNCH((k_mg_build<false, false>), dim3(grid_for(C.n)), dim3(kBlock), 0, s, g, F, C);
}

void mg_launch_split_codes(hipStream_t s, const Geo &g, const MgLevel &L0)
{
    FFR_LAUNCH(k_mg_split_codes, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, L0.nbs, L0.H);
}

// level-0 launches on uniform (ST false) or stretched (ST true) spacing
template <bool ST>
void smooth0_launch(hipStream_t s, const Geo &g, const MgLevel &L, int color, int zero_start,
                    const MgState *m, const RedScratch *dot)
{
    // level 0 is FP32 in the split layout
    const uint32_t H = L.H, cnt = color ? g.nc - H : H, base = color ? H : 0u;
    const uint32_t obase = color ? 0u : H;
    float *z = (float *)L.z;
    const float *r = (const float *)L.r;
    if (dot && color == 0)
        FFR_GANG((KB_mg_smooth0<true, ST>), GB_RES, (k_mg_smooth0<true, ST>), dim3(grid_resident(cnt)),
                 dim3(kBlock), 0, s, g, L, L.nbs + base, z + obase, z + base, r + base, r + obase,
                 cnt, g.nc - cnt, (uint32_t)color, zero_start, m, *dot);
    else
        FFR_GANG((KB_mg_smooth0<false, ST>), GB_RES, (k_mg_smooth0<false, ST>),
                 dim3(grid_resident(cnt)), dim3(kBlock), 0, s, g, L, L.nbs + base, z + obase,
                 z + base, r + base, r + obase, cnt, g.nc - cnt, (uint32_t)color, zero_start, m,
                 RedScratch{});
}

template <typename T>
void mg_launch_smooth(hipStream_t s, const Geo &g, const MgLevel &L, bool l0, int color,
                      int zero_start, const MgState *m, const RedScratch *dot)
{
    if (l0 && g.uniform)
        smooth0_launch<false>(s, g, L, color, zero_start, m, dot);
    else if (l0)
        smooth0_launch<true>(s, g, L, color, zero_start, m, dot);
    else
        FFR_GANG((KB_mg_smooth<T>), GB_PLAIN, k_mg_smooth<T>, dim3(grid_for(L.n)), dim3(kBlock), 0,
                 s, L, (T *)L.z, (const T *)L.r, color, zero_start, m);
}

template <bool ST>
void resid0_launch(hipStream_t s, const Geo &g, const MgLevel &L0, float *res, const MgState *m)
{
    FFR_GANG((KB_mg_resid0<ST>), GB_RES, (k_mg_resid0<ST>), dim3(grid_resident(L0.H)), dim3(kBlock),
             0, s, g, L0, (const float *)L0.z, (const float *)L0.r, res, m);
}

template <typename TF, typename TC, bool ST>
void restrict0_launch(hipStream_t s, const Geo &g, const MgLevel &F, const MgLevel &C, bool red0,
                      const MgState *m)
{
    FFR_GANG((KB_mg_restrict<TF, TC, true, ST>), GB_PLAIN, (k_mg_restrict<TF, TC, true, ST>),
             dim3(grid_for(8ull * C.n)), dim3(kBlock), 0, s, g, F, C, (const TF *)F.z,
             (const TF *)F.r, (TC *)C.r, (TC *)C.z, (int)red0, m);
}

template <typename TF, typename TC>
void mg_launch_restrict(hipStream_t s, const Geo &g, const MgLevel &F, const MgLevel &C, bool l0,
                        bool red0, const MgState *m, float *scratch, int pass)
{
    if (l0 && scratch) {  // two passes through the scratch residual
        if (pass != 2) mg_launch_resid0(s, g, F, scratch, m);
        if (pass != 1)
            FFR_GANG((KB_mg_restrict_sum0<TC>), GB_PLAIN, (k_mg_restrict_sum0<TC>), dim3(grid_for(C.n)),
                     dim3(kBlock), 0, s, g, F, C, (const float *)scratch, (TC *)C.r, (TC *)C.z,
                     (int)red0, m);
    } else if (l0 && g.uniform)
        restrict0_launch<TF, TC, false>(s, g, F, C, red0, m);
    else if (l0)
        restrict0_launch<TF, TC, true>(s, g, F, C, red0, m);
    else if (scratch && restrict_two_pass(false, F.n)) {  // residual, then the child sums
        if (pass != 2)
            FFR_GANG((KB_mg_cresid<TF>), GB_PLAIN, (k_mg_cresid<TF>), dim3(grid_for(F.n)),
                     dim3(kBlock), 0, s, F, (const TF *)F.z, (const TF *)F.r, scratch, m);
        if (pass != 1)
            FFR_GANG((KB_mg_csum<TC>), GB_PLAIN, (k_mg_csum<TC>), dim3(grid_for(C.n)), dim3(kBlock),
                     0, s, F, C, (const float *)scratch, (TC *)C.r, (TC *)C.z, (int)red0, m);
    } else
        FFR_GANG((KB_mg_restrict<TF, TC, false, false>), GB_PLAIN, (k_mg_restrict<TF, TC, false, false>),
                 dim3(grid_for(8ull * C.n)), dim3(kBlock), 0, s, g, F, C, (const TF *)F.z,
                 (const TF *)F.r, (TC *)C.r, (TC *)C.z, (int)red0, m);
}

void mg_launch_resid0(hipStream_t s, const Geo &g, const MgLevel &L0, float *res, const MgState *m)
{
    if (g.uniform)
        resid0_launch<false>(s, g, L0, res, m);
    else
        resid0_launch<true>(s, g, L0, res, m);
}

template <typename TF, typename TC>
void mg_launch_post_black(hipStream_t s, const Geo &g, const MgLevel &F, const MgLevel &C, bool l0,
                          const MgState *m)
{
    if (l0 && g.uniform)
        FFR_GANG((KB_mg_post_black0<TC, false>), GB_RES, (k_mg_post_black0<TC, false>),
                 dim3(grid_resident(g.nc - F.H)), dim3(kBlock), 0, s, g, F, C, (const TC *)C.z, m);
    else if (l0)
        FFR_GANG((KB_mg_post_black0<TC, true>), GB_RES, (k_mg_post_black0<TC, true>),
                 dim3(grid_resident(g.nc - F.H)), dim3(kBlock), 0, s, g, F, C, (const TC *)C.z, m);
    else
        FFR_GANG((KB_mg_post_black<TF, TC>), GB_PLAIN, (k_mg_post_black<TF, TC>),
                 dim3(grid_for(F.n)), dim3(kBlock), 0, s, F, C, (TF *)F.z, (const TF *)F.r,
                 (const TC *)C.z, m);
}

// Dynamic LDS beyond 64 KiB must be allowed per kernel before a launch; done
// at hierarchy creation, outside any stream capture, with its status checked.
int mg_tail_allow_lds()
{
    if (hipFuncSetAttribute((const void *)k_mg_tail<float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTailLdsMax) != hipSuccess)
        return -1;
    if (hipFuncSetAttribute((const void *)k_mg_tail<__hip_bfloat16>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTailLdsMax) != hipSuccess)
        return -1;
    // and the gang launches of the same kernels
    if (hipFuncSetAttribute(gang_kernel<KB_mg_tail<float>, GB_WIDE, TailLevels, const MgState *>(),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTailLdsMax) != hipSuccess)
        return -1;
    if (hipFuncSetAttribute(gang_kernel<KB_mg_tail<__hip_bfloat16>, GB_WIDE, TailLevels, const MgState *>(),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTailLdsMax) != hipSuccess)
        return -1;
    return 0;
}

template <typename T>
void mg_launch_tail(hipStream_t s, const TailLevels &tl, size_t lds_bytes, const MgState *m)
{
    if (t_rec) {  // a gang member's tail runs as one workgroup
        TailLevels one = tl;
        one.groups = 1;
        FFR_GANG((KB_mg_tail<T>), GB_WIDE, k_mg_tail<T>, dim3(1), dim3(kTailThreads), lds_bytes, s,
                 one, m);
        return;
    }
    // a group: kXcds blocks per member, of which the one on XCD tl.xcd stays
    const unsigned grid = tl.groups > 1 && tl.first < tl.lds_from ? kXcds * (unsigned)tl.groups : 1u;
    FFR_LAUNCH(k_mg_tail<T>, dim3(grid), dim3(kTailThreads), lds_bytes, s, tl, m);
}

template <typename T>
void mg_launch_coarse(hipStream_t s, const Geo &g, const MgLevel &L, bool l0, int sweeps,
                      const MgState *m)
{
    if (l0 && g.uniform)
        FFR_GANG((KB_mg_coarse<T, true, false>), GB_WIDE, (k_mg_coarse<T, true, false>), dim3(1),
                 dim3(1024), 0, s, g, L, (T *)L.z, (const T *)L.r, sweeps, m);
    else if (l0)
        FFR_GANG((KB_mg_coarse<T, true, true>), GB_WIDE, (k_mg_coarse<T, true, true>), dim3(1),
                 dim3(1024), 0, s, g, L, (T *)L.z, (const T *)L.r, sweeps, m);
    else
        FFR_GANG((KB_mg_coarse<T, false, false>), GB_WIDE, (k_mg_coarse<T, false, false>), dim3(1),
                 dim3(1024), 0, s, g, L, (T *)L.z, (const T *)L.r, sweeps, m);
}

template void mg_launch_smooth<float>(hipStream_t, const Geo &, const MgLevel &, bool, int, int, const MgState *, const RedScratch *);
template void mg_launch_smooth<__hip_bfloat16>(hipStream_t, const Geo &, const MgLevel &, bool, int, int, const MgState *, const RedScratch *);
template void mg_launch_restrict<float, float>(hipStream_t, const Geo &, const MgLevel &, const MgLevel &, bool, bool, const MgState *, float *, int);
template void mg_launch_restrict<float, __hip_bfloat16>(hipStream_t, const Geo &, const MgLevel &, const MgLevel &, bool, bool, const MgState *, float *, int);
template void mg_launch_restrict<__hip_bfloat16, __hip_bfloat16>(hipStream_t, const Geo &, const MgLevel &, const MgLevel &, bool, bool, const MgState *, float *, int);
template void mg_launch_post_black<float, float>(hipStream_t, const Geo &, const MgLevel &, const MgLevel &, bool, const MgState *);
template void mg_launch_post_black<float, __hip_bfloat16>(hipStream_t, const Geo &, const MgLevel &, const MgLevel &, bool, const MgState *);
template void mg_launch_post_black<__hip_bfloat16, __hip_bfloat16>(hipStream_t, const Geo &, const MgLevel &, const MgLevel &, bool, const MgState *);
template void mg_launch_tail<float>(hipStream_t, const TailLevels &, size_t, const MgState *);
template void mg_launch_tail<__hip_bfloat16>(hipStream_t, const TailLevels &, size_t, const MgState *);
template void mg_launch_coarse<float>(hipStream_t, const Geo &, const MgLevel &, bool, int, const MgState *);
template void mg_launch_coarse<__hip_bfloat16>(hipStream_t, const Geo &, const MgLevel &, bool, int, const MgState *);

void pcg_launch_dirq(hipStream_t s, const Geo &g, const MgLevel &L0, const float *r, float *p,
                     float *q, MgState *m, RedScratch red, uint32_t gdot)
{
    if (g.uniform)
        FFR_GANG((KB_pcg_dirq<false>), GB_PLAIN, (k_pcg_dirq<false>), dim3(grid_resident(L0.H)),
                 dim3(kBlock), 0, s, g, L0, (const float *)L0.z, r, p, q, m, red, gdot);
    else
        FFR_GANG((KB_pcg_dirq<true>), GB_PLAIN, (k_pcg_dirq<true>), dim3(grid_resident(L0.H)),
                 dim3(kBlock), 0, s, g, L0, (const float *)L0.z, r, p, q, m, red, gdot);
}
void pcg_launch_dir(hipStream_t s, uint32_t n, const float *z, float *p, MgState *m, RedScratch red,
                    uint32_t gdot)
{
    FFR_GANG((KB_pcg_dir), GB_PLAIN, k_pcg_dir, dim3(grid_for(n)), dim3(kBlock), 0, s, n, z, p, m,
             red, gdot);
}
void pcg_launch_spmv(hipStream_t s, const Geo &g, const MgLevel &L0, const float *p, float *q,
                     const MgState *m, RedScratch red)
{
    if (g.uniform)
        FFR_GANG((KB_pcg_spmv<false>), GB_PLAIN, (k_pcg_spmv<false>), dim3(grid_for(g.nc)),
                 dim3(kBlock), 0, s, g, L0, p, q, m, red);
    else
        FFR_GANG((KB_pcg_spmv<true>), GB_PLAIN, (k_pcg_spmv<true>), dim3(grid_for(g.nc)),
                 dim3(kBlock), 0, s, g, L0, p, q, m, red);
}
void pcg_launch_update(hipStream_t s, const Geo &g, const MgLevel &L0, float *e, float *r,
                       const float *p, const float *q, MgState *m, RedScratch red, uint32_t gq)
{
    if (g.uniform)
        FFR_GANG((KB_pcg_update<false>), GB_PLAIN, (k_pcg_update<false>), dim3(grid_for(g.nc)),
                 dim3(kBlock), 0, s, g, L0, e, r, p, q, (float *)L0.z, m, red, gq);
    else
        FFR_GANG((KB_pcg_update<true>), GB_PLAIN, (k_pcg_update<true>), dim3(grid_for(g.nc)),
                 dim3(kBlock), 0, s, g, L0, e, r, p, q, (float *)L0.z, m, red, gq);
}
void pcg_launch_check(hipStream_t s, uint32_t n, MgState *m, RedScratch red)
{
    FFR_GANG((KB_pcg_check), GB_PLAIN, k_pcg_check, dim3(1), dim3(kBlock), 0, s, m, red, grid_for(n));
}
void pcg_launch_dot_rz(hipStream_t s, uint32_t n, const float *r, const float *z, const MgState *m,
                       RedScratch red)
{
    FFR_GANG((KB_pcg_dot_rz), GB_PLAIN, k_pcg_dot_rz, dim3(grid_for(n)), dim3(kBlock), 0, s, n, r,
             z, m, red);
}
void dc_launch_init(hipStream_t s, const Geo &g, const double *b, double *x, int warm,
                    const double *h1, double *out, RedScratch red)
{
    FFR_GANG((KB_dc_init), GB_PLAIN, k_dc_init, dim3(grid_for(g.nc)), dim3(kBlock), 0, s, g, b, x,
             warm, h1, out, red);
}
void dc_launch_start(hipStream_t s, const Geo &g, MgState *m, RedScratch red, double outer_tol,
                     int max_outer, int max_inner, double inner_tol, int total_cap)
{
    FFR_GANG((KB_dc_start), GB_PLAIN, k_dc_start, dim3(1), dim3(kBlock), 0, s, m, red,
             grid_for(g.nc), outer_tol, max_outer, max_inner, inner_tol, total_cap);
}
void dc_launch_residual(hipStream_t s, const Geo &g, const double *b, double *xa, double *xb,
                        const float *e, float *r32, float *z0, const MgState *m, RedScratch red,
                        bool first)
{
    const dim3 grid(grid_for(g.nc));
    if (g.uniform && first)
        FFR_GANG((KB_dc_residual<false, true>), GB_PLAIN, (k_dc_residual<false, true>), grid,
                 dim3(kBlock), 0, s, g, b, xa, xb, e, r32, z0, m, red);
    else if (g.uniform)
        FFR_GANG((KB_dc_residual<false, false>), GB_PLAIN, (k_dc_residual<false, false>), grid,
                 dim3(kBlock), 0, s, g, b, xa, xb, e, r32, z0, m, red);
    else if (first)
        FFR_GANG((KB_dc_residual<true, true>), GB_PLAIN, (k_dc_residual<true, true>), grid,
                 dim3(kBlock), 0, s, g, b, xa, xb, e, r32, z0, m, red);
    else
        FFR_GANG((KB_dc_residual<true, false>), GB_PLAIN, (k_dc_residual<true, false>), grid,
                 dim3(kBlock), 0, s, g, b, xa, xb, e, r32, z0, m, red);
}
void dc_launch_check(hipStream_t s, const Geo &g, MgState *m, RedScratch red)
{
    FFR_GANG((KB_dc_check), GB_PLAIN, k_dc_check, dim3(1), dim3(kBlock), 0, s, m, red,
             grid_for(g.nc));
}
void dc_launch_finalize(hipStream_t s, const Geo &g, const double *b, const double *xa,
                        const double *xb, double *p, int pin_outflow, bool inplace,
                        const MgState *m)
{
    if (g.uniform)
        FFR_GANG((KB_dc_finalize<false>), GB_PLAIN, (k_dc_finalize<false>), dim3(grid_for(g.nc)),
                 dim3(kBlock), 0, s, g, b, xa, xb, p, pin_outflow, inplace ? 1 : 0, m);
    else
        FFR_GANG((KB_dc_finalize<true>), GB_PLAIN, (k_dc_finalize<true>), dim3(grid_for(g.nc)),
                 dim3(kBlock), 0, s, g, b, xa, xb, p, pin_outflow, inplace ? 1 : 0, m);
}

}  // namespace ffr
