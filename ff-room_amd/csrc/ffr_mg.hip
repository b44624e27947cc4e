// ffr_mg.hip - host side of the multigrid-preconditioned pressure solve
// (method in ffr_mg.hpp): hierarchy set-up on the device and the enqueue
// logic of the outer FP64 / inner FP32 iteration.
#include "ffr_gang.hpp"
#include "ffr_scene.hpp"
#include "ffr_mg_launch.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ffr {

namespace {

struct Dims {
    int levels;
    int n[kMaxLevels][3];
    int sh[kMaxLevels][3];  // axes halved from level l-1 to l
};

// Level extents: coarsen the axes (with at least 4 cells) whose spacing is
// at most 1/sqrt(2) of the largest spacing (their coupling is at least twice
// the weakest), or all such axes when none is (near-cubic cells); stop when
// no axis has 4 cells, at the level cap, or once a level has at most
// kCoarsest cells (solved with a dense inverse in one workgroup).
Dims level_dims(int nx, int ny, int nz, const double h0[3], const ffr_params &p)
{
    Dims d{};
    const int cap = p.mg_levels > 0 ? std::min(p.mg_levels, kMaxLevels) : kMaxLevels;
    d.n[0][0] = nx; d.n[0][1] = ny; d.n[0][2] = nz;
    double h[3] = {h0[0], h0[1], h0[2]};
    d.levels = 1;
    while (d.levels < cap) {
        const int *f = d.n[d.levels - 1];
        if ((long)f[0] * f[1] * f[2] <= kCoarsest) break;
        bool can[3], any = false;
        for (int a = 0; a < 3; a++) {
            can[a] = f[a] >= 4;
            any |= can[a];
        }
        if (!any) break;
        const double hmax = std::max({h[0], h[1], h[2]});
        int sh[3];
        bool anys = false;
        for (int a = 0; a < 3; a++) {
            sh[a] = (can[a] && h[a] <= 0.70710678 * hmax) ? 1 : 0;
            anys |= sh[a] != 0;
        }
        if (!anys)
            for (int a = 0; a < 3; a++) sh[a] = can[a] ? 1 : 0;
        for (int a = 0; a < 3; a++) {
            d.sh[d.levels][a] = sh[a];
            d.n[d.levels][a] = sh[a] ? (f[a] + 1) / 2 : f[a];
            if (sh[a]) h[a] *= 2.0;
        }
        d.levels++;
    }
    return d;
}

// arena layout (sizing pass when base == nullptr)
size_t mg_layout(char *base, const Dims &D, bool bf16, bool dense, Multigrid *M, MgState **state)
{
    size_t off = 0;
    auto take = [&](size_t bytes) -> char * {
        char *ptr = base ? base + off : nullptr;
        off = (off + bytes + 255) & ~(size_t)255;
        return ptr;
    };
    const uint32_t n0 = (uint32_t)D.n[0][0] * D.n[0][1] * D.n[0][2];
    const uint32_t H = (n0 + 1) / 2;  // split layout: two colour arrays of H
    MgState *st = (MgState *)take(sizeof(MgState));
    float *e = (float *)take(8ull * H), *ri = (float *)take(8ull * H);
    float *pi = (float *)take(8ull * H), *qi = (float *)take(8ull * H);
    float *sc = (float *)take(8ull * H);
    float *z0 = (float *)take(8ull * H);
    uint16_t *nbs = (uint16_t *)take(4ull * H);
    const size_t na = (size_t)D.n[0][0] + D.n[0][1] + D.n[0][2];
    float4 *mf = (float4 *)take(16 * na);
    if (M) {
        M->e = e; M->ri = ri; M->pi = pi; M->qi = qi; M->sc = sc;
        M->mgf = mf;
        M->lv[0].z = z0;
        M->lv[0].r = ri;
        M->lv[0].nbs = nbs;
        M->lv[0].H = H;
        M->lv[0].active = nullptr;
        M->lv[0].cx = M->lv[0].cy = M->lv[0].cz = M->lv[0].diag = M->lv[0].rdiag = nullptr;
        M->lv[0].bf16 = 0;
    }
    {
        const int L = D.levels - 1;
        const uint32_t n = (uint32_t)D.n[L][0] * D.n[L][1] * D.n[L][2];
        const bool has = dense && L >= 1 && n <= (uint32_t)kCoarsest;
        float *gi = has ? (float *)take(4ull * n * n) : nullptr;
        if (M) M->ginv = gi;
    }
    for (int l = 1; l < D.levels; l++) {
        const uint32_t n = (uint32_t)D.n[l][0] * D.n[l][1] * D.n[l][2];
        const size_t vs = bf16 ? 2 : 4;
        uint8_t *act = (uint8_t *)take(n);
        float *cx = (float *)take(4ull * n), *cy = (float *)take(4ull * n), *cz = (float *)take(4ull * n);
        float *dg = (float *)take(4ull * n), *rd = (float *)take(4ull * n);
        float *gx = (float *)take(4ull * n), *gy = (float *)take(4ull * n), *gz = (float *)take(4ull * n);
        // z with si zero pads at both ends (the tail reads a cell's six
        // neighbours without bounds tests when the level runs from global memory)
        const size_t si = (size_t)D.n[l][1] * D.n[l][2];
        char *zp = take(vs * (n + 2 * si));
        void *z = zp ? zp + vs * si : nullptr, *r = take(vs * n);
        if (M) {
            MgLevel &L = M->lv[l];
            L.active = act; L.cx = cx; L.cy = cy; L.cz = cz; L.diag = dg; L.rdiag = rd;
            L.dgx = gx; L.dgy = gy; L.dgz = gz;
            L.z = z; L.r = r; L.bf16 = bf16 ? 1 : 0;
        }
    }
    if (state) *state = st;
    return off;
}

double mean(const double *h, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += h[i];
    return s / n;
}

// Level-0 factors of a stretched grid (see Multigrid::mgf), one axis: w[i] =
// d[i] / m and t(i) = 1 / (0.5 (d[i-1] + d[i]) m), the centre distance of
// PoissonSolver.cpp:34-49, of the face below i for 1 <= i < n (0 at both
// walls); entry i holds {w[i], t(i), t(i+1), 0}.
void stretch_factors(const double *d, int n, std::vector<float4> &f)
{
    const double m = mean(d, n);
    auto t = [&](int i) { return (i <= 0 || i >= n) ? 0.f : (float)(1.0 / (0.5 * (d[i - 1] + d[i]) * m)); };
    for (int i = 0; i < n; i++) f.push_back(make_float4((float)(d[i] / m), t(i), t(i + 1), 0.f));
}

}  // namespace

size_t mg_bytes(int nx, int ny, int nz, const double *dxv, const double *dyv, const double *dzv,
                const ffr_params &p)
{
    // level extents follow the mean spacing (the uniform spacing when uniform)
    const double h0[3] = {mean(dxv, nx), mean(dyv, ny), mean(dzv, nz)};
    const Dims D = level_dims(nx, ny, nz, h0, p);
    return mg_layout(nullptr, D, p.mg_bf16 != 0, p.mg_coarse_sweeps == 0, nullptr, nullptr);
}

// Dense generalised inverse of the coarsest level's operator A (n <= kCoarsest),
// computed once per scene in FP64 and stored in FP32: A = L D L^T without
// pivoting (A is symmetric positive semidefinite); a pivot at or below
// 1e-12 of the largest diagonal marks a direction A does not see (a pocket
// with no Dirichlet neighbour; for a semidefinite A its column below is 0
// too) and gets D^+ = 0. G = L^-T D^+ L^-1 is symmetric and semidefinite, so
// the V-cycle stays a symmetric preconditioner. Rows of cells with no
// unknown child are the identity.
static int coarse_inverse(ffr_scene *s)
{
    Multigrid &M = s->mg;
    const MgLevel &C = M.lv[M.levels - 1];
    const int n = (int)C.n, si = C.ny * C.nz, sj = C.nz;
    std::vector<float> cx(n), cy(n), cz(n), dg(n);
    std::vector<uint8_t> act(n);
    hipStream_t st = s->stream;
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(cx.data(), C.cx, 4ull * n, hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(cy.data(), C.cy, 4ull * n, hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(cz.data(), C.cz, 4ull * n, hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(dg.data(), C.diag, 4ull * n, hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(act.data(), C.active, n, hipMemcpyDeviceToHost, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(st), FFR_E_DEVICE);
    // A (FP64): diag, and -coupling to each neighbour (couplings stored on
    // the upper cell of each pair)
    std::vector<double> A((size_t)n * n, 0.0);
    double dmax = 0.0;
    for (int c = 0; c < n; c++) {
        if (!act[c]) {
            A[(size_t)c * n + c] = 1.0;
            continue;
        }
        A[(size_t)c * n + c] = dg[c];
        dmax = std::max(dmax, (double)dg[c]);
        const int i = c / si, j = (c / sj) % C.ny, k = c % sj;
        const int lo[3] = {i > 0 ? c - si : -1, j > 0 ? c - sj : -1, k > 0 ? c - 1 : -1};
        const float cp[3] = {cx[c], cy[c], cz[c]};
        for (int a = 0; a < 3; a++)
            if (lo[a] >= 0 && cp[a] != 0.f) {
                A[(size_t)c * n + lo[a]] = -(double)cp[a];
                A[(size_t)lo[a] * n + c] = -(double)cp[a];
            }
    }
    // L D L^T in place: strict lower triangle of A holds L, d the pivots
    std::vector<double> d(n, 0.0), w(n);
    const double tol = 1e-12 * std::max(dmax, 1e-300);
    for (int k = 0; k < n; k++) {
        double dk = A[(size_t)k * n + k];
        for (int p = 0; p < k; p++) dk -= A[(size_t)k * n + p] * A[(size_t)k * n + p] * d[p];
        for (int p = 0; p < k; p++) w[p] = A[(size_t)k * n + p] * d[p];
        if (dk <= tol) {
            d[k] = 0.0;
            for (int i = k + 1; i < n; i++) A[(size_t)i * n + k] = 0.0;
            continue;
        }
        d[k] = dk;
        for (int i = k + 1; i < n; i++) {
            double v = A[(size_t)i * n + k];
            const double *Li = &A[(size_t)i * n];
            for (int p = 0; p < k; p++) v -= Li[p] * w[p];
            A[(size_t)i * n + k] = v / dk;
        }
    }
    // X = L^-1 (unit lower triangular) by rows: X_i = e_i - sum_{p<i} L_ip X_p;
    // then G = X^T D^+ X as rank-one updates by the rows of X. Every inner
    // loop runs along a row (n = 512: ~35 ms with the factorisation).
    std::vector<double> X((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) {
        double *Xi = &X[(size_t)i * n];
        Xi[i] = 1.0;
        const double *Li = &A[(size_t)i * n];
        for (int p = 0; p < i; p++) {
            const double l = Li[p];
            if (l == 0.0) continue;
            const double *Xp = &X[(size_t)p * n];
            for (int c = 0; c <= p; c++) Xi[c] -= l * Xp[c];
        }
    }
    std::vector<double> Gd((size_t)n * n, 0.0);
    for (int k = 0; k < n; k++) {
        if (!(d[k] > 0.0)) continue;
        const double *Xk = &X[(size_t)k * n];
        const double sk = 1.0 / d[k];
        for (int i = 0; i <= k; i++) {
            const double a = Xk[i] * sk;
            if (a == 0.0) continue;
            double *Gi = &Gd[(size_t)i * n];
            for (int j = 0; j <= i; j++) Gi[j] += a * Xk[j];
        }
    }
    std::vector<float> G((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j <= i; j++)
            G[(size_t)i * n + j] = G[(size_t)j * n + i] = (float)Gd[(size_t)i * n + j];
    HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(M.ginv, G.data(), 4ull * n * n, hipMemcpyHostToDevice, st), FFR_E_DEVICE);
    HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(st), FFR_E_DEVICE);
    return FFR_OK;
}

static MgState *g_unused_state;  // placeholder for sizing calls

int mg_create(ffr_scene *s)
{
    if (s->mg.levels) return FFR_OK;
    if (mg_tail_allow_lds() != 0) {
        s->ctx->fail("hipFuncSetAttribute: %zu bytes of LDS for the multigrid tail refused",
                     (size_t)kTailLdsMax);
        return FFR_E_DEVICE;
    }
    const double h0[3] = {mean(s->dxv.data(), s->nx), mean(s->dyv.data(), s->ny),
                          mean(s->dzv.data(), s->nz)};
    const Dims D = level_dims(s->nx, s->ny, s->nz, h0, s->prm);
    Multigrid &M = s->mg;
    const bool bf16 = s->prm.mg_bf16 != 0;
    const bool dense = s->prm.mg_coarse_sweeps == 0;
    const size_t bytes = mg_layout(nullptr, D, bf16, dense, nullptr, nullptr);
    if (s->ctx->in_use + bytes > s->ctx->budget) {
        s->ctx->fail("multigrid hierarchy needs %zu bytes beyond the context budget", bytes);
        return FFR_E_NOMEM;
    }
    M.arena = s->ctx->dev_take(bytes, &M.bytes);
    if (!M.arena) {
        s->ctx->fail("hipMalloc of %zu bytes for the multigrid hierarchy failed", bytes);
        return FFR_E_NOMEM;
    }
    MgState *st = nullptr;
    mg_layout((char *)M.arena, D, bf16, dense, &M, &st);
    s->mgs = st;
    M.levels = D.levels;
    for (int l = 0; l < D.levels; l++) {
        MgLevel &L = M.lv[l];
        L.nx = D.n[l][0]; L.ny = D.n[l][1]; L.nz = D.n[l][2];
        L.n = (uint32_t)L.nx * L.ny * L.nz;
        L.shx = D.sh[l][0]; L.shy = D.sh[l][1]; L.shz = D.sh[l][2];
        L.f_nz = make_fastdiv(L.nz);
        L.f_nynz = make_fastdiv(L.ny * L.nz);
    }
    for (int l = 0; l < D.levels; l++) {  // colour-slot parity (level 0: split layout; tail)
        MgLevel &L = M.lv[l];
        L.par_nz = (L.nz % 2) == 0;
        L.par_ny = (L.ny % 2) == 0;
        L.f_hz = make_fastdiv(L.par_nz ? L.nz / 2 : 1);
        L.f_hyz = make_fastdiv(L.par_ny ? L.ny * L.nz / 2 : 1);
    }
    // the small levels that fit one workgroup's LDS (operator, z and r of
    // each) run as a single kernel, and so does the next finer level when its
    // z and r fit too and it is small (its operator is then read from global
    // memory); level 0 never does (it is FP32 in the split layout)
    {
        const size_t vs = bf16 ? 2 : 4;
        const size_t budget = s->prm.mg_tail_kb > 0
                                  ? std::min<size_t>((size_t)s->prm.mg_tail_kb * 1024, kTailLdsMax)
                                  : kTailLdsDefault;
        M.tail_first = M.tail_staged = -1;
        M.tail_lds = 0;
        // with more than one level, the restriction adds an FP32 scratch of
        // the first level's size
        auto scratch = [&](int l) { return l < D.levels - 1 ? tail_scratch_bytes(M.lv[l].n) : 0; };
        size_t acc = 0;
        for (int l = D.levels - 1; l >= 1; l--) {
            const size_t b = tail_level_bytes(M.lv[l], vs, true);
            if (acc + b + scratch(l) > budget) break;
            acc += b;
            M.tail_first = M.tail_staged = l;
        }
        const int u = M.tail_first - 1;
        if (M.tail_first > 1 && M.lv[u].n <= kTailUnstagedMax &&
            acc + tail_level_bytes(M.lv[u], vs, false) + scratch(u) <= budget) {
            acc += tail_level_bytes(M.lv[u], vs, false);
            M.tail_first = u;
        }
        M.tail_lds = M.tail_first >= 0 ? acc + scratch(M.tail_first) : 0;
        // the finer levels of at most mg_tail_cells cells join the tail with
        // their vectors and operator left in global memory
        M.tail_lds_from = M.tail_first;
        if (s->prm.mg_tail_cells > 0 && M.tail_first > 1)
            while (M.tail_first > 1 && M.lv[M.tail_first - 1].n <= (uint32_t)s->prm.mg_tail_cells)
                M.tail_first--;
        // workgroups sharing the global levels (mg_tail_groups; 0: one per
        // kTailGroupCells cells of the largest), on an XCD of their own per scene
        M.tail_groups = 1;
        if (M.tail_first >= 1 && M.tail_first < M.tail_lds_from) {
            int grp = s->prm.mg_tail_groups;
            if (grp <= 0)
                grp = (int)((M.lv[M.tail_first].n + kTailGroupCells - 1) / kTailGroupCells);
            M.tail_groups = std::max(1, std::min(grp, kTailGroupsMax));
        }
        M.tail_xcd = (int)(s->ctx->tail_seq++ % (unsigned)kXcds);
    }
    M.pre_sweeps = std::max(1, s->prm.mg_pre_sweeps);
    M.level_sweeps = std::max(1, s->prm.mg_level_sweeps);
    M.coarse_sweeps = s->prm.mg_coarse_sweeps > 0 ? s->prm.mg_coarse_sweeps : 8;
    hipStream_t str = s->stream;
    HIP_CHECK_CTX(s->ctx, hipMemsetAsync(M.arena, 0, bytes, str), FFR_E_DEVICE);
    s->geo.mgf = M.mgf;
    std::vector<float4> ff;  // (stay zero on a uniform grid, where no kernel reads them)
    if (!s->geo.uniform) {
        stretch_factors(s->dxv.data(), s->nx, ff);
        stretch_factors(s->dyv.data(), s->ny, ff);
        stretch_factors(s->dzv.data(), s->nz, ff);
        HIP_CHECK_CTX(s->ctx, hipMemcpyAsync(M.mgf, ff.data(), sizeof(float4) * ff.size(),
                                             hipMemcpyHostToDevice, str), FFR_E_DEVICE);
    }
    mg_launch_split_codes(str, s->geo, M.lv[0]);
    for (int l = 0; l + 1 < D.levels; l++) mg_launch_build(str, s->geo, M.lv[l], M.lv[l + 1], l == 0);
    HIP_CHECK_CTX(s->ctx, hipStreamSynchronize(str), FFR_E_DEVICE);
    if (M.ginv) {
        const int rc = coarse_inverse(s);
        if (rc != FFR_OK) return rc;
    }
    (void)g_unused_state;
    return FFR_OK;
}

void mg_destroy(ffr_scene *s)
{
    if (s->mg.g_iter) hipGraphExecDestroy(s->mg.g_iter);
    if (s->mg.g_start) hipGraphExecDestroy(s->mg.g_start);
    s->mg.g_iter = s->mg.g_start = nullptr;
    s->mg.graphs_tried = false;
    if (s->mg.arena) {  // the scene's stream is idle: the block goes back to the cache
        s->ctx->dev_give(s->mg.arena, s->mg.bytes);
        s->mg.arena = nullptr;
    }
    s->mg.levels = 0;
}

// ---- V-cycle enqueue (z_0 = M r_0 on level 0) --------------------------------
// kernel names per level for the profiler (bytes differ per level)
static const char *kSmoothName[kMaxLevels] = {
    "mg_smooth_l0", "mg_smooth_l1", "mg_smooth_l2", "mg_smooth_l3", "mg_smooth_l4",
    "mg_smooth_l5", "mg_smooth_l6", "mg_smooth_l7", "mg_smooth_l8", "mg_smooth_l9",
    "mg_smooth_l10", "mg_smooth_l11"};
static const char *kRestrictName[kMaxLevels] = {
    "mg_restrict_l0", "mg_restrict_l1", "mg_restrict_l2", "mg_restrict_l3", "mg_restrict_l4",
    "mg_restrict_l5", "mg_restrict_l6", "mg_restrict_l7", "mg_restrict_l8", "mg_restrict_l9",
    "mg_restrict_l10", "mg_restrict_l11"};
static const char *kResidName[kMaxLevels] = {
    "mg_resid_l0", "mg_resid_l1", "mg_resid_l2", "mg_resid_l3", "mg_resid_l4", "mg_resid_l5",
    "mg_resid_l6", "mg_resid_l7", "mg_resid_l8", "mg_resid_l9", "mg_resid_l10", "mg_resid_l11"};
static const char *kPostName[kMaxLevels] = {
    "mg_post_l0", "mg_post_l1", "mg_post_l2", "mg_post_l3", "mg_post_l4", "mg_post_l5",
    "mg_post_l6", "mg_post_l7", "mg_post_l8", "mg_post_l9", "mg_post_l10", "mg_post_l11"};

// V-cycle from level l. first_red_done: the level's first pass (the red
// sweep from z = 0) was done by the kernel that produced its r (the PCG
// update or the FP64 residual on level 0, the restriction on the coarse
// levels). The prolongation is fused into the first post-smoothing sweep.
template <typename TC>
// Returns whether the last pass wrote the r.z partials (level 0 with a
// smoothed level 0: the last red pass does, when `dot`).
static bool vcycle(ffr_scene *s, int l, bool first_red_done, bool dot = false)
{
    Multigrid &M = s->mg;
    const Geo &g = s->geo;
    hipStream_t st = s->stream;
    const MgState *m = s->mgs;
    const MgLevel &L = M.lv[l];
    const bool l0 = l == 0;
    const double n = (double)L.n;
    if (l == M.tail_first) {
        TailLevels tl{};
        uint32_t off = 0;
        for (int k = l; k < M.levels; k++) {
            tl.lv[k] = M.lv[k];
            tl.off[k] = off;
            if (k >= M.tail_lds_from)  // levels before it run from global memory
                off += (uint32_t)tail_level_bytes(M.lv[k], sizeof(TC), k >= M.tail_staged);
        }
        tl.scratch_off = off;  // FP32 restriction scratch of the first LDS level's size
        tl.gscratch = M.sc;    // level 0's restriction scratch, free while the tail runs
        tl.lds_from = M.tail_lds_from;
        tl.first = l;
        tl.last = M.levels - 1;
        tl.staged_from = M.tail_staged;
        tl.sweeps = M.level_sweeps;
        tl.coarse_sweeps = M.coarse_sweeps;
        tl.ginv = M.ginv;
        MgState *ms = s->mgs;
        tl.groups = M.tail_groups;
        tl.xcd = M.tail_xcd;
        tl.gsync = ms->gsync;
        tl.fault = &ms->tail_fault;
        ProfScope ps(s->ctx, st, "mg_tail", 0);
        mg_launch_tail<TC>(st, tl, M.tail_lds, m);
        return false;
    }
    if (l == M.levels - 1) {
        ProfScope ps(s->ctx, st, l0 ? "mg_coarse_l0" : "mg_coarse", 0);
        if (l0) mg_launch_coarse<float>(st, g, L, true, M.coarse_sweeps, m);
        else mg_launch_coarse<TC>(st, g, L, false, M.coarse_sweeps, m);
        return false;
    }
    const MgLevel &C = M.lv[l + 1];
    // algorithmic bytes of one colour pass: z and r read, the code (level 0,
    // 2 B) or operator (coarse: active + 3 couplings + rdiag), z written
    // (level 0, split layout: r, code and z of the colour, z of the other colour)
    const double bsm = l0 ? n * (4 + 4 + 2 + 4) / 2 : n * (2 * sizeof(TC) + 1 + 16 + sizeof(TC));
    // fused restriction (small coarse levels): the fine level's r, z and
    // operator read, the coarse r written; two passes: the residual (level 0:
    // red cells' r, code and z, the black z, the residual written; coarse:
    // r, z, operator, residual written), then P^T of it
    const double bres = n * (2 * sizeof(TC) + 1 + 16) + C.n * (double)sizeof(TC);
    const double bresid = l0 ? n / 2 * (4 + 2 + 4 + 4) + n / 2 * 4 : n * (2 * sizeof(TC) + 1 + 16 + 4);
    const double bsum = (l0 ? n / 2 * 4 : n * 4) + C.n * (double)sizeof(TC);
    const double bpost = bsm + C.n * (double)sizeof(TC);
    const int sweeps = l0 ? M.pre_sweeps : M.level_sweeps;
    const RedScratch red = s->red;
    auto smooth = [&](int color, int zero_start, bool with_dot) {
        ProfScope ps(s->ctx, st, kSmoothName[l], bsm + (with_dot ? n * 4 : 0.0));
        if (l0) mg_launch_smooth<float>(st, g, L, true, color, zero_start, m, with_dot ? &red : nullptr);
        else mg_launch_smooth<TC>(st, g, L, false, color, zero_start, m);
    };
    for (int sw = 0; sw < sweeps; sw++) {
        if (!(sw == 0 && first_red_done)) smooth(0, sw == 0, false);
        smooth(1, 0, false);
    }
    // the next level's first sweep rides on the restriction unless that level
    // is the single-workgroup tail or the coarsest (both start from z = 0)
    const bool red0 = l + 1 != M.tail_first && l + 1 != M.levels - 1;
    auto restrict_pass = [&](int pass) {
        if (l0) mg_launch_restrict<float, TC>(st, g, L, C, true, red0, m, M.sc, pass);
        else mg_launch_restrict<TC, TC>(st, g, L, C, false, red0, m, M.sc, pass);
    };
    if (restrict_two_pass(l0, L.n)) {  // each pass in its own profiler scope
        {
            ProfScope ps(s->ctx, st, kResidName[l], bresid);
            restrict_pass(1);
        }
        ProfScope ps(s->ctx, st, kRestrictName[l], bsum);
        restrict_pass(2);
    } else {
        ProfScope ps(s->ctx, st, kRestrictName[l], bres);
        restrict_pass(0);
    }
    vcycle<TC>(s, l + 1, red0);
    {
        ProfScope ps(s->ctx, st, kPostName[l], bpost);
        if (l0) mg_launch_post_black<float, TC>(st, g, L, C, true, m);
        else mg_launch_post_black<TC, TC>(st, g, L, C, false, m);
    }
    // post-smoothing: the red half of the first sweep, then whole sweeps
    // (at least one sweep, so the V-cycle ends on a red pass)
    const int post = std::max(sweeps, 1);
    for (int sw = 0; sw < post; sw++) {
        if (sw > 0) smooth(1, 0, false);
        smooth(0, 0, l0 && dot && sw == post - 1);
    }
    return l0 && dot;
}

// Whether the V-cycle ends with a red pass on level 0 (level 0 smoothed by
// launches: neither the coarsest level nor in the tail). That pass then
// writes the partials of r.z, and the PCG takes its main form (k_pcg_dirq).
static bool main_form(const Multigrid &M) { return M.levels > 1 && M.tail_first != 0; }

// z = M r (level 0's first red sweep is done by the PCG update / FP64
// residual), then the partials of r.z
static void precondition(ffr_scene *s)
{
    const bool fused = s->mg.levels > 1 && s->mg.lv[1].bf16
                           ? vcycle<__hip_bfloat16>(s, 0, true, true)
                           : vcycle<float>(s, 0, true, true);
    if (!fused) {
        ProfScope ps(s->ctx, s->stream, "pcg_dot_rz", (double)s->geo.nc * 8);
        pcg_launch_dot_rz(s->stream, s->geo.nc, s->mg.ri, (const float *)s->mg.lv[0].z, s->mgs,
                          s->red);
    }
}

// the number of r.z partials precondition() leaves: one per block of the
// last red pass (over the H red slots, a resident grid) or of k_pcg_dot_rz
// (all slots)
static uint32_t gdot_of(const ffr_scene *s)
{
    return main_form(s->mg) ? grid_resident(s->mg.lv[0].H) : grid_for(s->geo.nc);
}

static void inner_iteration(ffr_scene *s)
{
    Multigrid &M = s->mg;
    const Geo &g = s->geo;
    hipStream_t st = s->stream;
    const double n = (double)g.nc;
    uint32_t gq;
    if (main_form(M)) {
        // z, r (red half), code, p and q read; p and q written
        ProfScope ps(s->ctx, st, "pcg_dirq", n * (4 + 2 + 2 + 4 + 4 + 4 + 4));
        pcg_launch_dirq(st, g, M.lv[0], M.ri, M.pi, M.qi, s->mgs, s->red, gdot_of(s));
        gq = grid_resident(M.lv[0].H);
    } else {
        pcg_launch_dir(st, g.nc, (const float *)M.lv[0].z, M.pi, s->mgs, s->red, gdot_of(s));
        pcg_launch_spmv(st, g, M.lv[0], M.pi, M.qi, s->mgs, s->red);
        gq = grid_for(g.nc);
    }
    {
        ProfScope ps(s->ctx, st, "pcg_update", n * 24);
        pcg_launch_update(st, g, M.lv[0], M.e, M.ri, M.pi, M.qi, s->mgs, s->red, gq);
    }
    pcg_launch_check(st, g.nc, s->mgs, s->red);
    precondition(s);
}

// ---- the solve as a task (see ffr_task.hip) ------------------------------------
// Outer pass k: r = b - A x in FP64 and the stopping test; between passes k
// and k+1 the inner solve k (z = M r, p = z, k_i CG steps). A solve enqueues
// the passes and inner iterations the previous solve of the scene needed and
// then waits once; an inner solve with no history is grown under host control
// (a few iterations, then two at a time until the device reports it done).

// an inner solve starts with z = M r and r.z; its first iteration takes
// p = z and e = 0 (k_pcg_dirq or k_pcg_dir, and k_pcg_update)
static void inner_start(ffr_scene *s) { precondition(s); }

// ---- graph replay ------------------------------------------------------------
// Every launch of an inner iteration (and of an inner solve's start) has the
// same arguments each time: iteration state lives on the device. Each is
// captured once per scene as a graph and replayed, which replaces ~45 kernel
// launches by one graph launch. No memset or host work is captured. The
// profiler needs per-launch events, so profiled runs launch eagerly.
static hipGraphExec_t capture(ffr_scene *s, void (*body)(ffr_scene *))
{
    hipStream_t st = s->stream;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    body(s);
    if (hipStreamEndCapture(st, &graph) != hipSuccess || !graph) return nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) exec = nullptr;
    hipGraphDestroy(graph);
    return exec;
}

static bool graphs_ready(ffr_scene *s)
{
    Multigrid &M = s->mg;
    if (!s->prm.graphs || s->ctx->prof_on || t_rec) return false;
    if (!M.graphs_tried) {
        M.graphs_tried = true;
        M.g_iter = capture(s, inner_iteration);
        M.g_start = capture(s, inner_start);
        hipGetLastError();  // a failed capture leaves the eager path
    }
    return M.g_iter && M.g_start;
}

// A replay the runtime refused: the scene drops its graphs and launches
// eagerly from then on. The iteration state is on the device, so an
// iteration is either replayed or launched, never half of each.
static void drop_graphs(ffr_scene *s)
{
    Multigrid &M = s->mg;
    if (M.g_iter) hipGraphExecDestroy(M.g_iter);
    if (M.g_start) hipGraphExecDestroy(M.g_start);
    M.g_iter = M.g_start = nullptr;
    hipGetLastError();
}

static void launch_iterations(ffr_scene *s, int k)
{
    if (k > 0 && graphs_ready(s)) {
        int i = 0;
        while (i < k && hipGraphLaunch(s->mg.g_iter, s->stream) == hipSuccess) i++;
        if (i == k) return;
        drop_graphs(s);
        k -= i;
    }
    if (k > 0 && t_rec) {  // recorded once, repeated with the same arguments
        const size_t b = t_rec->ops.size();
        inner_iteration(s);
        t_rec->repeat(b, t_rec->ops.size(), k - 1);
        return;
    }
    for (int i = 0; i < k; i++) inner_iteration(s);
}

// inner solve: z = M r, p = z, then k CG steps
static void enqueue_inner(ffr_scene *s, int k)
{
    const bool replay = graphs_ready(s);
    if (!replay || hipGraphLaunch(s->mg.g_start, s->stream) != hipSuccess) {
        if (replay) drop_graphs(s);
        inner_start(s);
    }
    launch_iterations(s, k);
}

// outer pass: x += e, r = b - A x (FP64), r32 = r; stopping test
static void enqueue_pass(ffr_scene *s)
{
    hipStream_t st = s->stream;
    const Geo &g = s->geo;
    const double n = (double)g.nc;
    {
        // the solve's first pass: no correction to add yet (x read, not written)
        const bool first = s->task.passes_enq == 0;
        ProfScope ps(s->ctx, st, "dc_residual",
                     first ? n * (8 + 8 + 1 + 4 + 2) : n * (8 + 8 + 4 + 1 + 8 + 4 + 2));
        dc_launch_residual(st, g, s->b, s->p, s->d, s->mg.e, s->mg.ri, (float *)s->mg.lv[0].z,
                           s->mgs, s->red, first);
    }
    dc_launch_check(st, g, s->mgs, s->red);
    s->task.passes_enq++;
}

static int max_inner_of(const ffr_scene *s) { return std::max(1, s->prm.inner_max_iter); }

// enqueues passes up to task.target_passes, stopping early at an inner solve
// with no history; ends with a wait point
static int mg_schedule(ffr_scene *s)
{
    StepTask &t = s->task;
    const int max_inner = max_inner_of(s);
    while (t.passes_enq < t.target_passes) {
        const int idx = t.passes_enq - 1;  // the inner solve after residual idx
        int k = idx < kMgHist ? s->pred.mg_inner[idx] : 0;
        if (k <= 0) {
            k = std::min(max_inner, 6);
            enqueue_inner(s, k);
            if (idx < kMgHist) t.k_used[idx] = 0;
            t.inner_wait = idx;
            return wait_point(s, s->h_mgs, s->mgs, sizeof(MgState));
        }
        k = std::min(k, max_inner);
        enqueue_inner(s, k);
        if (idx < kMgHist) t.k_used[idx] = k;
        enqueue_pass(s);
    }
    return wait_point(s, s->h_mgs, s->mgs, sizeof(MgState));
}

int mg_begin(ffr_scene *s)
{
    if (!s->mg.levels) {
        int rc = mg_create(s);
        if (rc != FFR_OK) return rc;
    }
    Multigrid &M = s->mg;
    const Geo &g = s->geo;
    hipStream_t st = s->stream;
    StepTask &t = s->task;
    const double n = (double)g.nc;
    {
        // a step starts from the pressure of the previous step (warm_start);
        // a bare solve, like PoissonSolver::solve, from 0
        // (warm_start 2 and 3: extrapolated linearly or quadratically from
        // the pressures of the last two or three steps, as far as the scene
        // has them)
        // With a history, the start is written to the spare buffer (p_hist2,
        // which holds p_{n-2} or nothing yet) and the three buffers rotate:
        // the spare becomes the solve's x, x's buffer (p_n) p_{n-1}, and
        // p_{n-1}'s buffer the next spare.
        const int ws = std::max(0, std::min(3, s->prm.warm_start));
        const bool warm = ws != 0 && !t.solve_only;
        const bool keep = warm && ws >= 2 && s->p_hist;
        const int mode = warm ? std::min(ws, 1 + (keep ? s->p_hist_n : 0)) : 0;
        ProfScope ps(s->ctx, st, "dc_init",
                     n * (8 + 8 + 1) + (keep ? n * 8 * (mode == 1 ? 1 : mode) : 0));
        dc_launch_init(st, g, s->b, s->p, mode, keep ? s->p_hist : nullptr,
                       keep ? s->p_hist2 : nullptr, s->red);
        if (keep) {
            double *const x = s->p_hist2;
            s->p_hist2 = s->p_hist;
            s->p_hist = s->p;
            s->p = x;
        }
        s->p_hist_n = keep ? std::min(2, s->p_hist_n + 1) : 0;
    }
    dc_launch_start(st, g, s->mgs, s->red, s->prm.poisson_tol, /*max_outer=*/50, max_inner_of(s),
                    s->prm.inner_tol, std::max(1, s->prm.poisson_max_iter));
    t.passes_enq = 0;
    t.inner_wait = -1;
    t.target_passes = std::max(1, s->pred.mg_passes);
    enqueue_pass(s);
    return mg_schedule(s);
}

int mg_continue(ffr_scene *s, bool *finished)
{
    StepTask &t = s->task;
    const MgState &h = *s->h_mgs;
    *finished = false;
    if (h.tail_fault) {
        // a tail group's barrier timed out: the solve is void; the scene's
        // tails run as one workgroup from now on (its graphs hold the group)
        drop_graphs(s);
        s->mg.tail_groups = 1;
        HIP_CHECK_CTX(s->ctx, hipMemsetAsync(s->mgs->gsync, 0, 2 * sizeof(unsigned) + sizeof(int),
                                             s->stream), FFR_E_DEVICE);
        s->ctx->fail("a multigrid tail group barrier timed out; the scene's tails now run as one "
                     "workgroup");
        return FFR_E_DEVICE;
    }
    if (t.inner_wait >= 0) {  // an inner solve grown under host control
        if (!h.solve_done && !h.inner_done && h.inner_it < max_inner_of(s)) {
            launch_iterations(s, 2);
            return wait_point(s, s->h_mgs, s->mgs, sizeof(MgState));
        }
        t.inner_wait = -1;
        enqueue_pass(s);
        return mg_schedule(s);
    }
    if (!h.solve_done) {  // the solve needs more passes than enqueued
        t.target_passes = t.passes_enq + 1;
        return mg_schedule(s);
    }
    {
        ProfScope ps(s->ctx, s->stream, "dc_finalize", (double)s->geo.nc * 2);  // codes; OUTFLOW rows
        // the last pass wrote its x into s->d after an even, non-zero number
        // of passes: the two buffers swap roles, so that x is p in place
        if (h.passes > 0 && !(h.passes & 1)) std::swap(s->p, s->d);
        dc_launch_finalize(s->stream, s->geo, s->b, s->p, s->d, s->p, s->prm.use_multigrid, true,
                           s->mgs);
    }
    t.iters = h.total_iters;
    t.outer = h.passes;
    t.res = h.res;
    // predictions for the next solve: an inner solve that used every enqueued
    // iteration may have wanted more, so it gets one more next time
    SolvePred &P = s->pred;
    P.mg_passes = h.passes;
    for (int k = 0; k < kMgHist; k++) {
        if (k + 1 >= h.passes) {
            P.mg_inner[k] = 0;
            continue;
        }
        const int used = h.hist[k];
        const bool saturated = t.k_used[k] > 0 && used >= t.k_used[k];
        P.mg_inner[k] = std::min(max_inner_of(s), used + (saturated ? 1 : 0));
    }
    *finished = true;
    return FFR_OK;
}

double mg_prof_launch(ffr_scene *s, const char *name, std::function<void()> *launch)
{
    Multigrid &M = s->mg;
    const Geo g = s->geo;
    hipStream_t st = s->stream;
    MgState *m = s->mgs;
    const double n = (double)g.nc;
    const std::string k(name);
    const MgLevel L0 = M.lv[0];
    if (k == "mg_smooth_l0") {
        *launch = [=] { mg_launch_smooth<float>(st, g, L0, true, 1, 0, m); };
        return (double)(g.nc - L0.H) * (4 + 4 + 2 + 4);
    }
    if (k == "mg_resid0" && M.levels > 1) {  // the first pass of the level-0 restriction
        // red cells: r, code and z read, the residual written; black z read
        float *res = M.sc;
        *launch = [=] { mg_launch_resid0(st, g, L0, res, m); };
        return n / 2 * (4 + 2 + 4 + 4) + n / 2 * 4;
    }
    if (k == "pcg_dirq" && main_form(M)) {
        float *r = M.ri, *p = M.pi, *q = M.qi;
        RedScratch red = s->red;
        const uint32_t gd = gdot_of(s);
        *launch = [=] { pcg_launch_dirq(st, g, L0, r, p, q, m, red, gd); };
        return n * (4 + 2 + 2 + 4 + 4 + 4 + 4);
    }
    if (k == "pcg_update") {
        float *e = M.e, *r = M.ri, *p = M.pi, *q = M.qi;
        RedScratch red = s->red;
        const uint32_t gq = main_form(M) ? grid_resident(L0.H) : grid_for(g.nc);
        *launch = [=] { pcg_launch_update(st, g, L0, e, r, p, q, m, red, gq); };
        return n * 24;
    }
    if (k == "dc_residual") {
        double *b = s->b, *x0 = s->p, *x1 = s->d;
        float *e = M.e, *r = M.ri;
        RedScratch red = s->red;
        float *z0 = (float *)L0.z;
        *launch = [=] { dc_launch_residual(st, g, b, x0, x1, e, r, z0, m, red, false); };
        return n * (8 + 8 + 4 + 2 + 8 + 4);
    }
    *launch = [] {};
    return 0.0;
}

}  // namespace ffr

extern "C" int ffr_scene_mg_layout(const ffr_scene *s, int *levels, int *tail_first,
                                   int *tail_lds_from, unsigned int *cells)
{
    if (!s || !levels) return FFR_E_INVALID;
    const ffr::Multigrid &M = s->mg;
    *levels = M.levels;
    if (tail_first) *tail_first = M.levels ? M.tail_first : -1;
    if (tail_lds_from) *tail_lds_from = M.levels ? M.tail_lds_from : -1;
    if (cells)
        for (int l = 0; l < M.levels; l++) cells[l] = M.lv[l].n;
    return FFR_OK;
}

extern "C" int ffr_scene_mg_tail_groups(const ffr_scene *s)
{
    if (!s) return FFR_E_INVALID;
    return s->mg.levels ? s->mg.tail_groups : 0;
}
