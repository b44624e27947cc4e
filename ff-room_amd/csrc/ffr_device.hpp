// ffr_device.hpp - device-side types and helpers shared by the gfx950 kernels.
//
// Data layout in HBM: every field keeps the reference's flat layout
// (k fastest; u (nx+1,ny,nz), v (nx,ny+1,nz), w (nx,ny,nz+1), cells
// (nx,ny,nz)), so a thread's flat index is its element and a wavefront reads
// 64 consecutive elements. Neighbours at +-1 share the cache line, +-nz and
// +-ny*nz come from L2; launch blocks are remapped so that each XCD (own L2)
// walks one contiguous eighth of an array (see xcd_block()).
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace ffr {

constexpr int kBlock = 256;  // threads per block (4 wave64)
constexpr int kWave = 64;
constexpr int kXcds = 8;     // MI355X: 8 XCDs, blocks dispatched round-robin

enum : uint8_t { FLUID = 0, SOLID = 1, INFLOW = 2, OUTFLOW = 3 };
enum { UPWIND1 = 0, QUICK = 1, LAX_WENDROFF = 2 };

// Division by a runtime-invariant divisor d (1 <= d < 2^31) for n < 2^31:
// q = (umulhi(n, m) + n) >> s, with s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1.
struct FastDiv {
    uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    f.d = d;
    uint32_t s = 0;
    while ((1ull << s) < d) s++;
    f.s = s;
    f.m = (uint32_t)((((1ull << s) - d) << 32) / d + 1);
    return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f)
{
    return (__umulhi(n, f.m) + n) >> f.s;
}

struct FanDev {      // FanBC with its box clamped to the grid (BoundaryManager.cpp:16-22)
    int i0, i1, j0, j1, k0, k1;
    double vel[3];
    float velf[3];
};

struct OpeningDev {  // OpeningBC (BoundaryManager.hpp:27-29)
    int axis, side, a0, a1, b0, b1;
    double T_out;
    int offset;      // first thread of this patch in the launch
};

// Per-cell neighbour code of the pressure operator (built once per scene):
//   bit 0      the cell is an unknown (FLUID or INFLOW)
//   bits 1..6  neighbour -x,+x,-y,+y,-z,+z exists and is an unknown (couples)
//   bits 7..12 that neighbour is OUTFLOW (adds 2*scale to the diagonal)
//   bit 13     red-black colour (i+j+k) & 1
//   bit 14     the cell is OUTFLOW, bit 15 the cell is SOLID
// Walls and SOLID neighbours set neither bit (Neumann), exactly the cases of
// PoissonSolver.cpp:52-87.
enum : uint16_t {
    NB_SELF = 1u << 0,
    NB_UNK0 = 1u << 1,
    NB_OUT0 = 1u << 7,
    NB_COLOR = 1u << 13,
    NB_SELF_OUT = 1u << 14,
    NB_SELF_SOLID = 1u << 15
};

// Per-face mask code (built once per scene, k_face_codes), concatenated
// over the u, v and w face arrays: what the no-slip, fan and outflow
// conditions (BoundaryManager.cpp:60-130, 180-220) do to that face.
//   FC_BOUND   the face lies on the domain boundary (not updated by the stencils)
//   FC_ZERO    no-slip: zeroed (wall face whose cell is not OUTFLOW, or a face of a SOLID cell)
//   FC_FAN     set to the velocity of fan (code >> 8), the last covering fan with a
//              non-zero component along the face's axis
//   FC_OUT_LO / FC_OUT_HI  boundary face of an outflow patch on the low / high
//              wall: after the correction it copies the face one step inside
enum : uint16_t { FC_BOUND = 1, FC_ZERO = 2, FC_FAN = 4, FC_OUT_LO = 8, FC_OUT_HI = 16 };
constexpr int kMaxFans = 255;

// Geometry of one scene as the kernels see it (passed by value).
struct Geo {
    int nx, ny, nz;
    uint32_t nu, nv, nw, nc;
    FastDiv f_nz, f_nz1, f_nynz, f_ny1nz, f_nynz1;
    const uint8_t *flags;
    // FP64 per-axis arrays
    const double *hx, *hy, *hz;        // cell widths dxv, dyv, dzv
    const double *i2x, *i2y, *i2z;     // 1/(h*h), FluidSolver.cpp:101-103
    const double *cx, *cy, *cz;        // centre distance 0.5*(d[i-1]+d[i]) at index i (i>=1)
    const double *ihx, *ihy, *ihz;     // 1/h and 1/centre distance (the FP32 form
    const double *icx, *icy, *icz;     // multiplies where the reference divides)
    const double *lap;                 // slx srx | sly sry | slz srz (PoissonSolver.cpp:34-49)
    // FP32 copies
    const float *hxf, *hyf, *hzf, *ihxf, *ihyf, *ihzf, *i2xf, *i2yf, *i2zf;
    const float *cxf, *cyf, *czf, *icxf, *icyf, *iczf, *lapf;
    const uint16_t *nbm;               // neighbour codes (see NB_*)
    // FP32 per-axis factors of the multigrid's symmetric level-0 operator on a
    // stretched grid: one {w, t below, t above, 0} per index of x | y | z
    // (mg_kernels.hip)
    const float4 *mgf;
    const uint16_t *fcode;             // face mask codes (see FC_*), u | v | w
    double sxd, syd, szd;              // uniform-grid coupling 1/h^2 per axis (0 if not uniform)
    float sx, sy, sz;
    int uniform;
    const FanDev *fans;
    int n_fans;
    const uint8_t *wall_of;            // outflow-patch marks per boundary face
    const OpeningDev *openings;
    int n_openings;
    int open_threads;                  // total patch cells over all openings
};

// Offsets into wall_of: x walls [side][j][k], y walls [side][i][k], z walls [side][i][j].
__host__ __device__ inline uint32_t wall_x(const Geo &g, int side, int j, int k)
{
    return (uint32_t)((side * g.ny + j) * g.nz + k);
}
__host__ __device__ inline uint32_t wall_y(const Geo &g, int side, int i, int k)
{
    return (uint32_t)(2 * g.ny * g.nz + (side * g.nx + i) * g.nz + k);
}
__host__ __device__ inline uint32_t wall_z(const Geo &g, int side, int i, int j)
{
    return (uint32_t)(2 * g.ny * g.nz + 2 * g.nx * g.nz + (side * g.nx + i) * g.ny + j);
}

__device__ __forceinline__ uint32_t cidx(const Geo &g, int i, int j, int k)
{
    return ((uint32_t)i * g.ny + j) * g.nz + k;
}
__device__ __forceinline__ uint32_t uidx(const Geo &g, int i, int j, int k)
{
    return ((uint32_t)i * g.ny + j) * g.nz + k;
}
__device__ __forceinline__ uint32_t vidx(const Geo &g, int i, int j, int k)
{
    return ((uint32_t)i * (g.ny + 1) + j) * g.nz + k;
}
__device__ __forceinline__ uint32_t widx(const Geo &g, int i, int j, int k)
{
    return ((uint32_t)i * g.ny + j) * (g.nz + 1) + k;
}

// Level-0 gather kernels are launched with up to 2048 blocks of 4 waves, 8
// waves per SIMD on 256 CUs: asking for 8 waves per EU (<= 64 VGPRs) keeps
// every block resident, so no block waits for a second round.
#define FFR_RESIDENT __attribute__((amdgpu_waves_per_eu(8)))

// Block b of G remapped so that blocks dispatched to the same XCD (b mod 8)
// own one contiguous range of logical blocks.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t G)
{
    const uint32_t q = G / kXcds, r = G % kXcds;
    const uint32_t x = b % kXcds, slot = b / kXcds;
    return (x < r) ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}
__device__ __forceinline__ uint32_t xcd_block() { return xcd_remap(blockIdx.x, gridDim.x); }

// A block's place in the launch of one scene: its logical (XCD-remapped)
// block, the number of blocks of that launch and its raw index. A plain
// launch takes them from the hardware; a gang launch, which runs the same
// launch of several scenes as one grid, from its table (ffr_gang.hpp).
struct VB {
    uint32_t blk, G, raw;
};
__device__ __forceinline__ VB vb_here() { return VB{xcd_block(), gridDim.x, blockIdx.x}; }

// Contiguous element range [b, e) of logical block `blk` when n elements are
// split over gridDim.x blocks in whole 256-element chunks.
__device__ __forceinline__ void block_span(uint32_t n, uint32_t blk, uint32_t &b, uint32_t &e)
{
    const uint32_t chunks = (n + kBlock - 1) / kBlock;
    const uint32_t per = (chunks + gridDim.x - 1) / gridDim.x;
    const uint64_t b64 = (uint64_t)blk * per * kBlock;
    b = (uint32_t)(b64 < n ? b64 : n);
    const uint64_t e64 = b64 + (uint64_t)per * kBlock;
    e = (uint32_t)(e64 < n ? e64 : n);
}

// The same split of n elements over G of the grid's blocks (a section of
// the grid handling one array).
__device__ __forceinline__ void block_span_of(uint32_t n, uint32_t blk, uint32_t G, uint32_t &b,
                                              uint32_t &e)
{
    const uint32_t chunks = (n + kBlock - 1) / kBlock;
    const uint32_t per = (chunks + G - 1) / G;
    const uint64_t b64 = (uint64_t)blk * per * kBlock;
    b = (uint32_t)(b64 < n ? b64 : n);
    const uint64_t e64 = b64 + (uint64_t)per * kBlock;
    e = (uint32_t)(e64 < n ? e64 : n);
}

// The chunks (256 elements) of n split over the blocks of G so that the
// blocks of one XCD (a contiguous range of logical blocks, see xcd_remap)
// share one contiguous range of chunks and walk it interleaved: block slot s
// of the XCD's g blocks takes chunks first + s, first + s + g, ... < end. The
// XCD's resident blocks then advance through their range together, so data
// a stencil reads again one plane later is still in that XCD's L2.
__device__ __forceinline__ void xcd_chunks(uint32_t n, uint32_t blk, uint32_t G, uint32_t &first,
                                           uint32_t &step, uint32_t &end)
{
    const uint32_t q = G / kXcds, r = G % kXcds;
    const uint32_t x = blk < r * (q + 1) ? blk / (q + 1) : r + (blk - r * (q + 1)) / (q ? q : 1u);
    const uint32_t base = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
    const uint32_t gx = x < r ? q + 1 : q;
    const uint64_t chunks = (n + kBlock - 1) / kBlock;
    first = (uint32_t)(chunks * base / G) + (blk - base);
    end = (uint32_t)(chunks * (base + gx) / G);
    step = gx;
}

// Blocks of a grid of G split over the u, v and w face arrays in proportion
// to their sizes (each at least one), so a block works on one array only.
struct FaceSections {
    uint32_t gu, gv;  // blocks [0, gu) on u, [gu, gu+gv) on v, the rest on w
};
inline FaceSections face_sections(uint64_t nu, uint64_t nv, uint64_t nw, uint32_t G)
{
    const uint64_t n = nu + nv + nw;
    uint32_t gu = (uint32_t)std::max<uint64_t>(1, (G * nu + n / 2) / n);
    uint32_t gv = (uint32_t)std::max<uint64_t>(1, (G * nv + n / 2) / n);
    if (gu + gv >= G) gu = gv = G / 3 > 0 ? G / 3 : 1;
    return {gu, gv};
}
// section of logical block blk of a grid of Gall: axis, block index within
// it, its block count
__device__ __forceinline__ int face_section_of(const FaceSections &fs, uint32_t blk, uint32_t Gall,
                                               uint32_t &lb, uint32_t &G)
{
    if (blk < fs.gu) { lb = blk; G = fs.gu; return 0; }
    if (blk < fs.gu + fs.gv) { lb = blk - fs.gu; G = fs.gv; return 1; }
    lb = blk - fs.gu - fs.gv;
    G = Gall - fs.gu - fs.gv;
    return 2;
}

// ---- reductions -----------------------------------------------------------
// Wave64 tree reductions via cross-lane shuffles, then one LDS stage.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// Sum over the block; the result is valid in thread 0. Fixed order.
__device__ __forceinline__ double block_sum(double v, double *lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x / kWave); w++) s += lds[w];
    }
    __syncthreads();
    return s;
}

__device__ __forceinline__ double block_max(double v, double *lds)
{
    v = wave_max(v);
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x / kWave); w++) s = fmax(s, lds[w]);
    }
    __syncthreads();
    return s;
}

// Max of non-negative doubles through the ordered bit pattern; NaN is never
// passed in (fmax drops it), matching std::max's handling in the reference.
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *slot, double v)
{
    atomicMax(slot, (unsigned long long)__double_as_longlong(v));
}

// Grid reductions without inter-block synchronisation: every block stores
// its partial (block_partial), and the consumer - the next kernel on the
// stream, where the kernel boundary makes the partials visible - re-adds all
// of them in one fixed order (sum_partials), so every block of the consumer
// obtains the same bits. (A last-block-done scheme would need a device-scope
// fence per block, which on gfx950 writes back and invalidates the L2.)
__device__ __forceinline__ void block_partial(double v, double *partials, uint32_t blk,
                                              double *lds)
{
    const double s = block_sum(v, lds);
    if (threadIdx.x == 0) partials[blk] = s;
}

// Sum of partials[0..G) in a fixed order, returned to every thread.
__device__ __forceinline__ double sum_partials(const double *partials, uint32_t G, double *lds)
{
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < G; i += blockDim.x) acc += partials[i];
    const double s = block_sum(acc, lds);
    __shared__ double bcast;
    if (threadIdx.x == 0) bcast = s;
    __syncthreads();
    const double r = bcast;
    __syncthreads();
    return r;
}

// ---- raw buffer access ------------------------------------------------------
// A buffer resource over n elements: element offsets are 32-bit, and a read
// whose offset lies outside [0, n) returns 0 (the hardware range check), so a
// gather at an array end needs no clamping. An index that underflowed as an
// unsigned value ((c - s) >> 1 with c < s) maps to a byte offset of at least
// 2^32 - 2s, outside any array below 2 GiB. Word 3 is the gfx9-family data
// format (32-bit, untyped).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float buf_ldf(__amdgpu_buffer_rsrc_t r, uint32_t i)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(i * 4u), 0, 0));
}
__device__ __forceinline__ double buf_ldd(__amdgpu_buffer_rsrc_t r, uint32_t i)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(i * 8u), 0, 0));
}
__device__ __forceinline__ uint16_t buf_ldu16(__amdgpu_buffer_rsrc_t r, uint32_t i)
{
    return __builtin_amdgcn_raw_buffer_load_b16(r, (int)(i * 2u), 0, 0);
}
__device__ __forceinline__ uint8_t buf_ldu8(__amdgpu_buffer_rsrc_t r, uint32_t i)
{
    return __builtin_amdgcn_raw_buffer_load_b8(r, (int)i, 0, 0);
}
__device__ __forceinline__ void buf_stf(__amdgpu_buffer_rsrc_t r, uint32_t i, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)(i * 4u), 0, 0);
}

// Profiling hook: when a profiler scope has armed a start/stop event pair,
// the next launch records them at the dispatch's own begin and end
// (hipExtLaunchKernelGGL), i.e. the kernel's execution time without the
// inter-kernel gap that events placed around a launch would include.
struct ProfEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local ProfEvents g_prof_events;

// While a gang records (ffr_gang.hpp), launches are recorded, not made; a
// kernel without a gang form marks the record unusable.
struct Recorder;
extern thread_local Recorder *t_rec;
void rec_unsupported();

}  // namespace ffr

#define FFR_LAUNCH(kernel, grid, block, shm, stream, ...)                                   \
    do {                                                                                    \
        if (ffr::t_rec) {                                                                   \
            ffr::rec_unsupported();                                                         \
        } else if (ffr::g_prof_events.start) {                                              \
            hipEvent_t a_ = ffr::g_prof_events.start, b_ = ffr::g_prof_events.stop;        \
            ffr::g_prof_events = ffr::ProfEvents{};                                         \
            hipExtLaunchKernelGGL(kernel, grid, block, shm, stream, a_, b_, 0, __VA_ARGS__); \
        } else {                                                                            \
            hipLaunchKernelGGL(kernel, grid, block, shm, stream, __VA_ARGS__);              \
        }                                                                                   \
    } while (0)
