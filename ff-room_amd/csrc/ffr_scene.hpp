// ffr_scene.hpp - host-side definitions of the opaque ABI objects.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <set>
#include <memory>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ffroom.h"
#include "ffr_kernels.hpp"
#include "ffr_lbm.hpp"
#include "ffr_mg.hpp"

namespace ffr {
// one pinned host block per scene holds its StepRec, CgState and MgState mirrors
constexpr size_t pinned_slot(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kPinnedBlock =
    pinned_slot(sizeof(StepRec)) + pinned_slot(sizeof(CgState)) + pinned_slot(sizeof(MgState));
}  // namespace ffr

struct ffr_ctx {
    int device = 0;
    hipDeviceProp_t prop{};
    size_t budget = 0;  // HBM bytes this context may hold
    size_t in_use = 0;
    char err[512] = {0};
    std::mutex err_mu;  // serving threads may fail at once
    void fail(const char *fmt, ...);

    // profiling
    struct PendingEvent {
        hipEvent_t a, b;
        const char *name;
        double bytes;
    };
    struct ProfEntry {
        long launches = 0;
        double total_ms = 0;
        double bytes = 0;
    };
    bool prof_on = false;
    bool open_event = false;
    std::vector<PendingEvent> pending;
    std::vector<hipEvent_t> ev_pool;
    std::map<std::string, ProfEntry> prof_map;
    void prof_begin(hipStream_t s, const char *name, double bytes);
    void prof_end(hipStream_t s);
    void prof_resolve();
    void prof_clear();

    // scenes created on this context (destroyed with it)
    std::set<ffr_scene *> scenes;
    // hierarchies built so far: spreads the scenes' multigrid tail groups over
    // the XCDs (hierarchies are built by serving worker threads too)
    std::atomic<unsigned> tail_seq{0};

    // Device blocks, pinned host blocks, streams and events of destroyed
    // scenes, kept for the scenes created next (ffr_pool.hip). hipFree and
    // hipHostFree wait for the whole device, so freeing them each time a
    // request ends would stall every other scene's work in a serving loop.
    // `in_use` counts the blocks scenes hold; cached device blocks are freed,
    // largest first, when a new block would not fit the budget beside them.
    struct Pool {
        std::mutex mu;
        std::multimap<size_t, void *> dev;  // free device blocks by size
        size_t dev_bytes = 0;
        std::vector<void *> pinned;         // free blocks of kPinnedBlock bytes
        std::vector<hipStream_t> streams;
        std::vector<hipEvent_t> events;
    } pool;
    void *dev_take(size_t bytes, size_t *got);  // nullptr: allocation failed
    void dev_give(void *p, size_t bytes);
    void *pinned_take();
    void pinned_give(void *p);
    hipStream_t stream_take();
    void stream_give(hipStream_t st);
    hipEvent_t event_take();
    void event_give(hipEvent_t ev);
    size_t pool_trim();  // frees every cached block; returns the device bytes freed

    // serving (ffr_serve.hip)
    struct Server;
    Server *server = nullptr;
    void serve_shutdown();
};

namespace ffr {
// An in-flight projection step (or a stand-alone pressure solve) of one
// scene. The host enqueues as far as the scene's predictions allow, records
// `ev`, and resumes when the device has reached it; kernels past the point
// where the solve stopped return at once (device flags), so enqueuing ahead
// is safe. See ffr_task.hip.
struct StepTask {
    enum Phase { IDLE = 0, SOLVE, RESULT } phase = IDLE;
    bool solve_only = false;
    hipEvent_t ev = nullptr;
    // CG: iterations enqueued
    int cg_enq = 0;
    // MG-PCG: residual passes enqueued / wanted, the inner solve waiting for
    // a host check (-1: none), iterations enqueued per inner solve
    int passes_enq = 0, target_passes = 0, inner_wait = -1;
    int k_used[kMgHist] = {};
    // outcome
    int iters = 0, outer = 0;
    double res = 0.0;
    ffr_step_result result{};
};
// Iteration counts of the previous solve, used to enqueue the next one.
struct SolvePred {
    int cg_iters = 0;
    int mg_passes = 0;
    int mg_inner[kMgHist] = {};
};
}  // namespace ffr

// RAII-free scope marker used around each launch.
struct ProfScope {
    ffr_ctx *c;
    hipStream_t s;
    ProfScope(ffr_ctx *c_, hipStream_t s_, const char *name, double bytes) : c(c_), s(s_)
    {
        c->prof_begin(s, name, bytes);
    }
    ~ProfScope() { c->prof_end(s); }
};

struct ffr_scene {
    ffr_ctx *ctx = nullptr;
    ffr_params prm{};
    int nx = 0, ny = 0, nz = 0;
    uint64_t nu = 0, nv = 0, nw = 0, nc = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;

    // host copies of the geometry
    std::vector<double> dxv, dyv, dzv;
    std::vector<uint8_t> flags;
    std::vector<ffr_fan> fans;
    std::vector<ffr_outflow> outflows;
    std::vector<ffr_opening> openings;
    bool opening_touches_fluid = false;
    long n_fluid = 0;

    // device memory
    void *arena = nullptr;
    size_t arena_bytes = 0;
    ffr::Geo geo{};
    int fp64 = 1;            // state precision: 1 = double, 0 = float
    void *u = nullptr, *v = nullptr, *w = nullptr, *T = nullptr;
    void *us = nullptr, *vs = nullptr, *ws = nullptr, *Tn = nullptr;
    double *p = nullptr, *b = nullptr;
    double *r = nullptr, *d = nullptr, *Ad = nullptr;  // FP64 CG / outer residual
    ffr::StepRec *rec = nullptr;
    ffr::CgState *cg = nullptr;
    ffr::RedScratch red{};
    double *scratch64 = nullptr;  // nc doubles for host transfers of fp32 state
    // mixed form: the pressures of the two steps before the last (warm_start
    // 2 and 3 extrapolate from them); p_hist_n of them are valid
    double *p_hist = nullptr, *p_hist2 = nullptr;
    int p_hist_n = 0;
    ffr::Multigrid mg{};          // multigrid hierarchy (when allocated)
    ffr::MgState *mgs = nullptr;  // device scalars of the multigrid solve

    // pinned host mirrors (three slots of one kPinnedBlock-byte block)
    void *h_block = nullptr;
    ffr::StepRec *h_rec = nullptr;
    ffr::CgState *h_cg = nullptr;
    ffr::MgState *h_mgs = nullptr;

    int step = 0;
    ffr::StepTask task;
    ffr::SolvePred pred;

    // snapshot egress (ffr_snapshot.hip): device slot, pinned mirror, events
    struct Snap {
        bool busy = false;
        int step = 0, kind = 0;
        void *dev = nullptr, *host = nullptr;
        hipEvent_t copied = nullptr, landed = nullptr;
    };
    Snap snaps[FFR_SNAPSHOT_SLOTS];
    ffr::LbmState lbm;  // the D3Q19 lattice, once attached (ffr_lbm.hip)
    size_t snap_bytes = 0;  // bytes of one slot
    hipStream_t copy_stream = nullptr;
};

// Returns `code` from the calling function after recording the HIP error.
#define HIP_CHECK_CTX(ctx, call, code)                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            (ctx)->fail("%s failed: %s", #call, hipGetErrorString(e_));                 \
            return code;                                                                \
        }                                                                               \
    } while (0)

namespace ffr {
// internal entry points shared between translation units
int resolve_method(const ffr_params &p);
StepParams make_step_params(const ffr_params &p);
double bits_to_double(unsigned long long b);
size_t scene_bytes(const ffr_scene_desc *d, const ffr_params *p, bool with_mg);
bool scene_uses_mg(const ffr_params &p);
// step / solve tasks (ffr_task.hip)
int task_begin(ffr_scene *s, bool solve_only);
int task_advance(ffr_scene *s, bool block, bool *done);
bool task_busy(const ffr_scene *s);
int scene_step(ffr_scene *s, ffr_step_result *out);
int wait_point(ffr_scene *s, void *host, const void *dev, size_t bytes);
// one iteration of a task whose last event the device has reached: handles
// it and enqueues (or records) the next part; *done when the step finished
int task_resume(ffr_scene *s, bool *done);
// what must exist before a scene's step can be recorded for a gang (the
// multigrid hierarchy); false when the scene cannot take part in a gang
int task_prepare_gang(ffr_scene *s, bool *eligible);
// serving (ffr_serve.hip)
void serve_forget(ffr_ctx *c, ffr_scene *s);
// snapshot egress (ffr_snapshot.hip)
void snapshot_release(ffr_scene *s);
// the lattice of a scene (ffr_lbm.hip)
void lbm_release(ffr_scene *s);

struct Recorder;
// Several scenes' step tasks advanced in rounds of merged launches on one
// stream (ffr_gang.hip).
struct Gang {
    ffr_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    uint8_t *d_tab = nullptr, *h_stage = nullptr;  // tables: device / pinned staging
    size_t cap = 0;
    std::vector<uint8_t> stage_buf;
    bool in_flight = false;
    // rounds, launches issued, operations recorded by the members, the
    // longest member's operations (summed over rounds), host time spent
    // recording and issuing rounds
    unsigned long long rounds = 0, merged_launches = 0, recorded_ops = 0, longest_ops = 0;
    double host_us = 0.0;
    Gang() = default;
    Gang(const Gang &) = delete;
    Gang &operator=(const Gang &) = delete;
    ~Gang();
    int init(ffr_ctx *c);
    void release();
    int reserve(size_t bytes);
    int launch_round(const std::vector<Recorder *> &recs);
};
// multigrid (ffr_mg.hip)
size_t mg_bytes(int nx, int ny, int nz, const double *dxv, const double *dyv, const double *dzv,
                const ffr_params &p);
int mg_create(ffr_scene *s);
void mg_destroy(ffr_scene *s);
int mg_begin(ffr_scene *s);
int mg_continue(ffr_scene *s, bool *finished);
// sets *launch to one launch of a level-0 multigrid / inner-PCG / outer kernel
// on the scene's buffers and returns its algorithmic bytes
double mg_prof_launch(ffr_scene *s, const char *name, std::function<void()> *launch);
}  // namespace ffr
