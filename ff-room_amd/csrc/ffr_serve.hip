// ffr_serve.hip - continuous batching of many scenes on one device.
//
// A ticket asks for up to max_steps projection steps of one scene (stopping
// early on FluidSolver::run's convergence test, FluidSolver.cpp:552-560).
// Up to `slots` tickets are active at once; the rest wait in arrival order
// and are admitted as active ones finish. Every scene has its own stream, and
// each pass of the loop advances every active scene's step task without
// blocking (ffr_task.hip): a scene whose event has been reached gets its next
// phase or next step enqueued, so the kernels of all active scenes overlap on
// the device. Admission by HBM footprint happens where the memory is taken:
// ffr_scene_create refuses a scene beyond the context's budget, and the
// caller creates scenes as earlier ones are destroyed.
#include "ffr_gang.hpp"
#include "ffr_scene.hpp"

#include <algorithm>
#include <chrono>
#include <atomic>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

using namespace ffr;


struct ffr_ctx::Server {
    struct Ticket {
        ffr_scene *scene = nullptr;
        int max_steps = 0, steps = 0;
        bool active = false, done = false;
        int rc = FFR_OK;
        ffr_step_result last{};
        int gang = -2;  // -2: not yet placed, -1: on the scene's own stream
        int worker = -1;  // the host thread advancing it (-1: the caller's)
        std::unique_ptr<Recorder> rec;
    };
    std::map<uint64_t, Ticket> tickets;
    std::deque<uint64_t> waiting;
    std::vector<uint64_t> active;
    uint64_t next_id = 1;
    int slots = 16;
    // gangs (ffr_gang.hip): 0 = every scene on its own stream
    int ngangs = 0;
    std::vector<std::unique_ptr<Gang>> gangs;
    std::vector<std::vector<uint64_t>> members;
    ffr_ctx *ctx = nullptr;
    // Host threads advancing the scenes on their own streams. A scene's step
    // is a few hundred launches issued by the host, and one thread issuing
    // for every scene caps the served steps per second below what the device
    // sustains (measured on MI355X with 72x49x58 rooms: about 1000
    // scene-steps/s from one thread, 1600 from four). With nthreads > 1 (and
    // no gangs) each admitted ticket is given to one of nthreads workers,
    // which run their tickets one at a time (work); the caller's passes then
    // only admit, collect and report. `mu` guards
    // the tickets' bookkeeping, the queues and the workers' lists; a scene's
    // step task is touched by its worker alone while its ticket is active.
    static constexpr int kDefaultThreads = 8;
    int nthreads = kDefaultThreads;
    std::mutex mu;
    std::vector<std::thread> workers;
    std::vector<std::vector<uint64_t>> own;  // per worker: its tickets
    std::vector<uint64_t> passes;            // per worker: passes completed
    std::vector<char> inflight;              // per worker: advancing outside the lock
    std::atomic<bool> stop{false};
    int completed_pending = 0;  // steps the workers completed since the last pass

    void finish(Ticket &t, int rc)
    {
        if (t.done) return;  // the first outcome stands (a forgotten scene's, say)
        t.done = true;
        t.active = false;
        t.rc = rc;
    }

    // a step of ticket t completed: bookkeeping; whether it goes on
    bool step_done(Ticket &t, int &completed)
    {
        if (t.done) return false;
        completed++;
        t.steps++;
        t.last = t.scene->task.result;
        if (t.last.converged || t.steps >= t.max_steps) {
            finish(t, FFR_OK);
            return false;
        }
        return true;
    }

    int set_gangs(int n)
    {
        for (const uint64_t id : active)
            if (tickets[id].gang != -2) {
                ctx->fail("ffr_serve_gangs: tickets are in flight");
                return FFR_E_STATE;
            }
        gangs.clear();
        members.assign(n, {});
        for (int g = 0; g < n; g++) {
            gangs.emplace_back(new Gang());
            const int rc = gangs.back()->init(ctx);
            if (rc != FFR_OK) {
                gangs.clear();
                members.clear();
                ngangs = 0;
                return rc;
            }
        }
        ngangs = n;
        return FFR_OK;
    }

    // places newly admitted tickets: a gang when the scene can take part
    void place()
    {
        for (const uint64_t id : active) {
            Ticket &t = tickets[id];
            if (t.gang != -2 || t.done) continue;
            t.gang = -1;
            if (ngangs == 0) continue;
            bool eligible = false;
            const int rc = task_prepare_gang(t.scene, &eligible);
            if (rc != FFR_OK) {
                finish(t, rc);
                continue;
            }
            if (!eligible) continue;
            int g = 0;
            for (int k = 1; k < ngangs; k++)
                if (members[k].size() < members[g].size()) g = k;
            t.gang = g;
            if (!t.rec) t.rec.reset(new Recorder());
            members[g].push_back(id);
        }
    }

    // one round of gang g, whose previous round (if any) has completed: every
    // member resumes or begins its step with its recorder active, then the
    // records go out as merged launches
    int gang_round(int g, int &completed)
    {
        Gang &G = *gangs[g];
        const auto t0 = std::chrono::steady_clock::now();
        struct HostTime {  // host time of the round, however it ends
            Gang &G;
            std::chrono::steady_clock::time_point t0;
            ~HostTime()
            {
                G.host_us += std::chrono::duration<double, std::micro>(
                                 std::chrono::steady_clock::now() - t0).count();
            }
        } ht{G, t0};
        std::vector<Recorder *> recs;
        for (const uint64_t id : members[g]) {
            Ticket &t = tickets[id];
            ffr_scene *s = t.scene;
            Recorder &r = *t.rec;
            r.clear();
            t_rec = &r;
            int rc = FFR_OK;
            bool go = true;
            if (task_busy(s)) {
                bool done = false;
                rc = task_resume(s, &done);
                go = rc == FFR_OK && done && step_done(t, completed);
            } else if (t.steps >= t.max_steps) {
                finish(t, FFR_OK);
                go = false;
            }
            if (rc == FFR_OK && go) rc = task_begin(s, false);
            t_rec = nullptr;
            if (rc == FFR_OK && r.unsupported) {
                ctx->fail("gang: a kernel without a gang form was reached");
                rc = FFR_E_UNSUPPORTED;
            }
            if (rc != FFR_OK) {
                s->task.phase = StepTask::IDLE;
                finish(t, rc);
                continue;
            }
            if (!t.done && !r.ops.empty()) recs.push_back(&r);
        }
        members[g].erase(std::remove_if(members[g].begin(), members[g].end(),
                                        [&](uint64_t id) { return tickets[id].done; }),
                         members[g].end());
        if (recs.empty()) return FFR_OK;
        int rc = G.launch_round(recs);
        if (rc != FFR_OK) {
            // the members' launches that did go out finish before they are failed
            if (hipStreamSynchronize(G.stream) != hipSuccess) rc = FFR_E_DEVICE;
            for (const uint64_t id : members[g]) {
                tickets[id].scene->task.phase = StepTask::IDLE;
                finish(tickets[id], rc);
            }
            members[g].clear();
            return rc;
        }
        G.in_flight = true;
        return FFR_OK;
    }

    int progress_gangs()
    {
        int completed = 0;
        for (int g = 0; g < ngangs; g++) {
            Gang &G = *gangs[g];
            if (G.in_flight) {
                const hipError_t q = hipEventQuery(G.ev);
                if (q == hipErrorNotReady) continue;
                G.in_flight = false;
                if (q != hipSuccess) {
                    ctx->fail("hipEventQuery failed: %s", hipGetErrorString(q));
                    for (const uint64_t id : members[g]) {
                        tickets[id].scene->task.phase = StepTask::IDLE;
                        finish(tickets[id], FFR_E_DEVICE);
                    }
                    members[g].clear();
                    continue;
                }
            }
            if (!members[g].empty()) gang_round(g, completed);
        }
        return completed;
    }

    void admit()
    {
        while (!waiting.empty() && (int)active.size() < slots) {
            const uint64_t id = waiting.front();
            waiting.pop_front();
            Ticket &t = tickets[id];
            t.active = true;
            active.push_back(id);
        }
    }

    // advances ticket t (on its scene's own stream) without blocking; the
    // ticket's bookkeeping is changed under `lock` when given (a worker), and
    // t is not touched once it is done (its owner may release it at once).
    // Returns whether anything happened.
    bool advance(Ticket &t, int &completed, std::mutex *lock)
    {
        auto book = [&](auto &&f) {
            if (lock) {
                std::lock_guard<std::mutex> lk(*lock);
                f();
            } else {
                f();
            }
        };
        ffr_scene *s = t.scene;
        if (!task_busy(s)) {
            if (t.steps >= t.max_steps) {
                book([&] { finish(t, FFR_OK); });
                return true;
            }
            const int rc = task_begin(s, false);
            if (rc != FFR_OK) book([&] { finish(t, rc); });
            return true;
        }
        bool done = false;
        const int rc = task_advance(s, false, &done);
        if (rc != FFR_OK) {
            book([&] { finish(t, rc); });
            return true;
        }
        if (!done) return false;
        bool more = false;
        book([&] { more = step_done(t, completed); });
        if (!more) return true;
        const int rb = task_begin(s, false);  // the scene's next step, at once
        if (rb != FFR_OK) book([&] { finish(t, rb); });
        return true;
    }

    bool threaded() const { return nthreads > 1 && ngangs == 0; }

    // worker w: runs one ticket at a time to its end, waiting on the device
    // (hipEventSynchronize) at each of its step's host checks, so that the
    // ticket's next launches follow at once; measured on MI355X, blocking
    // waits sustain more scene-steps per second than polling from a few
    // threads, whose event queries contend with the other threads' launches
    void work(int w)
    {
        hipSetDevice(ctx->device);
        while (!stop.load(std::memory_order_relaxed)) {
            Ticket *t = nullptr;
            {
                std::lock_guard<std::mutex> lk(mu);
                for (const uint64_t id : own[w]) {
                    auto it = tickets.find(id);
                    if (it != tickets.end() && !it->second.done) {
                        t = &it->second;
                        break;
                    }
                }
                inflight[w] = t ? 1 : 0;
            }
            if (!t) {
                std::this_thread::sleep_for(std::chrono::microseconds(50));
                continue;
            }
            run_ticket(*t);
            std::lock_guard<std::mutex> lk(mu);
            inflight[w] = 0;
            passes[w]++;
            auto &o = own[w];
            o.erase(std::remove_if(o.begin(), o.end(),
                                   [&](uint64_t id) {
                                       auto it = tickets.find(id);
                                       return it == tickets.end() || it->second.done;
                                   }),
                    o.end());
        }
    }

    // ticket t's steps until it ends (its budget, convergence, an error, or
    // being marked done by serve_forget) or the workers stop; bookkeeping
    // under mu, completed steps counted as they happen
    void run_ticket(Ticket &t)
    {
        ffr_scene *s = t.scene;
        while (!stop.load(std::memory_order_relaxed)) {
            {
                std::lock_guard<std::mutex> lk(mu);
                if (t.done) return;
                if (!task_busy(s) && t.steps >= t.max_steps) {
                    finish(t, FFR_OK);
                    return;
                }
            }
            if (!task_busy(s)) {
                const int rc = task_begin(s, false);
                if (rc != FFR_OK) {
                    std::lock_guard<std::mutex> lk(mu);
                    finish(t, rc);
                    return;
                }
            }
            bool done = false;
            const int rc = task_advance(s, true, &done);
            std::lock_guard<std::mutex> lk(mu);
            if (rc != FFR_OK) {
                finish(t, rc);
                return;
            }
            if (!done) continue;
            int one = 0;
            const bool more = step_done(t, one);
            completed_pending += one;
            if (!more) return;
        }
    }

    void start_workers()
    {
        if (!workers.empty()) return;
        stop = false;
        own.assign(nthreads, {});
        passes.assign(nthreads, 0);
        inflight.assign(nthreads, 0);
        for (int w = 0; w < nthreads; w++) workers.emplace_back([this, w] { work(w); });
    }

    // stops the workers; their unfinished tickets return to the caller's passes
    void stop_workers()
    {
        if (workers.empty()) return;
        stop = true;
        for (auto &t : workers) t.join();
        workers.clear();
        own.clear();
        for (auto &kv : tickets) kv.second.worker = -1;
    }

    // gives each active ticket without a worker to the least loaded one (mu held)
    void assign_locked()
    {
        for (const uint64_t id : active) {
            Ticket &t = tickets[id];
            if (t.done || t.worker >= 0) continue;
            int w = 0;  // the worker with the fewest tickets (an idle one first)
            for (int k = 1; k < (int)own.size(); k++)
                if (own[k].size() < own[w].size()) w = k;
            t.worker = w;
            t.gang = -1;
            own[w].push_back(id);
        }
    }

    // waits until no worker is advancing ticket id's scene (mu held on entry
    // and exit; released while waiting)
    void wait_unused(std::unique_lock<std::mutex> &lk, const Ticket &t)
    {
        const int w = t.worker;
        if (w < 0 || w >= (int)inflight.size() || !inflight[w]) return;
        const uint64_t p = passes[w];
        while (inflight[w] && passes[w] == p) {
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        }
    }

    // one non-blocking pass over the active tickets; returns the number of
    // steps that completed (mu held)
    int progress()
    {
        admit();
        if (threaded()) {
            start_workers();
            assign_locked();
            const int completed = completed_pending;
            completed_pending = 0;
            active.erase(std::remove_if(active.begin(), active.end(),
                                        [&](uint64_t id) { return tickets[id].done; }),
                         active.end());
            admit();
            assign_locked();
            return completed;
        }
        place();
        int completed = ngangs > 0 ? progress_gangs() : 0;
        for (const uint64_t id : active) {
            Ticket &t = tickets[id];
            if (!t.done && t.gang < 0) advance(t, completed, nullptr);
        }
        active.erase(std::remove_if(active.begin(), active.end(),
                                    [&](uint64_t id) { return tickets[id].done; }),
                     active.end());
        admit();
        place();
        return completed;
    }

    bool idle() const { return active.empty() && waiting.empty(); }
};

void ffr_ctx::serve_shutdown()
{
    if (server) {
        server->stop_workers();
        // let in-flight work finish before the scenes' memory can go away
        for (const uint64_t id : server->active) {
            ffr_scene *s = server->tickets[id].scene;
            if (s) hipStreamSynchronize(s->stream);
        }
        for (auto &g : server->gangs) hipStreamSynchronize(g->stream);
    }
    delete server;
    server = nullptr;
}

namespace ffr {
// A scene being destroyed leaves the serving loop: its ticket ends with
// FFR_E_STATE unless it had finished.
void serve_forget(ffr_ctx *c, ffr_scene *s)
{
    if (!c->server) return;
    auto *sv = c->server;
    std::unique_lock<std::mutex> lk(sv->mu);
    for (auto &kv : sv->tickets) {
        if (kv.second.scene != s) continue;
        // done first, so that no worker takes it up again; then wait for a
        // worker that may be advancing it
        if (!kv.second.done) sv->finish(kv.second, FFR_E_STATE);
        sv->wait_unused(lk, kv.second);
        const int w = kv.second.worker;
        if (w >= 0 && w < (int)sv->own.size()) {
            auto &o = sv->own[w];
            o.erase(std::remove(o.begin(), o.end(), kv.first), o.end());
        }
        const int g = kv.second.gang;
        if (g >= 0 && g < sv->ngangs) {  // its gang's launches may still use the scene
            hipStreamSynchronize(sv->gangs[g]->stream);
            auto &m = sv->members[g];
            m.erase(std::remove(m.begin(), m.end(), kv.first), m.end());
            s->task.phase = StepTask::IDLE;
        }
        if (!kv.second.done) sv->finish(kv.second, FFR_E_STATE);
        kv.second.scene = nullptr;
    }
    sv->active.erase(std::remove_if(sv->active.begin(), sv->active.end(),
                                    [&](uint64_t id) { return sv->tickets[id].done; }),
                     sv->active.end());
    sv->waiting.erase(std::remove_if(sv->waiting.begin(), sv->waiting.end(),
                                     [&](uint64_t id) { return sv->tickets[id].done; }),
                      sv->waiting.end());
}
}  // namespace ffr

static ffr_ctx::Server *server_of(ffr_ctx *c)
{
    if (!c->server) {
        c->server = new (std::nothrow) ffr_ctx::Server();
        if (c->server) c->server->ctx = c;
    }
    return c->server;
}

extern "C" {

int ffr_serve_config(ffr_ctx *c, int slots)
{
    if (!c || slots < 1) return FFR_E_INVALID;
    auto *sv = server_of(c);
    if (!sv) return FFR_E_NOMEM;
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->slots = slots;
    return FFR_OK;
}

int ffr_serve_threads(ffr_ctx *c, int threads)
{
    if (!c || threads < 1 || threads > 64) return FFR_E_INVALID;
    auto *sv = server_of(c);
    if (!sv) return FFR_E_NOMEM;
    sv->stop_workers();  // restarted with the new count at the next pass
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->nthreads = threads;
    return FFR_OK;
}

int ffr_serve_gang_stats(ffr_ctx *c, ffr_gang_stats *out)
{
    if (!c || !out) return FFR_E_INVALID;
    *out = ffr_gang_stats{};
    if (c->server)
        for (const auto &g : c->server->gangs) {
            out->rounds += g->rounds;
            out->launches += g->merged_launches;
            out->ops += g->recorded_ops;
            out->longest_ops += g->longest_ops;
            out->host_us += g->host_us;
        }
    return FFR_OK;
}

int ffr_serve_gangs(ffr_ctx *c, int gangs)
{
    if (!c || gangs < 0 || gangs > 64) return FFR_E_INVALID;
    auto *sv = server_of(c);
    if (!sv) return FFR_E_NOMEM;
    sv->stop_workers();  // gangs are advanced by the caller's passes
    std::lock_guard<std::mutex> lk(sv->mu);
    return sv->set_gangs(gangs);
}

int ffr_submit(ffr_ctx *c, ffr_scene *s, int max_steps, uint64_t *ticket)
{
    if (!c || !s || !ticket || max_steps < 0 || s->ctx != c) return FFR_E_INVALID;
    auto *sv = server_of(c);
    if (!sv) return FFR_E_NOMEM;
    std::lock_guard<std::mutex> lk(sv->mu);
    for (const auto &kv : sv->tickets)
        if (kv.second.scene == s && !kv.second.done) {
            c->fail("ffr_submit: the scene already has an unfinished ticket");
            return FFR_E_STATE;
        }
    if (task_busy(s)) {
        c->fail("ffr_submit: the scene has a step in flight");
        return FFR_E_STATE;
    }
    const uint64_t id = sv->next_id++;
    ffr_ctx::Server::Ticket &t = sv->tickets[id];
    t.scene = s;
    t.max_steps = max_steps;
    sv->waiting.push_back(id);
    sv->admit();
    *ticket = id;
    return FFR_OK;
}

int ffr_poll(ffr_ctx *c, uint64_t ticket, ffr_step_result *last, int *done)
{
    if (!c || !done) return FFR_E_INVALID;
    auto *sv = c->server;
    if (!sv) {
        c->fail("ffr_poll: unknown ticket %llu", (unsigned long long)ticket);
        return FFR_E_NOT_FOUND;
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    if (!sv->tickets.count(ticket)) {
        c->fail("ffr_poll: unknown ticket %llu", (unsigned long long)ticket);
        return FFR_E_NOT_FOUND;
    }
    sv->progress();
    const auto &t = sv->tickets[ticket];
    *done = t.done ? 1 : 0;
    if (last) *last = t.last;
    return t.done ? t.rc : FFR_OK;
}

int ffr_progress(ffr_ctx *c, int *steps_completed)
{
    if (!c) return FFR_E_INVALID;
    auto *sv = c->server;
    int n = 0;
    if (sv) {
        std::lock_guard<std::mutex> lk(sv->mu);
        n = sv->progress();
    }
    if (steps_completed) *steps_completed = n;
    return FFR_OK;
}

int ffr_drain(ffr_ctx *c)
{
    if (!c) return FFR_E_INVALID;
    auto *sv = c->server;
    if (!sv) return FFR_OK;
    for (;;) {
        int n = 0;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            if (sv->idle()) break;
            n = sv->progress();
        }
        if (n == 0) {  // with worker threads, leave them the lock
            if (sv->threaded()) std::this_thread::sleep_for(std::chrono::microseconds(100));
            else std::this_thread::yield();
        }
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    for (const auto &kv : sv->tickets)
        if (kv.second.rc != FFR_OK) return kv.second.rc;
    return FFR_OK;
}

int ffr_ticket_info(ffr_ctx *c, uint64_t ticket, int *steps, int *done)
{
    if (!c) return FFR_E_INVALID;
    auto *sv = c->server;
    if (!sv) return FFR_E_NOT_FOUND;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (!sv->tickets.count(ticket)) return FFR_E_NOT_FOUND;
    const auto &t = sv->tickets[ticket];
    if (steps) *steps = t.steps;
    if (done) *done = t.done ? 1 : 0;
    return t.done ? t.rc : FFR_OK;
}

int ffr_release(ffr_ctx *c, uint64_t ticket)
{
    if (!c) return FFR_E_INVALID;
    auto *sv = c->server;
    if (!sv) return FFR_E_NOT_FOUND;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (!sv->tickets.count(ticket)) return FFR_E_NOT_FOUND;
    if (!sv->tickets[ticket].done) {
        c->fail("ffr_release: ticket %llu has not finished", (unsigned long long)ticket);
        return FFR_E_STATE;
    }
    sv->tickets.erase(ticket);
    return FFR_OK;
}
}
