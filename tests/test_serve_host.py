"""Host logic of the serving layer, without a GPU: admission in arrival
order under slot and HBM limits, completion bookkeeping, latency statistics.
A fake context stands in for the native loop (each submitted scene finishes
after a fixed number of progress passes)."""
import numpy as np
import pytest

from ff_room import serve


class _FakeScene:
    def __init__(self, ctx, ps):
        self.ctx, self.ps = ctx, ps
        ctx.in_use += ps.need
        self.closed = False

    def download(self):
        return {"p": np.full(4, float(self.ps.need))}

    def close(self):
        if not self.closed:
            self.ctx.in_use -= self.ps.need
            self.closed = True


class _FakePS:
    def __init__(self, need, cells):
        self.need, self.cells = need, cells

    def footprint(self):
        return self.need


class _FakeCtx:
    def __init__(self, passes=2):
        self.in_use = 0
        self.passes = passes
        self.tickets = {}
        self.next = 1
        self.max_live = 0

    def serve_config(self, slots):
        self.slots = slots

    def serve_gangs(self, gangs):
        self.gangs = gangs

    def serve_threads(self, threads):
        self.threads = threads

    def device_info(self):
        return {"hbm_bytes": 10_000}

    def bytes_in_use(self):
        return self.in_use

    def submit(self, scene, max_steps):
        t = self.next
        self.next += 1
        self.tickets[t] = [scene, max_steps, 0]
        self.max_live = max(self.max_live, sum(1 for v in self.tickets.values() if v[2] < self.passes))
        return t

    def progress(self):
        n = 0
        for v in self.tickets.values():
            if v[2] < self.passes:
                v[2] += 1
                n += 1
        return n

    def ticket_info(self, t):
        v = self.tickets[t]
        return (v[1] if v[2] >= self.passes else 0), v[2] >= self.passes

    def poll(self, t):
        return True, {"step": self.tickets[t][1]}

    def release(self, t):
        del self.tickets[t]


@pytest.fixture
def fake(monkeypatch):
    ctx = _FakeCtx()
    monkeypatch.setattr(serve.device, "prepare", lambda cfg, **kw: _FakePS(*cfg))
    monkeypatch.setattr(serve.device, "create", lambda c, ps: _FakeScene(c, ps))
    return ctx


def test_admission_respects_slots_and_budget(fake):
    srv = serve.Server(fake, slots=2, hbm_budget=1000, keep_fields=True, tuned=False)
    reqs = [serve.Request((400, 10), 5, 0.0) for _ in range(5)]
    done = srv.run(reqs, realtime=False)
    assert [c.index for c in done] == list(range(5))
    assert fake.max_live <= 2                  # 2 x 400 fits, a third does not
    order = sorted(done, key=lambda c: c.admitted)
    assert [c.index for c in order] == list(range(5))  # arrival order
    assert all(c.steps == 5 and c.fields["p"][0] == 400 for c in done)
    assert fake.in_use == 0                    # every scene freed


def test_head_of_line_waits_for_memory(fake):
    srv = serve.Server(fake, slots=4, hbm_budget=1000, tuned=False)
    done = srv.run([serve.Request((700, 1), 1, 0.0), serve.Request((600, 1), 1, 0.0),
                    serve.Request((100, 1), 1, 0.0)], realtime=False)
    a = {c.index: c for c in done}
    # request 1 (600) cannot join 700; request 2 is not admitted ahead of it
    assert a[1].admitted >= a[0].finished
    assert a[2].admitted >= a[1].admitted


def test_oversized_request_is_rejected(fake):
    srv = serve.Server(fake, slots=1, hbm_budget=100, tuned=False)
    with pytest.raises(MemoryError):
        srv.run([serve.Request((500, 1), 1, 0.0)], realtime=False)


def test_summary_statistics():
    cs = [serve.Completion(i, 100, 10, 0.0, 0.0, 1.0 + i, {}) for i in range(100)]
    s = serve.summary(cs)
    assert s["requests"] == 100 and s["steps"] == 1000
    assert s["p50_latency_s"] == pytest.approx(np.percentile(np.arange(1, 101), 50))
    assert s["p99_latency_s"] == pytest.approx(np.percentile(np.arange(1, 101), 99))
    assert s["cell_updates_per_s"] == pytest.approx(100 * 10 * 100 / 100.0)


def test_work_policies_admit_by_work(monkeypatch):
    from types import SimpleNamespace
    ctx = _FakeCtx()
    monkeypatch.setattr(serve.device, "prepare", lambda cfg, **kw: _FakePS(cfg.need, cfg.cells))
    monkeypatch.setattr(serve.device, "create", lambda c, ps: _FakeScene(c, ps))

    def req(cells, steps, need=100):
        room = SimpleNamespace(grid=(cells, 1, 1))
        return serve.Request(SimpleNamespace(room=room, need=need, cells=cells), steps, 0.0)

    reqs = [req(1000, 10), req(10, 5), req(100, 10), req(10, 50)]  # work 10000, 50, 1000, 500
    srv = serve.Server(ctx, slots=1, hbm_budget=1000, tuned=False, policy="shortest")
    done = srv.run(reqs, realtime=False)
    order = [c.index for c in sorted(done, key=lambda c: c.admitted)]
    assert order == [1, 3, 2, 0]
    srv = serve.Server(ctx, slots=1, hbm_budget=1000, tuned=False, policy="longest")
    done = srv.run(reqs, realtime=False)
    assert [c.index for c in sorted(done, key=lambda c: c.admitted)] == [0, 2, 3, 1]
    srv = serve.Server(ctx, slots=1, hbm_budget=1000, tuned=False)  # fifo keeps arrival order
    done = srv.run(reqs, realtime=False)
    assert [c.index for c in sorted(done, key=lambda c: c.admitted)] == [0, 1, 2, 3]
    with pytest.raises(ValueError):
        serve.Server(ctx, policy="random")
