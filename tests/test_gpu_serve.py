"""Continuous batching: many scenes served together on one GPU.

Pins: each scene's fields after K steps in a batch equal, bit for bit, the
fields of the same scene stepped alone with ffr_step (every reduction has a
fixed order and each scene's iteration schedule depends only on its own
history), for mixed-precision MG-PCG and FP64 CG scenes of uneven, odd sizes;
admission under an HBM budget that fits one scene at a time; the early stop
of FluidSolver::run's convergence test; and the ticket API's errors.
"""
import numpy as np
import pytest

from ff_room import autotune, capi, device
from ff_room.generator import random_scene
from ff_room.serve import Request, Server, summary

pytestmark = pytest.mark.gpu

K = 6


def _scenes():
    """Uneven grids (odd extents included), mixed and FP64 forms."""
    base = dict(dt=0.005, max_steps=K, convergence_tol=0.0, advection="quick", buoyancy=True,
                beta=3.4e-3, alpha=2.1e-5, T_initial=26.0)
    return [
        random_scene(101, grid=(40, 32, 16), n_fans=2, n_obstacles=2, precision="mixed", **base),
        random_scene(102, grid=(33, 27, 19), n_fans=1, n_obstacles=3, precision="mixed", **base),
        random_scene(103, grid=(24, 30, 12), n_fans=2, n_obstacles=1, precision="fp64", **base),
        random_scene(104, grid=(48, 20, 25), n_fans=3, n_obstacles=0, precision="mixed",
                     openings=True, **base),
    ]


def _alone(ctx, cfg, steps, tuned=False):
    ps = device.prepare(cfg)
    if tuned:  # the per-scene entries of the tuning table, as the Server applies them
        entry = autotune.lookup(ps.nx, ps.ny, ps.nz, ps.params.precision)
        if entry:
            autotune.apply(ps.params, {k: v for k, v in entry.items() if k != "max_blocks"})
    d = device.create(ctx, ps)
    last = None
    for _ in range(steps):
        last = d.step()
    f = d.download()
    d.close()
    return f, last


def test_batch_equals_alone_bitwise(ctx):
    cfgs = _scenes()
    alone = [_alone(ctx, c, K) for c in cfgs]
    srv = Server(ctx, slots=len(cfgs), keep_fields=True)
    done = srv.run([Request(c, K) for c in cfgs], realtime=False)
    assert [c.index for c in done] == list(range(len(cfgs)))
    for (fa, la), c in zip(alone, done):
        assert c.steps == K
        for k in ("u", "v", "w", "p", "T"):
            assert np.array_equal(fa[k], c.fields[k]), k
        for k in ("velocity_change", "pressure_residual", "divergence_max", "T_mean",
                  "poisson_iters", "outer_iters"):
            assert la[k] == c.last[k], k
    s = summary(done)
    assert s["steps"] == K * len(cfgs) and s["p99_latency_s"] >= s["p50_latency_s"] > 0


def test_slot_limit_and_budget_admission(ctx):
    """A budget that holds one scene at a time: requests run one after the
    other, in arrival order, with the same results as alone."""
    cfgs = _scenes()[:3]
    sizes = [device.prepare(c).footprint() for c in cfgs]
    srv = Server(ctx, slots=4, hbm_budget=ctx.bytes_in_use() + max(sizes) + 1024, keep_fields=True)
    done = srv.run([Request(c, 3) for c in cfgs], realtime=False)
    order = sorted(done, key=lambda c: c.admitted)
    assert [c.index for c in order] == [0, 1, 2]
    for a, b in zip(order, order[1:]):
        assert b.admitted >= a.finished - 1e-3  # admitted only after memory was freed
    for c, cfg in zip(done, cfgs):
        fa, _ = _alone(ctx, cfg, 3)
        assert np.array_equal(fa["p"], c.fields["p"])


def test_convergence_ends_a_ticket_early(ctx):
    cfg = _scenes()[0]
    cfg.solver.convergence_tol = 1e30  # converged after the first step
    srv = Server(ctx, slots=2)
    done = srv.run([Request(cfg, 50)], realtime=False)
    assert done[0].steps == 1 and done[0].last["converged"] == 1


def test_ticket_api(ctx):
    ps = device.prepare(_scenes()[0])
    d = device.create(ctx, ps)
    t = ctx.submit(d, 2)
    with pytest.raises(capi.FfrError):
        ctx.submit(d, 2)  # one unfinished ticket per scene
    with pytest.raises(capi.FfrError):
        ctx.release(t)    # not finished yet
    ctx.drain()
    steps, fin = ctx.ticket_info(t)
    assert fin and steps == 2
    fin2, last = ctx.poll(t)
    assert fin2 and last["step"] == 2
    ctx.release(t)
    with pytest.raises(capi.FfrError):
        ctx.poll(t)
    # the scene is usable alone afterwards
    assert d.step()["step"] == 3
    d.close()


@pytest.mark.parametrize("gangs", [1, 3])
def test_gangs_equal_alone_bitwise(ctx, gangs):
    """Merged launches over groups of scenes (ffr_serve_gangs): every mixed
    scene's fields and results equal its own-stream run bit for bit, with
    scenes of different sizes (different hierarchies and launch sequences)
    sharing launches; the FP64 scene keeps its own stream."""
    cfgs = _scenes()
    base = dict(dt=0.005, max_steps=K, convergence_tol=0.0, advection="quick", buoyancy=True,
                beta=3.4e-3, alpha=2.1e-5, T_initial=26.0)
    cfgs += [random_scene(105, grid=(64, 48, 30), n_fans=2, n_obstacles=4, precision="mixed", **base),
             random_scene(106, grid=(17, 45, 21), n_fans=1, n_obstacles=1, precision="mixed", **base)]
    alone = [_alone(ctx, c, K, tuned=True) for c in cfgs]
    srv = Server(ctx, slots=len(cfgs), keep_fields=True, gangs=gangs)  # counts start at 0
    done = srv.run([Request(c, K) for c in cfgs], realtime=False)
    st = ctx.gang_stats()
    ctx.serve_gangs(0)
    for (fa, la), c in zip(alone, sorted(done, key=lambda c: c.index)):
        assert c.steps == K
        for k in ("u", "v", "w", "p", "T"):
            assert np.array_equal(fa[k], c.fields[k]), (c.index, k)
        for k in ("velocity_change", "pressure_residual", "divergence_max", "T_mean",
                  "poisson_iters", "outer_iters"):
            assert la[k] == c.last[k], (c.index, k)
    rounds, launches, ops = st["rounds"], st["launches"], st["ops"]
    assert rounds > 0
    # five mixed scenes: fewer launches than their own streams would issue
    # (all five share launches with one gang; 2 + 2 + 1 with three)
    assert launches < (0.6 if gangs == 1 else 0.9) * ops, (launches, ops)


@pytest.mark.parametrize("gangs", [0, 1])
def test_stretched_scenes_served_equal_alone_bitwise(ctx, gangs):
    """Stretched rooms (the volume-weighted level-0 kernels) beside uniform
    ones, on their own streams and in one gang: each equals its alone run."""
    base = dict(dt=0.005, max_steps=K, convergence_tol=0.0, advection="quick", buoyancy=True,
                beta=3.4e-3, alpha=2.1e-5, T_initial=26.0)
    cfgs = [random_scene(107, grid=(40, 32, 20), n_fans=2, n_obstacles=3, precision="mixed", **base),
            random_scene(108, grid=(40, 32, 20), n_fans=1, n_obstacles=2, precision="mixed", **base),
            random_scene(109, grid=(24, 36, 18), n_fans=1, n_obstacles=1, precision="mixed", **base)]
    cfgs[0].room.stretch = 1.04
    cfgs[2].room.stretch = 1.06
    alone = [_alone(ctx, c, K, tuned=True) for c in cfgs]
    srv = Server(ctx, slots=len(cfgs), keep_fields=True, gangs=gangs)
    done = srv.run([Request(c, K) for c in cfgs], realtime=False)
    ctx.serve_gangs(0)
    for (fa, la), c in zip(alone, sorted(done, key=lambda c: c.index)):
        assert c.steps == K
        for k in ("u", "v", "w", "p", "T"):
            assert np.array_equal(fa[k], c.fields[k]), (c.index, k)
        assert la["pressure_residual"] == c.last["pressure_residual"], c.index
        assert not c.last["nonfinite"]
