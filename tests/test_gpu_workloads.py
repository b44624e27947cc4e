"""The batched (c4) and serving (c5) workloads themselves, on the GPU.

* c4: the first 8 rooms of the seeded 64-room batch (grids 39..111 x
  37..121 x 17..58, 1-3 fans, 0-5 boxes), 5 steps each, all admitted at
  once: every room's fields and step results in the batch equal its alone
  run bit for bit; then one step of each room from its batched state in the
  mixed form and in the all-FP64 form (CG to 1e-12) differ by relative L2
  <= 1e-6 in u, v, w and in p (p gauge-free: on a pocket of unknown cells
  that touches no OUTFLOW cell the operator is singular and each form's
  constant offset there is removed).
* c5: the first 10 requests of the seeded 50 req/s stream (grids up to
  148 x 196 x 38, each <= 200 steps or until max|du|/dt < 1e-4, the
  FluidSolver::run test), served with continuous batching: each request's
  step count, last result and fields equal its alone run (ffr_run) bit for
  bit.
"""
from types import SimpleNamespace

import numpy as np
import pytest

from ff_room import autotune, capi, device
from ff_room.generator import workload_c4, workload_c5
from ff_room.serve import Request, Server, summary

pytestmark = pytest.mark.gpu

MIXED = dict(precision=capi.PREC_MIXED, poisson_method=capi.POISSON_MGPCG, poisson_tol=1e-10,
             poisson_max_iter=500)


def _prepared(cfg, **overrides):
    ps = device.prepare(cfg, on_device=True, **overrides)
    entry = autotune.lookup(ps.nx, ps.ny, ps.nz, ps.params.precision)
    if entry:  # the table entry of the scene's size class, as the Server applies it
        autotune.apply(ps.params, {k: v for k, v in entry.items() if k != "max_blocks"})
    return ps


def _alone(ctx, cfg, steps):
    d = device.create(ctx, _prepared(cfg, **MIXED))
    last = d.run(steps)
    f = d.download()
    d.close()
    return f, last


def _check_same(alone, done, keys=("velocity_change", "pressure_residual", "divergence_max",
                                   "T_mean", "poisson_iters", "outer_iters", "step")):
    for (fa, la), c in zip(alone, done):
        assert c.steps == la["step"], c.index
        for k in ("u", "v", "w", "p", "T"):
            assert np.array_equal(fa[k], c.fields[k]), (c.index, k)
        for k in keys:
            assert la[k] == c.last[k], (c.index, k)


def test_c4_rooms_batch_equals_alone_and_meets_fp64(ctx):
    from bench import gauge_free_error, sealed_components
    steps = 5
    cfgs = workload_c4(n_scenes=8, steps=steps, precision="mixed")
    alone = [_alone(ctx, c, steps) for c in cfgs]
    srv = Server(ctx, slots=6, keep_fields=True, **MIXED)
    done = srv.run([Request(c, steps) for c in cfgs], realtime=False)
    _check_same(alone, done)
    assert summary(done)["steps"] == 8 * steps
    worst = {k: 0.0 for k in "uvwp"}
    for (fa, _), cfg in zip(alone, cfgs):
        ps = _prepared(cfg, **MIXED)
        p64 = device.prepare(cfg, on_device=True, precision=capi.PREC_FP64,
                             poisson_method=capi.POISSON_CG, poisson_tol=1e-12,
                             poisson_max_iter=100000)
        dm, d64 = device.create(ctx, ps), device.create(ctx, p64)
        for d in (dm, d64):
            d.upload(u=fa["u"], v=fa["v"], w=fa["w"], p=fa["p"], T=fa["T"])
            d.step()
        got, ref = dm.download(), d64.download()
        flags = dm.cell_type()
        dm.close()
        d64.close()
        view = SimpleNamespace(flags=flags, nx=ps.nx, ny=ps.ny, nz=ps.nz)
        worst["p"] = max(worst["p"], gauge_free_error(got["p"].copy(), ref["p"], view,
                                                      sealed_components(view)))
        for k in "uvw":
            worst[k] = max(worst[k], float(np.linalg.norm(got[k] - ref[k]) /
                                           np.linalg.norm(ref[k])))
    for k, e in worst.items():
        assert e <= 1e-6, (k, e)


def test_c5_stream_requests_equal_alone(ctx):
    stream = list(workload_c5(n_requests=10, precision="mixed"))
    alone = [_alone(ctx, c, c.solver.max_steps) for _, c in stream]
    srv = Server(ctx, slots=6, keep_fields=True, **MIXED)
    done = srv.run([Request(c, c.solver.max_steps, t) for t, c in stream], realtime=False)
    assert [c.index for c in done] == list(range(10))
    _check_same(alone, done)
    for c in done:  # each ran to its budget or stopped on FluidSolver::run's test
        assert c.steps == 200 or c.last["converged"] == 1
        assert not c.last["nonfinite"]
