"""The context's cache of scene blocks (device arenas, pinned mirrors,
streams, events), which lets a serving loop end requests without a
device-wide hipFree.

Pins: a scene built on blocks a destroyed scene left (full of that scene's
state) steps to the same fields, bit for bit, as the same scene on freshly
allocated blocks, in the mixed MG-PCG and the FP64 CG forms; a destroyed
scene's device bytes leave `in_use` and show as cached, a new scene of the
same size takes them back, and ffr_ctx_trim frees them; a scene that would
not fit the budget beside the cached blocks evicts enough of them and is
created, and one above the budget is still refused.
"""
import numpy as np
import pytest

from ff_room import capi, device
from ff_room.generator import random_scene

pytestmark = pytest.mark.gpu

BASE = dict(dt=0.005, max_steps=4, convergence_tol=0.0, advection="quick", buoyancy=True,
            beta=3.4e-3, alpha=2.1e-5, T_initial=26.0)


def _run(ctx, cfg, steps=4):
    d = device.create(ctx, device.prepare(cfg))
    for _ in range(steps):
        last = d.step()
    f = d.download()
    d.close()
    return f, last


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_reused_blocks_give_the_fresh_result(precision):
    cfg = random_scene(211, grid=(37, 29, 18), n_fans=2, n_obstacles=2, precision=precision, **BASE)
    ctx = capi.Context(0)
    try:
        fresh, lf = _run(ctx, cfg)  # fresh blocks; they hold this run's state when cached
        assert ctx.bytes_in_use() == 0
        cached = ctx.bytes_cached()
        assert cached > 0
        again, la = _run(ctx, cfg)
        assert ctx.bytes_cached() == cached  # taken back and returned, nothing new allocated
        for k in ("u", "v", "w", "p", "T"):
            assert np.array_equal(fresh[k], again[k]), k
        for k in ("velocity_change", "pressure_residual", "divergence_max", "poisson_iters"):
            assert lf[k] == la[k], k
        assert ctx.trim() == cached
        assert ctx.bytes_cached() == 0
    finally:
        ctx.close()


def test_budget_evicts_cached_blocks():
    small = random_scene(212, grid=(40, 30, 20), n_fans=1, n_obstacles=1, precision="mixed", **BASE)
    large = random_scene(213, grid=(64, 48, 24), n_fans=1, n_obstacles=1, precision="mixed", **BASE)
    fs = device.prepare(small).footprint()
    fl = device.prepare(large).footprint()
    assert fl > 1.5 * fs
    budget = fl + fs // 2  # the large scene fits, but not beside the cached small one
    ctx = capi.Context(0, budget)
    try:
        _run(ctx, small, steps=1)
        assert ctx.bytes_in_use() == 0 and ctx.bytes_cached() >= fs
        d = device.create(ctx, device.prepare(large))
        # some of the small scene's blocks went; what stays fits the budget beside the large
        assert ctx.bytes_cached() < fs
        assert fl <= ctx.bytes_in_use() and ctx.bytes_in_use() + ctx.bytes_cached() <= budget
        d.step()
        d.close()
        with pytest.raises(capi.FfrError):  # the budget itself still holds
            device.create(ctx, device.prepare(random_scene(
                214, grid=(80, 60, 30), n_fans=1, n_obstacles=1, precision="mixed", **BASE)))
        assert ctx.bytes_in_use() == 0
    finally:
        ctx.close()
