"""Snapshot egress (ffr_snapshot_begin/end) and SolverBridge.run_animated.

Pins: a snapshot holds the fields as they were when it was queued, bit for
bit (the synchronous download of the same state), after later steps have
overwritten the live fields, in both precisions; four slots, then
FFR_E_STATE until one is ended, and ending a free slot is an error;
run_animated's captures (queued, collected later) equal captures taken
synchronously between the steps of the same solver (the reference's
solver_bridge.py:224-245 quantities), and its SimResult equals run()'s.
"""
import numpy as np
import pytest

from ff_room import SceneConfig, SolverBridge, capi, device
from ff_room.generator import random_scene

from scenes import SMALL_THERMAL, variant

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_snapshot_holds_the_fields_at_capture(ctx, precision):
    cfg = random_scene(201, grid=(36, 28, 18), n_fans=2, n_obstacles=2, precision=precision,
                       dt=0.005, max_steps=10, convergence_tol=0.0, advection="quick",
                       buoyancy=True, beta=3.4e-3, alpha=2.1e-5, T_initial=26.0)
    d = device.create(ctx, device.prepare(cfg))
    for _ in range(2):
        d.step()
    slot = d.snapshot_begin()
    ref = d.download()  # the same state, synchronously
    d.step()
    d.step()
    step, snap = d.snapshot_end(slot)
    assert step == 2
    for k in ("u", "v", "w", "p", "T"):
        assert np.array_equal(snap[k], ref[k]), k
    assert not np.array_equal(d.download()["u"], ref["u"])  # the live fields moved on
    d.close()


def test_cell_snapshot_matches_numpy(ctx):
    """FFR_SNAP_CELLS against numpy on the fields of a FFR_SNAP_FIELDS capture
    of the same state (mixed form: FP32 state, FP64 arithmetic)."""
    cfg = random_scene(203, grid=(30, 22, 14), n_fans=2, n_obstacles=2, precision="mixed",
                       dt=0.005, max_steps=10, convergence_tol=0.0, buoyancy=True)
    d = device.create(ctx, device.prepare(cfg))
    for _ in range(3):
        d.step()
    sc, sf = d.snapshot_begin(cells=True), d.snapshot_begin()
    d.step()
    _, c = d.snapshot_end(sc)
    _, f = d.snapshot_end(sf)
    nx, ny, nz = d.nx, d.ny, d.nz
    u = f["u"].reshape(nx + 1, ny, nz)
    v = f["v"].reshape(nx, ny + 1, nz)
    w = f["w"].reshape(nx, ny, nz + 1)
    uc, vc, wc = 0.5 * (u[:-1] + u[1:]), 0.5 * (v[:, :-1] + v[:, 1:]), 0.5 * (w[:, :, :-1] + w[:, :, 1:])
    assert np.array_equal(c["vel"].reshape(nx, ny, nz, 3), np.stack([uc, vc, wc], axis=-1))
    assert np.array_equal(c["speed"].reshape(nx, ny, nz), np.sqrt(uc ** 2 + vc ** 2 + wc ** 2))
    assert np.array_equal(c["T"], f["T"])
    d.close()


def test_snapshot_slots_run_out(ctx):
    cfg = random_scene(202, grid=(20, 16, 12), n_fans=1, n_obstacles=1, precision="mixed",
                       dt=0.005, max_steps=10, convergence_tol=0.0)
    d = device.create(ctx, device.prepare(cfg))
    slots = []
    for i in range(4):
        d.step()
        slots.append(d.snapshot_begin())
    assert sorted(slots) == [0, 1, 2, 3]
    with pytest.raises(capi.FfrError):
        d.snapshot_begin()
    step0, _ = d.snapshot_end(slots[0])
    assert step0 == 1
    with pytest.raises(capi.FfrError):
        d.snapshot_end(slots[0])  # already ended
    assert d.snapshot_begin() == slots[0]
    steps = sorted(d.snapshot_end(s)[0] for s in range(4))
    assert steps == [2, 3, 4, 4]
    d.close()


def _captures_between_steps(cfg, save_every):
    """The reference's capture, taken synchronously between steps."""
    bridge = SolverBridge(cfg)
    scene, params, solver = bridge._solver()
    g = scene.grid
    out = []
    for _ in range(cfg.solver.max_steps):
        r = solver.step()
        if r.step % save_every == 0 or r.step == cfg.solver.max_steps:
            u, v, w = np.array(g.u_shaped()), np.array(g.v_shaped()), np.array(g.w_shaped())
            uc, vc, wc = 0.5 * (u[:-1] + u[1:]), 0.5 * (v[:, :-1] + v[:, 1:]), 0.5 * (w[:, :, :-1] + w[:, :, 1:])
            out.append((r.step, r.T_mean, np.sqrt(uc ** 2 + vc ** 2 + wc ** 2),
                        np.array(g.T_shaped()), np.stack([uc, vc, wc], axis=-1)))
    return out


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_run_animated_equals_synchronous_captures(precision):
    d = variant(SMALL_THERMAL, max_steps=11, convergence_tol=0.0)
    d["solver"]["precision"] = precision
    if precision == "mixed":
        d["solver"]["poisson_method"] = "mgpcg"
    cfg = SceneConfig.from_dict(d)
    res, snaps = SolverBridge(cfg).run_animated(save_every=2, print_interval=0)
    ref = _captures_between_steps(cfg, 2)
    assert [s.step for s in snaps] == [r[0] for r in ref] == [2, 4, 6, 8, 10, 11]
    for s, (step, tmean, speed, T, vel) in zip(snaps, ref):
        assert s.time_s == step * cfg.solver.dt and s.T_mean == tmean
        assert np.array_equal(s.speed, speed) and np.array_equal(s.T_field, T)
        assert np.array_equal(s.vel, vel)
    plain = SolverBridge(cfg).run(print_interval=0)
    assert res.steps == plain.steps == 11
    for k in ("u_field", "v_field", "w_field", "p_field", "T_field"):
        assert np.array_equal(getattr(res, k), getattr(plain, k)), k
