"""D3Q19 lattice Boltzmann on the GPU (ffr_lbm_attach / ffr_lbm_step,
LBMSolver behind the reference's interface).

Pins, bit for bit: every step's result and the fields against the oracle's
restatement (itself pinned to the reference's LBMSolver by
test_oracle_lbm.py) over 25 steps on a room with two openings, basic_room
(fan, automatic outflow, obstacle) and a Lax-Wendroff room with a ceiling
outflow; the golden fixture written by the reference; SolverBridge with
method 'lbm' (pybind LBMSolver) against the oracle; an FP32 scene is
refused.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from ff_room import SceneConfig, SolverBridge, capi
from oracle.pyoracle import OracleLBM
from oracle.scene_flags import build_scene
from scenes import BASIC_ROOM, SMALL_LW, SMALL_THERMAL, stretched, variant

pytestmark = pytest.mark.gpu

CASES = {"thermal": (SMALL_THERMAL, 0.8, 0.01), "basic": (BASIC_ROOM, 0.6, 0.01),
         "lw": (SMALL_LW, 0.7, 0.005)}


def _device(ctx, sc, tau, dt, precision=capi.PREC_FP64):
    dxv, dyv, dzv = sc.spacings()
    outs = [tuple(o[:6]) for o in sc.openings] + list(sc.outflows)  # BoundaryManager::outflows()
    prm = capi.default_params(dt=dt, precision=precision, poisson_method=capi.POISSON_CG)
    d = capi.DeviceScene(ctx, sc.nx, sc.ny, sc.nz, dxv, dyv, dzv, sc.flags, sc.fans, outs,
                         sc.openings, prm)
    return d


def _res(r):
    return np.array([r["step"], r["velocity_change"], r["pressure_residual"],
                     r["divergence_max"], r["converged"], r["T_mean"]])


@pytest.mark.parametrize("tag", sorted(CASES))
def test_lattice_equals_oracle(ctx, tag):
    cfg, tau, dt = CASES[tag]
    sc = build_scene(cfg)
    d = _device(ctx, sc, tau, dt)
    d.lbm_attach(tau, dt, 0.0, sc.lx / sc.nx)
    o = OracleLBM(sc, tau, dt, 0.0)
    for _ in range(25):
        rd, ro = d.lbm_step(), o.step()
        assert np.array_equal(_res(rd), _res(ro)), (rd, ro)
    fd, fo = d.download(), o.fields()
    for k in ("u", "v", "w", "p"):
        assert np.array_equal(fd[k], fo[k]), k
    d.close()


def test_lattice_equals_reference_fixture(ctx):
    g = np.load(os.path.join(GOLDEN, "lbm_small.npz"), allow_pickle=False)
    for tag in ("thermal", "basic"):
        cfg, tau, dt = CASES[tag]
        sc = build_scene(cfg)
        d = _device(ctx, sc, tau, dt)
        d.lbm_attach(tau, dt, 0.0, sc.lx / sc.nx)
        res = np.array([_res(d.lbm_step()) for _ in range(len(g[f"{tag}_results"]))])
        assert np.array_equal(res, g[f"{tag}_results"]), tag
        f = d.download()
        for k in ("u", "v", "w", "p"):
            assert np.array_equal(f[k], g[f"{tag}_{k}"]), (tag, k)
        d.close()


def test_solver_bridge_lbm_equals_oracle():
    d = variant(SMALL_THERMAL, max_steps=12, convergence_tol=0.0)
    d["solver"]["method"] = "lbm"
    d["solver"]["tau"] = 0.8
    cfg = SceneConfig.from_dict(d)
    res = SolverBridge(cfg).run(print_interval=0)
    o = OracleLBM(build_scene(SMALL_THERMAL), 0.8, cfg.solver.dt, 0.0)
    for _ in range(12):
        ro = o.step()
    fo = o.fields()
    assert res.steps == 12 and res.velocity_change == ro["velocity_change"]
    assert res.divergence_max == ro["divergence_max"]
    assert np.array_equal(res.u_field.reshape(-1), fo["u"])
    assert np.array_equal(res.p_field.reshape(-1), fo["p"])


def test_lattice_needs_fp64(ctx):
    cfg, tau, dt = CASES["basic"]
    sc = build_scene(cfg)
    d = _device(ctx, sc, tau, dt, precision=capi.PREC_MIXED)
    with pytest.raises(capi.FfrError):
        d.lbm_attach(tau, dt, 0.0, sc.lx / sc.nx)
    d.close()


def _edge_scene(fan_at_outflow: bool):
    """12x10x8 room: an outflow wall at x = max, a SOLID box, an INFLOW cell
    that no fan box holds (it streams but neither collides nor takes an
    equilibrium), and a fan either inside the room or on the cells the
    outflow copies from."""
    from oracle.pyoracle import INFLOW, OUTFLOW, SOLID, SceneArrays
    nx, ny, nz = 12, 10, 8
    f = np.zeros((nx, ny, nz), dtype=np.uint8)
    f[nx - 1] = OUTFLOW
    f[4:6, 2:4, 0:3] = SOLID
    f[2, 7, 5] = INFLOW
    i = nx - 2 if fan_at_outflow else 3
    f[i, 3:6, 2:5] = INFLOW
    fans = [(i, i, 3, 5, 2, 4, 1.5, 0.2, 0.0)]
    return SceneArrays(nx, ny, nz, 2.4, 2.0, 1.6, f.reshape(-1), fans, [(0, 1, 0, ny - 1, 0, nz - 1)])


@pytest.mark.parametrize("fan_at_outflow", [False, True])
def test_lattice_fused_and_separate_forms(ctx, fan_at_outflow):
    """The streaming pass applies the next step's inflow and collision unless
    a patch copies from an INFLOW cell; either form equals the oracle bit for
    bit, and the profile shows which ran (a separate relaxation pass only in
    the second)."""
    sc = _edge_scene(fan_at_outflow)
    tau, dt = 0.7, 0.01
    d = _device(ctx, sc, tau, dt)
    d.lbm_attach(tau, dt, 0.0, sc.lx / sc.nx)
    o = OracleLBM(sc, tau, dt, 0.0)
    ctx.prof_reset()
    ctx.prof_enable(True)
    for _ in range(15):
        rd, ro = d.lbm_step(), o.step()
        assert np.array_equal(_res(rd), _res(ro)), (rd, ro)
    table = ctx.prof_read()
    ctx.prof_enable(False)
    assert ("lbm_relax" in table) == fan_at_outflow, sorted(table)
    fd, fo = d.download(), o.fields()
    for k in ("u", "v", "w", "p"):
        assert np.array_equal(fd[k], fo[k]), k
    d.close()


def test_lattice_on_stretched_grid_equals_oracle(ctx):
    """A stretched scene: the lattice steps as on a uniform one, and the
    divergence it reports divides by the stretched widths (LBMSolver.cpp:390-392)."""
    sc = stretched(build_scene(SMALL_THERMAL), 1.15)
    tau, dt = 0.8, 0.01
    d = _device(ctx, sc, tau, dt)
    d.lbm_attach(tau, dt, 0.0, sc.lx / sc.nx)
    o = OracleLBM(sc, tau, dt, 0.0)
    for _ in range(15):
        rd, ro = d.lbm_step(), o.step()
        assert np.array_equal(_res(rd), _res(ro)), (rd, ro)
    fd, fo = d.download(), o.fields()
    for k in ("u", "v", "w", "p"):
        assert np.array_equal(fd[k], fo[k]), k
    d.close()
