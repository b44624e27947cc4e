"""Parity of the FP64 GPU path (the reference algorithm) through the C ABI.

Bit-exact where the arithmetic is elementwise: the pressure operator, and a
whole step whose pressure solve is cut to zero iterations (every kernel then
does the reference's operations in the reference's order, contraction off).
Where the CG's dot products enter, the GPU sums in a different (fixed) order
than the reference's serial loop, so results agree to rounding, and an
iteration count that straddles the 1e-6 stopping test moves p by O(tol):
those comparisons use the tolerances stated per test.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from helpers import capi_params, device_scene, rel_l2
from oracle.pyoracle import OUTFLOW, SOLID, OracleSolver, SceneArrays
from oracle.scene_flags import build_scene
from scenes import BASIC_ROOM, SMALL_LW, SMALL_THERMAL, solver_params, variant

pytestmark = pytest.mark.gpu


def _golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def _random_state(sc: SceneArrays, seed: int, T0=20.0):
    rng = np.random.default_rng(seed)
    sh = sc.shapes
    return dict(u=rng.uniform(-2, 2, sh["u"]), v=rng.uniform(-2, 2, sh["v"]),
                w=rng.uniform(-2, 2, sh["w"]), p=np.zeros(sh["p"]),
                T=T0 + rng.uniform(-5, 5, sh["T"]))


@pytest.mark.parametrize("cfg", [BASIC_ROOM, SMALL_THERMAL, SMALL_LW])
def test_laplacian_bitwise(ctx, cfg):
    sc = build_scene(cfg)
    o = OracleSolver(sc, {}, init=False)
    d = device_scene(ctx, sc, capi_params({}))
    x = np.random.default_rng(1).standard_normal(sc.nx * sc.ny * sc.nz)
    assert np.array_equal(d.laplacian(x), o.laplacian(x))


CASES = [
    (BASIC_ROOM, dict(advection=0)),
    (BASIC_ROOM, dict(advection=1)),
    (BASIC_ROOM, dict(advection=2)),
    (SMALL_THERMAL, dict()),                       # QUICK + buoyancy + openings
    (SMALL_THERMAL, dict(advection=0)),
    (SMALL_THERMAL, dict(advection=2)),
    (SMALL_LW, dict(thermal=1, buoyancy=1)),
]


@pytest.mark.parametrize("cfg,extra", CASES)
def test_step_without_pressure_solve_is_bitwise(ctx, cfg, extra):
    """poisson_max_iter = 0: p stays 0 (PoissonSolver.cpp:219 loop skipped)."""
    sc = build_scene(cfg)
    params = solver_params(cfg)
    params.update(extra, poisson_max_iter=0, convergence_tol=0.0)
    st = _random_state(sc, 11, params["T_initial"])
    o = OracleSolver(sc, params, init=False)
    o.set_fields(**st)
    d = device_scene(ctx, sc, capi_params(params))
    d.upload(**st)
    for _ in range(3):
        ro, rd = o.step(), d.step()
        f = d.download()
        for k in ("u", "v", "w", "p", "T"):
            assert np.array_equal(f[k], getattr(o, k)), k
        assert rd["velocity_change"] == ro["velocity_change"]
        assert rd["divergence_max"] == ro["divergence_max"]
        assert rd["pressure_residual"] == ro["pressure_residual"] == 1.0
        assert rd["T_mean"] == pytest.approx(ro["T_mean"], rel=1e-14)
        assert rd["step"] == ro["step"]


@pytest.mark.parametrize("cfg", [BASIC_ROOM, SMALL_THERMAL])
def test_cg_solution_matches_oracle(ctx, cfg):
    sc = build_scene(cfg)
    o = OracleSolver(sc, solver_params(cfg), init=False)
    d = device_scene(ctx, sc, capi_params(solver_params(cfg), poisson_tol=1e-10,
                                          poisson_max_iter=5000))
    rng = np.random.default_rng(3)
    b = rng.standard_normal(sc.nx * sc.ny * sc.nz)
    b[sc.flags == SOLID] = 0.0
    p_gpu, info = d.poisson(b)
    res, it = o.poisson_cg(b, 5000, 1e-10)
    # same algorithm, different summation order of the dots: tolerance 1e-8
    assert rel_l2(p_gpu, o.p) < 1e-8
    assert abs(info["iterations"] - it) <= 3
    assert info["residual"] < 1e-10


@pytest.mark.parametrize("name,cfg,steps,tol", [
    ("c1_upwind_cg.npz", BASIC_ROOM, (1, 10, 100, 500), 1e-5),
    ("small_thermal.npz", SMALL_THERMAL, (20,), 1e-5),
    ("small_lw.npz", SMALL_LW, (10,), 1e-5),
])
def test_trajectory_matches_reference_fixture(ctx, name, cfg, steps, tol):
    """u, v, w, p, T against the reference's own output. Tolerance: relative
    L2 <= 1e-5 (the CG stops at ||r||/||b|| < 1e-6 and the GPU's summation
    order can end it one iteration earlier or later)."""
    g = _golden(name)
    sc = build_scene(cfg)
    params = solver_params(cfg)
    params["convergence_tol"] = 0.0
    d = device_scene(ctx, sc, capi_params(params))
    o = OracleSolver(sc, params)  # T initialisation (FluidSolver ctor) done by the oracle
    d.upload(u=o.u, v=o.v, w=o.w, p=o.p, T=o.T)
    for s in range(1, max(steps) + 1):
        r = d.step()
        if s in steps:
            f = d.download()
            for k in ("u", "v", "w", "p", "T"):
                assert rel_l2(f[k], g[f"{k}_{s}"]) < tol, (s, k, rel_l2(f[k], g[f"{k}_{s}"]))
            ref_r = g[f"result_{s}"]
            assert r["step"] == ref_r[0]
            assert r["velocity_change"] == pytest.approx(ref_r[1], rel=1e-4)
            assert r["T_mean"] == pytest.approx(ref_r[5], rel=1e-12)


def test_couette_case_matches_reference_fixture(ctx):
    g = _golden("couette.npz")
    nx, ny, nz = 4, 4, 10
    flags = np.zeros((nx, ny, nz), dtype=np.uint8)
    flags[:, :, nz - 1] = 2
    sc = SceneArrays(nx, ny, nz, 1.0, 1.0, 1.0, flags.reshape(-1),
                     fans=[(0, nx - 1, 0, ny - 1, nz - 1, nz - 1, 1.0, 0.0, 0.0)])
    params = dict(dt=0.005, max_steps=2000, convergence_tol=1e-5, nu=1.5e-5, poisson_tol=1e-7)
    d = device_scene(ctx, sc, capi_params(params))
    d.upload(T=np.full(sc.nx * sc.ny * sc.nz, 20.0))
    r = d.run(2000)
    assert r["step"] == 2000 and not r["converged"]
    f = d.download()
    for k in ("u", "v", "w"):
        assert rel_l2(f[k], g[k]) < 1e-6, k


def test_manufactured_poisson_second_order(ctx):
    from test_oracle import manufactured
    errs = []
    for N in (16, 32):
        sc, rhs, pe, inner = manufactured(N)
        d = device_scene(ctx, sc, capi_params({}, poisson_tol=1e-10))
        p, info = d.poisson(rhs)
        assert info["residual"] < 1e-9
        errs.append(np.max(np.abs(p[inner] - pe[inner])))
    assert errs[0] < 0.01 and errs[0] / errs[1] > 3.5


def test_deterministic(ctx):
    sc = build_scene(SMALL_THERMAL)
    params = solver_params(SMALL_THERMAL)
    outs = []
    for _ in range(2):
        d = device_scene(ctx, sc, capi_params(params))
        d.upload(T=np.full(sc.nx * sc.ny * sc.nz, params["T_initial"]))
        rs = [d.step() for _ in range(5)]
        outs.append((d.download(), rs))
    for k in outs[0][0]:
        assert np.array_equal(outs[0][0][k], outs[1][0][k]), k
    assert outs[0][1] == outs[1][1]


def test_nonfinite_flag_and_small_max_iter(ctx):
    sc = build_scene(BASIC_ROOM)
    params = solver_params(BASIC_ROOM)
    d = device_scene(ctx, sc, capi_params(params, poisson_max_iter=1))
    r = d.step()
    assert r["poisson_iters"] == 1 and r["pressure_residual"] > 1e-6  # not converged, reported
    d2 = device_scene(ctx, sc, capi_params(params, dt=50.0))  # CFL far above 1
    flagged = False
    for _ in range(30):
        r = d2.step()
        if r["nonfinite"]:
            flagged = True
            break
    assert flagged


def test_invalid_scene_rejected(ctx):
    from ff_room import capi
    sc = build_scene(BASIC_ROOM)
    with pytest.raises(capi.FfrError):
        dxv, dyv, dzv = sc.spacings()
        capi.DeviceScene(ctx, sc.nx, sc.ny, sc.nz, dxv, dyv, dzv, sc.flags,
                         outflows=[(0, 1, 0, sc.ny, 0, sc.nz - 1)])  # patch off the wall


def test_solver_bridge_matches_oracle(ctx):
    """The reference-compatible Python path (SolverBridge -> pybind -> C ABI)."""
    from ff_room import SceneConfig, SolverBridge
    cfg = SceneConfig.from_dict(variant(BASIC_ROOM, max_steps=30, convergence_tol=0.0))
    res = SolverBridge(cfg).run(print_interval=0)
    sc = build_scene(BASIC_ROOM)
    params = solver_params(BASIC_ROOM)
    params["convergence_tol"] = 0.0
    o = OracleSolver(sc, params)
    for _ in range(30):
        ro = o.step()
    assert res.steps == 30
    assert rel_l2(res.u_field.reshape(-1), o.u) < 1e-5
    assert rel_l2(res.p_field.reshape(-1), o.p) < 1e-5
    assert res.velocity_change == pytest.approx(ro["velocity_change"], rel=1e-4)
