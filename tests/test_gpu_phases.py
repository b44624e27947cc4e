"""Single phases of the step and the reference caller's contract, on the GPU.

* The right-hand side (FluidSolver::build_poisson_rhs, FluidSolver.cpp:311-330)
  and the projection with its post-solve masks (apply_correction,
  FluidSolver.cpp:332-371, then apply_noslip / apply_inflow / apply_outflow as
  step() calls them, :511-515), each run through ffr_build_rhs /
  ffr_apply_correction on random u*, v*, w* and a random non-zero pressure,
  against the oracle's own phases: bit for bit in the FP64 form (the kernels
  do the reference's operations in its order, contraction off), within FP32
  rounding in the mixed form.
* The reference caller (SolverBridge, solver_bridge.py:93-170, through the
  pybind module and FluidSolver's constructor, FluidSolver.cpp:18-32) on a
  thermal scene with openings and buoyancy: the constructor's initial T bit
  for bit, then a 20-step run against the oracle.
* Multigrid settings changed through ffr_scene_set_params rebuild the
  hierarchy: the solve equals a scene created with the new settings.
* Field views held across steps stay current, and writes through them reach
  the next step (the reference's views are of its live vectors).
"""
import numpy as np
import pytest

from helpers import capi_params, device_scene, rel_l2
from oracle.pyoracle import SOLID, OracleSolver
from oracle.scene_flags import build_scene
from scenes import BASIC_ROOM, SMALL_LW, SMALL_THERMAL, solver_params, stretched, variant

pytestmark = pytest.mark.gpu

SCENES = [
    ("basic", BASIC_ROOM, 1.0),
    ("thermal", SMALL_THERMAL, 1.0),
    ("lw", SMALL_LW, 1.0),
    ("thermal_stretched", SMALL_THERMAL, 1.15),
]


def _scene(cfg, ratio):
    sc = build_scene(cfg)
    return stretched(sc, ratio) if ratio != 1.0 else sc


def _velocity(sc, seed):
    rng = np.random.default_rng(seed)
    sh = sc.shapes
    return dict(u=rng.uniform(-2, 2, sh["u"]), v=rng.uniform(-2, 2, sh["v"]),
                w=rng.uniform(-2, 2, sh["w"]))


def _pressure(sc, seed):
    p = np.random.default_rng(seed).standard_normal(sc.nx * sc.ny * sc.nz) * 3.0
    p[sc.flags == SOLID] = 0.0
    return p


@pytest.mark.parametrize("name,cfg,ratio", SCENES)
def test_rhs_bitwise(ctx, name, cfg, ratio):
    sc = _scene(cfg, ratio)
    params = solver_params(cfg)
    vel = _velocity(sc, 21)
    o = OracleSolver(sc, params, init=False)
    o.set_fields(**vel)
    d = device_scene(ctx, sc, capi_params(params))
    d.upload(**vel)
    b = d.build_rhs()
    ref = o.rhs()
    assert np.array_equal(b, ref)
    assert np.count_nonzero(ref) > 0.5 * ref.size  # a non-trivial right-hand side


@pytest.mark.parametrize("name,cfg,ratio", SCENES)
def test_correction_bitwise(ctx, name, cfg, ratio):
    """u* random, p random and non-zero: the corrected faces, the masks and
    the outflow copies equal the oracle's phases bit for bit."""
    sc = _scene(cfg, ratio)
    params = solver_params(cfg)
    vel, p = _velocity(sc, 22), _pressure(sc, 23)
    o = OracleSolver(sc, params, init=False)
    o.set_fields(p=p, **vel)
    for fn in ("orc_correction", "orc_apply_noslip", "orc_apply_inflow", "orc_apply_outflow"):
        o.call(fn)
    d = device_scene(ctx, sc, capi_params(params))
    d.upload(**vel)
    dv = d.apply_correction(p)
    f = d.download()
    for k in ("u", "v", "w"):
        assert np.array_equal(f[k], getattr(o, k)), k
        assert not np.array_equal(f[k], vel[k]), k  # the correction changed the faces
    assert np.array_equal(f["p"], p)
    want = max(float(np.max(np.abs(getattr(o, k) - vel[k]))) for k in ("u", "v", "w"))
    assert dv == want


@pytest.mark.parametrize("name,cfg,ratio", SCENES[:2])
def test_mixed_phases_within_fp32_rounding(ctx, name, cfg, ratio):
    """The mixed form's phases take FP32 faces and multiply by rounded
    reciprocals: against the oracle on the same FP32-representable inputs,
    relative L2 <= 1e-6 (a few FP32 ulps of the stencil terms)."""
    from ff_room import capi
    sc = _scene(cfg, ratio)
    params = solver_params(cfg)
    vel = {k: v.astype(np.float32).astype(np.float64) for k, v in _velocity(sc, 24).items()}
    p = _pressure(sc, 25)
    d = device_scene(ctx, sc, capi_params(params, precision=capi.PREC_MIXED,
                                          poisson_method=capi.POISSON_MGPCG))
    d.upload(**vel)
    b = d.build_rhs()
    o = OracleSolver(sc, params, init=False)
    o.set_fields(**vel)
    assert rel_l2(b, o.rhs()) < 1e-6
    d.apply_correction(p)
    f = d.download()
    o.set_fields(p=p)
    for fn in ("orc_correction", "orc_apply_noslip", "orc_apply_inflow", "orc_apply_outflow"):
        o.call(fn)
    for k in ("u", "v", "w"):
        assert rel_l2(f[k], getattr(o, k)) < 1e-6, k


def test_solver_bridge_thermal_from_constructor(ctx):
    """SMALL_THERMAL (QUICK, buoyancy, two openings) through SceneConfig ->
    Scene -> FluidSolver (the product constructor sets T) -> run: the initial
    T equals the oracle's constructor bit for bit (T_initial, then the opening
    temperatures, FluidSolver.cpp:27-31), the bridge's beta override
    (1/(T_initial + 273.15), solver_bridge.py:118-124) and thermal rule hold,
    and after 20 steps u, v, w, p, T match the oracle within 1e-5 (the CG
    stops at 1e-6; the GPU's dot order may end it one iteration apart)."""
    from ff_room import SceneConfig, SolverBridge
    from ff_room.scene import Scene, _load_core
    from ff_room.solver_bridge import make_params
    core = _load_core()
    cfg = SceneConfig.from_dict(variant(SMALL_THERMAL, max_steps=20, convergence_tol=0.0))
    prm = make_params(cfg, core)
    assert prm.thermal and prm.buoyancy
    assert prm.beta == 1.0 / (30.0 + 273.15)
    sc = build_scene(SMALL_THERMAL)
    params = solver_params(SMALL_THERMAL)
    params["convergence_tol"] = 0.0
    o = OracleSolver(sc, params)  # orc_init: the reference constructor
    scene = Scene(cfg)
    solver = core.FluidSolver(scene.grid, scene.bm, prm)
    T0 = np.array(scene.grid.T_shaped()).reshape(-1)
    assert np.array_equal(T0, o.T)
    assert np.unique(T0).size == 2  # T_initial inside, T_outside on the opening cells
    del solver
    res = SolverBridge(cfg).run(print_interval=0)
    for _ in range(20):
        ro = o.step()
    assert res.steps == 20
    for k, a in (("u", res.u_field), ("v", res.v_field), ("w", res.w_field),
                 ("p", res.p_field), ("T", res.T_field)):
        assert rel_l2(a.reshape(-1), getattr(o, k)) < 1e-5, k
    assert res.T_mean == pytest.approx(ro["T_mean"], rel=1e-9)
    assert res.velocity_change == pytest.approx(ro["velocity_change"], rel=1e-4)


def test_set_params_rebuilds_multigrid(ctx):
    """mg_coarse_sweeps (and the other hierarchy settings) changed on a scene
    that already has its hierarchy: the next solve equals the solve of a scene
    created with the new settings, bit for bit."""
    from ff_room import capi
    sc = build_scene(variant(BASIC_ROOM))
    base = dict(precision=capi.PREC_MIXED, poisson_method=capi.POISSON_MGPCG, poisson_tol=1e-10,
                poisson_max_iter=50)
    rhs = np.random.default_rng(5).standard_normal(sc.nx * sc.ny * sc.nz)
    rhs[sc.flags == SOLID] = 0.0
    d = device_scene(ctx, sc, capi_params({}, mg_coarse_sweeps=8, **base))
    d.poisson(rhs)  # builds the hierarchy with 8 coarse sweeps
    for new in (dict(mg_coarse_sweeps=2), dict(mg_coarse_sweeps=2, mg_level_sweeps=1),
                dict(mg_coarse_sweeps=0, mg_pre_sweeps=1)):
        prm = capi_params({}, **base, **new)
        d.set_params(prm)
        x1, i1 = d.poisson(rhs)
        fresh = device_scene(ctx, sc, capi_params({}, **base, **new))
        x2, i2 = fresh.poisson(rhs)
        fresh.close()
        assert np.array_equal(x1, x2), new
        assert i1 == i2, new
    d.close()


def test_views_held_across_steps(ctx):
    """A view taken before two steps shows the fields after each step, and a
    write through it is what the next step starts from."""
    from ff_room import SceneConfig
    from ff_room.scene import Scene, _load_core
    from ff_room.solver_bridge import make_params
    core = _load_core()
    cfg = SceneConfig.from_dict(variant(BASIC_ROOM, max_steps=5, convergence_tol=0.0))
    scene = Scene(cfg)
    solver = core.FluidSolver(scene.grid, scene.bm, make_params(cfg, core))
    held_u, held_p = scene.grid.u, scene.grid.p_shaped()
    for _ in range(2):
        solver.step()
        fresh = np.array(scene.grid.u)
        assert np.array_equal(held_u, fresh)
        assert np.array_equal(held_p, np.array(scene.grid.p_shaped()))
    assert np.count_nonzero(held_u) > 0
    # a write through the held view reaches the device before the next step
    twin = Scene(cfg)
    twin_solver = core.FluidSolver(twin.grid, twin.bm, make_params(cfg, core))
    for _ in range(2):
        twin_solver.step()
    edit = np.array(held_u) * 0.5
    held_u[:] = edit
    tu = twin.grid.u
    tu[:] = edit
    del tu
    r1, r2 = solver.step(), twin_solver.step()
    assert r1.velocity_change == r2.velocity_change
    assert np.array_equal(np.array(scene.grid.u), np.array(twin.grid.u))
    del held_u, held_p
