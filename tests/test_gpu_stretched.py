"""Stretched grids: Grid::set_face_coords faces (reference scene.py:34-59,
82-87; Grid.cpp:39-47) through the step's variable-spacing stencils and the
pressure solve.

Pins: the oracle, bit for bit against the compiled reference on stretched
grids (test_oracle.py); here the FP64 device kernels bit for bit against the
oracle; the multigrid-preconditioned solve, whose inner FP32 solve runs on
the volume-weighted symmetric form of the operator, against a sparse direct
solve of the reference's own non-symmetric system (PoissonSolver.cpp:22-91);
the mixed-precision step against the FP64 step. Tolerances are stated per
test.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from ff_room import capi
from helpers import capi_params, device_scene, rel_l2
from oracle.pyoracle import OUTFLOW, SOLID, OracleSolver, SceneArrays
from oracle.scene_flags import build_scene
from scenes import BASIC_ROOM, SMALL_LW, SMALL_THERMAL, solver_params, stretched

pytestmark = pytest.mark.gpu


def _rhs(sc, seed=3):
    b = np.random.default_rng(seed).standard_normal(sc.nx * sc.ny * sc.nz)
    b[sc.flags == SOLID] = 0.0
    return b


def _sparse(sc):
    """The reference operator as a sparse matrix, column by column from the
    oracle's A x (SOLID rows are identity rows there)."""
    o = OracleSolver(sc, {}, init=False)
    n = sc.nx * sc.ny * sc.nz
    rows, cols, vals = [], [], []
    e = np.zeros(n)
    for c in range(n):
        e[c] = 1.0
        col = o.laplacian(e)
        e[c] = 0.0
        nz = np.flatnonzero(col)
        rows.append(nz)
        cols.append(np.full(nz.size, c))
        vals.append(col[nz])
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


def _random_state(sc, seed, T0=20.0):
    rng = np.random.default_rng(seed)
    sh = sc.shapes
    return dict(u=rng.uniform(-2, 2, sh["u"]), v=rng.uniform(-2, 2, sh["v"]),
                w=rng.uniform(-2, 2, sh["w"]), p=np.zeros(sh["p"]),
                T=T0 + rng.uniform(-5, 5, sh["T"]))


@pytest.mark.parametrize("cfg,ratio", [(BASIC_ROOM, 1.08), (SMALL_THERMAL, 1.2), (SMALL_LW, 1.1)])
def test_laplacian_bitwise_stretched(ctx, cfg, ratio):
    sc = stretched(build_scene(cfg), ratio)
    o = OracleSolver(sc, {}, init=False)
    d = device_scene(ctx, sc, capi_params({}))
    x = np.random.default_rng(1).standard_normal(sc.nx * sc.ny * sc.nz)
    assert np.array_equal(d.laplacian(x), o.laplacian(x))


@pytest.mark.parametrize("cfg,ratio,extra", [
    (BASIC_ROOM, 1.08, dict(advection=0)),
    (BASIC_ROOM, 1.15, dict(advection=1)),
    (SMALL_THERMAL, 1.2, dict()),            # QUICK + buoyancy + openings
    (SMALL_THERMAL, 1.1, dict(advection=2)),
    (SMALL_LW, 1.1, dict(thermal=1, buoyancy=1)),
])
def test_step_without_pressure_solve_is_bitwise_stretched(ctx, cfg, ratio, extra):
    """FP64 step with the pressure solve cut to zero iterations: every
    variable-spacing term (FluidSolver.cpp:98-100, 178-180, 245-247, 325-327,
    343-367, 407-409, 478-480) equals the oracle bit for bit."""
    sc = stretched(build_scene(cfg), ratio)
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


def test_cg_stretched_matches_oracle(ctx):
    """The reference CG on the non-symmetric operator: the same iteration on
    the device, dot products summed in another order (tolerance 1e-7; the
    iteration count within 2 %)."""
    sc = stretched(build_scene(BASIC_ROOM), 1.08)
    o = OracleSolver(sc, {}, init=False)
    d = device_scene(ctx, sc, capi_params({}, poisson_tol=1e-10, poisson_max_iter=20000))
    b = _rhs(sc, 9)
    p_gpu, info = d.poisson(b)
    res, it = o.poisson_cg(b, 20000, 1e-10)
    assert rel_l2(p_gpu, o.p) < 1e-7
    assert abs(info["iterations"] - it) <= max(3, it // 50)
    assert info["residual"] < 1e-10


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_MIXED])
@pytest.mark.parametrize("cfg,ratio", [(BASIC_ROOM, 1.08), (BASIC_ROOM, 1.15),
                                       (SMALL_THERMAL, 1.2)])
def test_mgpcg_stretched_solves_the_reference_system(ctx, cfg, ratio, precision):
    """The outer FP64 residual is the reference's non-symmetric row; the inner
    FP32 solve the symmetric volume-weighted form: the solution equals the
    sparse direct solve of the reference system to 1e-9 (relative L2). At
    ratio 1.15 the reference's own CG stalls near 1e-4 on this scene (oracle,
    20000 iterations); this solve converges in a few dozen."""
    sc = stretched(build_scene(cfg), ratio)
    A = _sparse(sc)
    b = _rhs(sc, 5)
    x = spla.spsolve(A, b)
    d = device_scene(ctx, sc, capi_params({}, poisson_method=capi.POISSON_MGPCG, poisson_tol=1e-11,
                                          precision=precision, poisson_max_iter=500))
    p, info = d.poisson(b)
    assert info["residual"] < 1e-11
    assert rel_l2(p, x) < 1e-9
    assert info["iterations"] <= 120, info


@pytest.mark.parametrize("coarse", [8, 0])
@pytest.mark.parametrize("dims,size", [((9, 7, 6), (1.8, 1.4, 1.2)),      # one level (<= 512 cells)
                                       ((16, 12, 14), (4.0, 3.0, 0.7)),   # semi-coarsened
                                       ((11, 9, 13), (1.0, 1.0, 1.0))])
def test_mgpcg_stretched_random_flags(ctx, dims, size, coarse):
    """Random SOLID / INFLOW cells and an OUTFLOW plane on a stretched grid
    (ratio 1.25), every hierarchy shape: equals the dense solve."""
    rng = np.random.default_rng(7)
    nx, ny, nz = dims
    f = rng.choice([0] * 8 + [SOLID, 2], size=(nx, ny, nz)).astype(np.uint8)
    f[:, -1, :] = OUTFLOW
    sc = stretched(SceneArrays(nx, ny, nz, *size, f.reshape(-1)), 1.25)
    A = _sparse(sc).toarray()
    b = _rhs(sc, 6)
    x = np.linalg.solve(A, b)
    d = device_scene(ctx, sc, capi_params({}, poisson_method=capi.POISSON_MGPCG, poisson_tol=1e-11,
                                          precision=capi.PREC_MIXED, poisson_max_iter=500,
                                          mg_coarse_sweeps=coarse))
    p, info = d.poisson(b)
    assert info["residual"] < 1e-11
    assert rel_l2(p, x) < 1e-9


def test_mixed_stretched_step_tracks_fp64_step(ctx):
    """FP32 stencils + mixed solve against the FP64 step (CG to 1e-10) on a
    stretched grid: relative L2 of u, v, w, T after 20 steps within 1e-3."""
    cfg = SMALL_THERMAL
    sc = stretched(build_scene(cfg), 1.2)
    params = solver_params(cfg)
    params["convergence_tol"] = 0.0
    o = OracleSolver(sc, params)
    d64 = device_scene(ctx, sc, capi_params(params, poisson_tol=1e-10, poisson_max_iter=20000))
    dmx = device_scene(ctx, sc, capi_params(params, precision=capi.PREC_MIXED, poisson_tol=1e-10,
                                            poisson_max_iter=500))
    for d in (d64, dmx):
        d.upload(T=o.T)
    for _ in range(20):
        r64, rmx = d64.step(), dmx.step()
    f64, fmx = d64.download(), dmx.download()
    for k in ("u", "v", "w", "T"):
        assert rel_l2(fmx[k], f64[k]) < 1e-3, (k, rel_l2(fmx[k], f64[k]))
    assert rmx["step"] == 20 and not rmx["nonfinite"]


def test_mixed_stretched_pressure_error_meets_1e6(ctx):
    """The pressure criterion on a stretched reduced c2 room (ratio 1.02 per
    axis): the mixed solve of a step's right-hand side is within relative L2
    1e-6 (non-SOLID cells) of the sparse direct solution."""
    from ff_room import device
    from ff_room.generator import workload_c2
    cfg = workload_c2(grid=(40, 32, 16), seed=2)
    cfg.room.stretch = 1.02
    ps = device.prepare(cfg)
    assert ps.dxv.max() / ps.dxv.min() > 1.3
    pm = capi.default_params(**{k: getattr(ps.params, k) for k, _ in capi.Params._fields_})
    pm.precision, pm.poisson_method, pm.poisson_tol, pm.poisson_max_iter = 1, 2, 1e-10, 500
    dm = device.create(ctx, ps, pm)
    for _ in range(3):
        dm.step()
    f = dm.download()
    b = dm.laplacian(f["p"])  # a realistic right-hand side: A p of a step's pressure
    b[ps.flags == SOLID] = 0.0
    x, info = dm.poisson(b)
    sa = SceneArrays(ps.nx, ps.ny, ps.nz, float(ps.dxv.sum()), float(ps.dyv.sum()),
                     float(ps.dzv.sum()), ps.flags)
    sa.faces = tuple(np.concatenate([[0.0], np.cumsum(h)]) for h in (ps.dxv, ps.dyv, ps.dzv))
    x_ref = spla.spsolve(_sparse(sa), b)
    m = ps.flags != SOLID
    assert rel_l2(x[m], x_ref[m]) < 1e-6
    assert info["residual"] < 1e-10
