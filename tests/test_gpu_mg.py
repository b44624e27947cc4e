"""The multigrid-preconditioned pressure solve and the mixed-precision step.

Pins: the dense solution of the same linear system (tiny grids), the FP64
reference CG converged tightly (oracle), the manufactured solution, and the
FP64 GPU step for the mixed-precision step's trajectory error. Tolerances are
stated per test.
"""
import numpy as np
import pytest

from ff_room import capi
from ff_room.generator import random_scene, workload_c2
from helpers import capi_params, device_scene, rel_l2
from oracle.pyoracle import OUTFLOW, SOLID, OracleSolver, SceneArrays
from oracle.scene_flags import build_scene
from scenes import BASIC_ROOM, SMALL_LW, SMALL_THERMAL, solver_params

pytestmark = pytest.mark.gpu


def _rhs(sc, seed=3):
    b = np.random.default_rng(seed).standard_normal(sc.nx * sc.ny * sc.nz)
    b[sc.flags == SOLID] = 0.0
    return b


def _dense(sc):
    o = OracleSolver(sc, {}, init=False)
    n = sc.nx * sc.ny * sc.nz
    A = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        A[:, c] = o.laplacian(e)
    return A


@pytest.mark.parametrize("bf16", [0, 1])
@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_MIXED])
def test_mgpcg_solves_the_reference_system(ctx, bf16, precision):
    """CG semantics (OUTFLOW pressure back-substituted): equals the dense solve."""
    rng = np.random.default_rng(1)
    nx, ny, nz = 9, 7, 6
    f = rng.choice([0, 0, 0, 0, 0, 0, SOLID, 2], size=(nx, ny, nz)).astype(np.uint8)
    f[-1] = OUTFLOW
    sc = SceneArrays(nx, ny, nz, 1.8, 1.4, 1.2, f.reshape(-1))
    A = _dense(sc)
    b = _rhs(sc)
    x = np.linalg.solve(A, b)
    d = device_scene(ctx, sc, capi_params({}, poisson_method=capi.POISSON_MGPCG, poisson_tol=1e-11,
                                          mg_bf16=bf16, precision=precision, poisson_max_iter=500))
    p, info = d.poisson(b)
    assert info["residual"] < 1e-11
    assert rel_l2(p, x) < 1e-9


@pytest.mark.parametrize("coarse", [8, 0])  # sweeps on the coarsest level, or its dense inverse
@pytest.mark.parametrize("dims,size", [((16, 12, 14), (4.0, 3.0, 0.7)),   # z-cells 3.4x thinner
                                       ((6, 20, 9), (3.0, 1.0, 2.4)),    # y strongly coupled
                                       ((11, 9, 13), (1.0, 1.0, 1.0))])  # odd extents, near-cubic
def test_mgpcg_semicoarsened_hierarchy_solves_the_system(ctx, dims, size, coarse):
    """Anisotropic cells coarsen the strongly coupled axes first: the solve
    still equals the dense solution, within few iterations, with the coarsest
    level swept or solved by its dense inverse."""
    rng = np.random.default_rng(11)
    nx, ny, nz = dims
    f = rng.choice([0] * 9 + [SOLID], size=(nx, ny, nz)).astype(np.uint8)
    f[:, :, -1] = OUTFLOW
    sc = SceneArrays(nx, ny, nz, *size, f.reshape(-1))
    A = _dense(sc)
    b = _rhs(sc, 5)
    x = np.linalg.solve(A, b)
    d = device_scene(ctx, sc, capi_params({}, poisson_method=capi.POISSON_MGPCG, poisson_tol=1e-11,
                                          precision=capi.PREC_MIXED, poisson_max_iter=500,
                                          mg_coarse_sweeps=coarse))
    p, info = d.poisson(b)
    assert info["residual"] < 1e-11
    assert rel_l2(p, x) < 1e-9
    assert info["iterations"] <= 60, info


def test_mgpcg_pinned_outflow_is_the_reference_multigrid_system(ctx):
    """use_multigrid: OUTFLOW pressure pinned to 0 (PoissonSolver.cpp:201)."""
    rng = np.random.default_rng(2)
    nx, ny, nz = 8, 6, 5
    f = rng.choice([0, 0, 0, 0, 0, SOLID], size=(nx, ny, nz)).astype(np.uint8)
    f[:, -1, :] = OUTFLOW
    sc = SceneArrays(nx, ny, nz, 1.6, 1.2, 1.0, f.reshape(-1))
    A = _dense(sc)
    b = _rhs(sc, 4)
    out = f.reshape(-1) == OUTFLOW
    A[out] = 0.0
    A[out, out] = 1.0
    bp = b.copy()
    bp[out] = 0.0
    x = np.linalg.solve(A, bp)
    d = device_scene(ctx, sc, capi_params({}, use_multigrid=1, poisson_tol=1e-11,
                                          poisson_max_iter=500))
    p, info = d.poisson(b)
    assert np.all(p[out] == 0.0)
    assert rel_l2(p, x) < 1e-9


@pytest.mark.parametrize("cfg", [BASIC_ROOM, SMALL_THERMAL])
def test_mgpcg_matches_tight_reference_cg(ctx, cfg):
    sc = build_scene(cfg)
    b = _rhs(sc, 9)
    o = OracleSolver(sc, {}, init=False)
    o.poisson_cg(b, 20000, 1e-13)
    d = device_scene(ctx, sc, capi_params({}, poisson_method=capi.POISSON_MGPCG, poisson_tol=1e-12,
                                          precision=capi.PREC_MIXED, poisson_max_iter=500))
    p, info = d.poisson(b)
    nonsolid = sc.flags != SOLID
    assert rel_l2(p[nonsolid], o.p[nonsolid]) < 1e-9
    assert info["outer_iterations"] >= 2  # FP32 inner solves need FP64 correction


def test_mgpcg_manufactured_second_order(ctx):
    from test_oracle import manufactured
    errs = []
    for N in (16, 32):
        sc, rhs, pe, inner = manufactured(N)
        d = device_scene(ctx, sc, capi_params({}, poisson_method=capi.POISSON_MGPCG,
                                              poisson_tol=1e-10, precision=capi.PREC_MIXED))
        p, info = d.poisson(rhs)
        assert info["residual"] < 1e-10
        errs.append(np.max(np.abs(p[inner] - pe[inner])))
    assert errs[0] < 0.01 and errs[0] / errs[1] > 3.5


def _workload_scene(grid, seed=2):
    from ff_room import device
    cfg = workload_c2(grid=grid, seed=seed)
    return device.prepare(cfg)


def test_zero_rhs_gives_zero_pressure_from_any_start(ctx):
    """A step whose right-hand side is 0 (a still room, no fan) returns p = 0,
    as the reference CG does from its zero start (PoissonSolver.cpp:96; a
    norm of b below 1e-30 ends it), whatever pressure the warm start
    carried in: here a random one uploaded before the step."""
    nx, ny, nz = 12, 10, 8
    f = np.zeros((nx, ny, nz), np.uint8)
    f[-1] = OUTFLOW
    f[3:5, 2:6, 0:3] = SOLID
    sc = SceneArrays(nx, ny, nz, 1.2, 1.0, 0.8, f.reshape(-1))
    for warm in (0, 1, 3):
        d = device_scene(ctx, sc, capi_params({}, precision=capi.PREC_MIXED,
                                              poisson_method=capi.POISSON_MGPCG,
                                              poisson_tol=1e-10, poisson_max_iter=500,
                                              warm_start=warm))
        d.upload(p=np.random.default_rng(5).standard_normal(nx * ny * nz))
        for _ in range(2):
            r = d.step()
            out = d.download()
            assert r["outer_iters"] == 0 and r["poisson_iters"] == 0, (warm, r)
            assert np.all(out["p"] == 0.0), warm
            assert all(np.all(out[k] == 0.0) for k in ("u", "v", "w")), warm
        d.close()


def test_warm_start_meets_the_same_tolerance(ctx):
    """A step's mixed pressure solve started from the previous pressure, or
    extrapolated from the last two, meets the same relative residual as one
    started from 0 (the reference CG's start, PoissonSolver.cpp:96), and the
    trajectories agree to the solve's accuracy; a bare solve starts from 0
    either way."""
    from ff_room import device
    ps = _workload_scene((64, 52, 20))
    out = []
    for warm in (0, 1, 2, 3):
        pm = capi.default_params(**{k: getattr(ps.params, k) for k, _ in capi.Params._fields_})
        pm.precision, pm.poisson_method, pm.poisson_tol, pm.poisson_max_iter = 1, 2, 1e-10, 500
        pm.warm_start = warm
        d = device.create(ctx, ps, pm)
        res = [d.step() for _ in range(6)]
        assert all(r["pressure_residual"] < 1e-10 for r in res), (warm, res)
        f = d.download()
        b = d.laplacian(f["p"])
        x, info = d.poisson(b)  # from 0: the previous pressure is this solution
        assert info["iterations"] > 0 and info["residual"] < 1e-10
        out.append(f)
        d.close()
    m = ps.flags != SOLID
    for k in ("u", "v", "w"):
        assert rel_l2(out[1][k], out[0][k]) < 1e-6, k
        assert rel_l2(out[2][k], out[0][k]) < 1e-6, k
        assert rel_l2(out[3][k], out[0][k]) < 1e-6, k
    assert rel_l2(out[1]["p"][m], out[0]["p"][m]) < 1e-6
    assert rel_l2(out[2]["p"][m], out[0]["p"][m]) < 1e-6
    assert rel_l2(out[3]["p"][m], out[0]["p"][m]) < 1e-6


def test_block_cap_beyond_the_resident_grid(ctx):
    """A block cap above one resident round (8192 > 2048 blocks) keeps the
    level-0 gather kernels at the resident grid, and the reductions they
    feed count their partials accordingly: the solve of a 1.3M-cell system
    converges as with the default cap, to the same solution within the
    tolerance of two independently rounded solves."""
    from ff_room import device
    ps = _workload_scene((128, 128, 80))
    rng = np.random.default_rng(7)
    b = rng.standard_normal(ps.cells)
    b[ps.flags == SOLID] = 0.0
    pm = capi.default_params(**{k: getattr(ps.params, k) for k, _ in capi.Params._fields_})
    pm.precision, pm.poisson_method, pm.poisson_tol, pm.poisson_max_iter = 1, 2, 1e-11, 500
    L = capi.lib()
    before = L.ffr_get_max_blocks()
    out = []
    try:
        for cap in (8192, 2048):
            assert L.ffr_set_max_blocks(cap) == 0
            d = device.create(ctx, ps, pm)
            x, info = d.poisson(b)
            d.close()
            assert info["residual"] < 1e-11, (cap, info)
            out.append((x, info["iterations"]))
    finally:
        L.ffr_set_max_blocks(before)
    m = ps.flags != SOLID
    assert rel_l2(out[0][0][m], out[1][0][m]) < 1e-8
    assert abs(out[0][1] - out[1][1]) <= 2


def test_mixed_pressure_error_meets_1e6(ctx):
    """The headline criterion at a reduced c2 grid: the mixed solve's pressure
    on a step's right-hand side is within relative L2 1e-6 (non-SOLID cells)
    of FP64 CG converged to 1e-12."""
    from ff_room import device
    ps = _workload_scene((64, 52, 20))
    p64 = capi.default_params(**{k: getattr(ps.params, k) for k, _ in capi.Params._fields_})
    p64.precision, p64.poisson_method, p64.poisson_tol, p64.poisson_max_iter = 0, 1, 1e-12, 100000
    d = device.create(ctx, ps, p64)
    for _ in range(3):
        d.step()
    f = d.download()
    b = d.laplacian(f["p"])  # a realistic right-hand side: A p of a step's pressure
    b[ps.flags == SOLID] = 0.0
    x_ref, _ = d.poisson(b)
    pm = capi.default_params(**{k: getattr(ps.params, k) for k, _ in capi.Params._fields_})
    pm.precision, pm.poisson_method, pm.poisson_tol, pm.poisson_max_iter = 1, 2, 1e-10, 500
    dm = device.create(ctx, ps, pm)
    x, info = dm.poisson(b)
    m = ps.flags != SOLID
    assert rel_l2(x[m], x_ref[m]) < 1e-6
    assert info["residual"] < 1e-10


@pytest.mark.parametrize("cfg,steps,tol", [(BASIC_ROOM, 20, 1e-3), (SMALL_THERMAL, 20, 1e-3)])
def test_mixed_step_tracks_fp64_step(ctx, cfg, steps, tol):
    """FP32 stencils + mixed pressure solve against the FP64 GPU step:
    relative L2 of u, v, w, T after 20 steps within 1e-3 (FP32 rounding of the
    state accumulating over the steps; the pressure solve itself is at 1e-10)."""
    sc = build_scene(cfg)
    params = solver_params(cfg)
    params["convergence_tol"] = 0.0
    o = OracleSolver(sc, params)
    d64 = device_scene(ctx, sc, capi_params(params, poisson_tol=1e-10, poisson_max_iter=20000))
    dmx = device_scene(ctx, sc, capi_params(params, precision=capi.PREC_MIXED, poisson_tol=1e-10,
                                            poisson_max_iter=500))
    for d in (d64, dmx):
        d.upload(T=o.T)
    for _ in range(steps):
        r64, rmx = d64.step(), dmx.step()
    f64, fmx = d64.download(), dmx.download()
    for k in ("u", "v", "w", "T"):
        assert rel_l2(fmx[k], f64[k]) < tol, (k, rel_l2(fmx[k], f64[k]))
    nonsolid = sc.flags != SOLID
    assert rel_l2(fmx["p"][nonsolid], f64["p"][nonsolid]) < 10 * tol
    assert rmx["step"] == steps and not rmx["nonfinite"]


def test_mixed_deterministic(ctx):
    sc = build_scene(SMALL_THERMAL)
    params = solver_params(SMALL_THERMAL)
    outs = []
    for _ in range(2):
        d = device_scene(ctx, sc, capi_params(params, precision=capi.PREC_MIXED, poisson_tol=1e-10,
                                              poisson_max_iter=500))
        d.upload(T=np.full(sc.nx * sc.ny * sc.nz, params["T_initial"]))
        rs = [d.step() for _ in range(4)]
        outs.append((d.download(), rs))
    for k in outs[0][0]:
        assert np.array_equal(outs[0][0][k], outs[1][0][k]), k
    assert outs[0][1] == outs[1][1]


def test_graph_replay_equals_eager_launches(ctx):
    """The inner PCG iteration replayed from a captured graph gives the same
    bits as launching its kernels one by one, over several steps with eager
    work (downloads, diagnostics) between the replays."""
    from ff_room import device
    cfg = random_scene(77, grid=(40, 36, 20), n_fans=2, n_obstacles=3, precision="mixed",
                       dt=0.005, advection="quick", buoyancy=True, beta=3.4e-3, alpha=2.1e-5,
                       T_initial=26.0, convergence_tol=0.0)
    ps = device.prepare(cfg)
    runs = []
    for graphs in (0, 1):
        prm = capi.default_params(**{k: getattr(ps.params, k) for k, _ in capi.Params._fields_})
        prm.graphs = graphs
        d = device.create(ctx, ps, prm)
        seq = []
        for _ in range(5):
            r = d.step()
            seq.append((r, d.divergence_max(), d.download()))
        d.close()
        runs.append(seq)
    for (ra, da, fa), (rb, db, fb) in zip(*runs):
        assert ra == rb and da == db
        for k in fa:
            assert np.array_equal(fa[k], fb[k]), k
