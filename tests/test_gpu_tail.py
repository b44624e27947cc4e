"""Coarse multigrid levels run from global memory inside the single-workgroup
V-cycle kernel (ffr_params.mg_tail_cells), the form used for small scenes
served side by side: one launch per V-cycle for all coarse levels. Those
global levels may be shared by several workgroups on one XCD
(ffr_params.mg_tail_groups), ordered by a barrier in that XCD's L2.

Pins: the dense solution of the same system (tiny semi-coarsened grids, the
hierarchy's tail forced to start at level 1 with its LDS levels after it), and
on a c4-sized room the solve of the per-level launches (both to a residual of
1e-11: relative L2 <= 1e-8, the condition number times the residual). The hierarchy introspection (ffr_scene_mg_layout)
shows that the global levels are in use.
"""
import numpy as np
import pytest

from ff_room import capi
from ff_room.generator import workload_c4
from helpers import capi_params, device_scene, rel_l2
from oracle.pyoracle import OUTFLOW, SOLID, SceneArrays
from test_gpu_mg import _dense, _rhs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("coarse", [8, 0])
@pytest.mark.parametrize("dims,size", [((16, 12, 14), (4.0, 3.0, 0.7)),
                                       ((14, 16, 12), (1.4, 3.2, 0.6)),
                                       ((11, 9, 13), (1.0, 1.0, 1.0))])
def test_global_tail_levels_solve_the_system(ctx, dims, size, coarse):
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
                                          mg_coarse_sweeps=coarse, mg_tail_kb=12,
                                          mg_tail_cells=1 << 20))
    lay = d.mg_layout()
    assert lay["levels"] >= 3, lay
    assert lay["tail_first"] == 1 and lay["tail_lds_from"] > 1, lay
    p, info = d.poisson(b)
    d.close()
    assert info["residual"] < 1e-11
    assert rel_l2(p, x) < 1e-9
    assert info["iterations"] <= 60, info


@pytest.mark.parametrize("room", [0, 3])
def test_global_tail_matches_per_level_launches_on_c4_rooms(ctx, room):
    """A consistent right-hand side b = A x* (FP64 operator; c4 rooms can hold
    sealed pockets, where A is singular and x is defined up to a constant):
    both hierarchies reach the 1e-11 residual and their solutions agree
    gauge-free to 1e-8 (the condition number, ~1e3, times the residual)."""
    from types import SimpleNamespace

    from bench import gauge_free_error, sealed_components
    from ff_room import device
    cfg = workload_c4(n_scenes=room + 1, steps=5, precision="mixed")[room]
    sols = {}
    for cells in (0, 1 << 17):
        ps = device.prepare(cfg, on_device=True, precision=capi.PREC_MIXED,
                            poisson_method=capi.POISSON_MGPCG, poisson_tol=1e-11,
                            poisson_max_iter=500, mg_tail_cells=cells)
        d = device.create(ctx, ps)
        lay = d.mg_layout()
        if cells:
            assert lay["tail_first"] == 1 and lay["tail_lds_from"] > 1, lay
        else:
            assert lay["tail_first"] == lay["tail_lds_from"], lay
        flags = d.cell_type()
        xs = np.random.default_rng(7).standard_normal(ps.nx * ps.ny * ps.nz)
        xs[flags == SOLID] = 0.0
        b = d.laplacian(xs)
        x, info = d.poisson(b)
        d.close()
        assert info["residual"] < 1e-11, (cells, info)
        sols[cells] = x
        view = SimpleNamespace(flags=flags, nx=ps.nx, ny=ps.ny, nz=ps.nz)
    sealed = sealed_components(view)
    assert gauge_free_error(sols[1 << 17].copy(), sols[0], view, sealed) < 1e-8



def _c4_room(room, **kw):
    from ff_room import device
    cfg = workload_c4(n_scenes=room + 1, steps=5, precision="mixed")[room]
    return device.prepare(cfg, on_device=True, precision=capi.PREC_MIXED,
                          poisson_method=capi.POISSON_MGPCG, **kw)


@pytest.mark.parametrize("room", [0, 3, 5])
def test_tail_groups_solve_equals_one_workgroup_bitwise(ctx, room):
    """The global levels' phases split over 2, 3, 8 and the automatic number
    of workgroups give the one-workgroup tail's solution bit for bit: every
    phase computes each cell from the same values in the same order, only the
    threads that take the cells differ."""
    from ff_room import device
    sols, layouts = {}, {}
    for groups in (1, 2, 3, 8, 0):
        ps = _c4_room(room, poisson_tol=1e-11, poisson_max_iter=500, mg_tail_cells=1 << 17,
                      mg_tail_groups=groups)
        d = device.create(ctx, ps)
        lay = d.mg_layout()
        assert lay["tail_first"] == 1 and lay["tail_lds_from"] > 1, lay
        flags = d.cell_type()
        xs = np.random.default_rng(3).standard_normal(ps.nx * ps.ny * ps.nz)
        xs[flags == SOLID] = 0.0
        b = d.laplacian(xs)
        x, info = d.poisson(b)
        x2, info2 = d.poisson(b)  # a second launch reuses the reset counters
        d.close()
        assert info["residual"] < 1e-11, (groups, info)
        assert np.array_equal(x, x2)
        sols[groups], layouts[groups] = x, lay
    auto = layouts[0]["tail_groups"]
    first = layouts[0]["cells"][1]
    assert auto == max(1, min(8, -(-first // 16384))), layouts[0]
    assert [layouts[g]["tail_groups"] for g in (1, 2, 3, 8)] == [1, 2, 3, 8]
    for g in (2, 3, 8, 0):
        assert np.array_equal(sols[g], sols[1]), g


@pytest.mark.parametrize("room", [1, 4])
def test_tail_groups_steps_equal_one_workgroup_bitwise(ctx, room):
    """Mixed steps (graph replay of the inner iteration included) with the
    grouped tail equal the one-workgroup tail's steps bit for bit."""
    from ff_room import device
    out = {}
    for groups in (1, 0):
        ps = _c4_room(room, mg_tail_cells=1 << 16, mg_tail_groups=groups)
        d = device.create(ctx, ps)
        for _ in range(4):
            d.step()
        out[groups] = (d.download(), d.mg_layout()["tail_groups"])
        d.close()
    assert out[1][1] == 1 and out[0][1] >= 1
    for k in ("u", "v", "w", "p", "T"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k


def test_tail_groups_served_side_by_side_equal_alone(ctx):
    """Six c4 rooms served together by worker threads, each scene's grouped
    tail on an XCD of its own (scenes take the XCDs in turn, so several
    groups run at once), equal their alone runs bit for bit."""
    from ff_room import device
    from ff_room.serve import Request, Server
    K = 3
    cfgs = workload_c4(n_scenes=6, steps=K, precision="mixed")
    params = dict(mg_tail_cells=1 << 16, mg_tail_groups=0)
    alone = []
    for cfg in cfgs:
        d = device.create(ctx, device.prepare(cfg, on_device=True, **params))
        for _ in range(K):
            d.step()
        alone.append(d.download())
        d.close()
    srv = Server(ctx, slots=len(cfgs), keep_fields=True, tuned=False, **params)
    done = srv.run([Request(c, K) for c in cfgs], realtime=False)
    assert sorted(c.index for c in done) == list(range(len(cfgs)))
    for c in done:
        assert c.steps == K
        for k in ("u", "v", "w", "p", "T"):
            assert np.array_equal(c.fields[k], alone[c.index][k]), (c.index, k)
