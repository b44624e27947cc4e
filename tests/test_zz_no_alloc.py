"""ffr_poisson_solve_noalloc against ffr_poisson_solve, through the C ABI.

The entry point that allocates nothing runs the same CG kernels with the
same reduction order as the allocating one, so its solution, iteration count
and residual are the same bits. Checked on two calls in a row, the second on
a smaller scene (the fixed array keeps the first call's contents), on a
thermal scene with openings, and then again on the first scene; plus the
rejections (a scene above FFR_NOALLOC_MAX_CELLS, a multigrid scene).
"""
import numpy as np
import pytest

from helpers import capi_params, device_scene
from oracle.pyoracle import SOLID
from oracle.scene_flags import build_scene
from scenes import BASIC_ROOM, SMALL_THERMAL, solver_params

pytestmark = pytest.mark.gpu

LARGER = {  # 48 x 40 x 24 = 46,080 cells: larger than the c1 room
    "room": {"size": [6.0, 5.0, 3.0], "grid": [48, 40, 24]},
    "fans": [{"name": "f", "position": [1.0, 2.5, 1.5], "direction": [0.0, 0.0],
              "velocity": 3.0, "radius": 0.3}],
    "obstacles": [{"name": "b", "bbox_min": [3.0, 1.0, 0.0], "bbox_max": [4.2, 2.6, 1.1]}],
}


def _rhs(sc, seed):
    b = np.random.default_rng(seed).standard_normal(sc.nx * sc.ny * sc.nz)
    b[sc.flags == SOLID] = 0.0
    return b


def _pair(ctx, cfg, seed, **extra):
    sc = build_scene(cfg)
    prm = capi_params(solver_params(cfg), **extra)
    b = _rhs(sc, seed)
    d_ref = device_scene(ctx, sc, prm)
    want, info_w = d_ref.poisson(b)
    d_ref.close()
    d = device_scene(ctx, sc, prm)
    got, info_g = d.poisson_noalloc(b)
    pressure = d.download()["p"]
    d.close()
    return want, info_w, got, info_g, pressure


def _same(want, info_w, got, info_g, pressure):
    assert np.array_equal(got, want)
    assert np.array_equal(pressure, want)  # the solution is also the scene's pressure
    assert info_g == info_w
    assert info_g["iterations"] > 0 and info_g["outer_iterations"] == 0


def test_noalloc_solve_equals_allocating_solve(ctx):
    # first call on the larger scene, the second on a smaller one
    r1 = _pair(ctx, LARGER, 1, poisson_tol=1e-10, poisson_max_iter=5000)
    _same(*r1)
    r2 = _pair(ctx, BASIC_ROOM, 2)
    _same(*r2)
    assert r2[0].size < r1[0].size
    _same(*_pair(ctx, SMALL_THERMAL, 3, poisson_tol=1e-9))
    _same(*_pair(ctx, LARGER, 4, poisson_max_iter=7))  # stopped by the iteration cap
    # the larger scene again, after smaller calls left their contents
    r5 = _pair(ctx, LARGER, 1, poisson_tol=1e-10, poisson_max_iter=5000)
    _same(*r5)
    assert np.array_equal(r5[2], r1[2])


def test_noalloc_rejections(ctx):
    from ff_room import capi
    sc = build_scene(BASIC_ROOM)
    d = device_scene(ctx, sc, capi_params({}, precision=capi.PREC_MIXED,
                                          poisson_method=capi.POISSON_MGPCG))
    with pytest.raises(capi.FfrError) as e:
        d.poisson_noalloc(_rhs(sc, 5))
    assert e.value.status == -6  # FFR_E_UNSUPPORTED
    d.close()
    nz = (1 << 18) + 1  # 2 x 2 x nz cells: 4 above the limit
    big = capi.DeviceScene(ctx, 2, 2, nz, np.ones(2), np.ones(2), np.ones(nz),
                           np.zeros(4 * nz, dtype=np.uint8))
    with pytest.raises(capi.FfrError) as e:
        big.poisson_noalloc(np.ones(4 * nz))
    assert e.value.status == -1 and "FFR_NOALLOC_MAX_CELLS" in str(e.value)
    big.close()
