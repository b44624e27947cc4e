"""The full-size rooms (c2: 256x204x80, c3: 512x408x160), checked through
properties that do not depend on size, where the oracle would take minutes:
every pressure solve meets its tolerance, the state stays finite, two runs
give the same bits, and one mixed step from a given state stays within 1e-6
(relative L2 of p, u, v, w and T) of the FP64 step from the same state, on
c2 and on c3. The
lattice and the stretched grid are checked the same way at c2 size.
"""
import numpy as np
import pytest

from ff_room import capi, device
from ff_room.generator import workload_c2, workload_c3

pytestmark = pytest.mark.gpu


def _params(ps, mixed: bool):
    p = capi.default_params(**{k: getattr(ps.params, k) for k, _ in capi.Params._fields_})
    if mixed:
        p.precision, p.poisson_method, p.poisson_tol, p.poisson_max_iter = 1, 2, 1e-10, 500
    else:
        p.precision, p.poisson_method, p.poisson_tol, p.poisson_max_iter = 0, 1, 1e-8, 200000
    return p


def _run(ctx, ps, params, steps):
    d = device.create(ctx, ps, params)
    rs = [d.step() for _ in range(steps)]
    f = d.download()
    d.close()
    return rs, f


@pytest.mark.parametrize("name,steps", [("c2", 3), ("c3", 2)])
def test_fullsize_mixed_steps_converge_and_repeat(ctx, name, steps):
    cfg = (workload_c2 if name == "c2" else workload_c3)(precision="mixed")
    ps = device.prepare(cfg, on_device=True)
    runs = [_run(ctx, ps, _params(ps, True), steps) for _ in range(2)]
    for r in runs[0][0]:
        assert not r["nonfinite"]
        assert r["pressure_residual"] < 1e-10, r
        assert r["outer_iters"] >= 1
    for k in runs[0][1]:
        assert np.isfinite(runs[0][1][k]).all(), k
        assert np.array_equal(runs[0][1][k], runs[1][1][k]), k
    assert runs[0][0] == runs[1][0]


@pytest.mark.parametrize("name,tol", [("c2", 1e-8), ("c3", 1e-9)])
def test_fullsize_mixed_step_matches_fp64_step(ctx, name, tol):
    """One step of each form from the same state (the mixed form's after
    three steps; FP32 values are exact in FP64), the FP64 form's CG run to the
    loosest tolerance whose pressure error meets 1e-6 on that room: p, u, v, w
    and T within 1e-6 relative L2 (p gauge-free: c3 holds a sealed pocket, on
    which each form's constant offset is removed)."""
    from types import SimpleNamespace

    from bench import gauge_free_error, sealed_components
    cfg = (workload_c2 if name == "c2" else workload_c3)(precision="mixed")
    ps = device.prepare(cfg, on_device=True)
    dm = device.create(ctx, ps, _params(ps, True))
    for _ in range(3):
        dm.step()
    st = dm.download()
    p64 = _params(ps, False)
    p64.poisson_tol = tol
    d64 = device.create(ctx, ps, p64)
    d64.upload(**st)
    r64 = d64.step()
    ref = d64.download()
    flags = d64.cell_type()
    d64.close()
    dm.step()
    got = dm.download()
    dm.close()
    assert r64["pressure_residual"] < tol
    for k in ("u", "v", "w", "T"):
        err = np.linalg.norm(got[k] - ref[k]) / max(np.linalg.norm(ref[k]), 1e-300)
        assert err < 1e-6, (k, err)
    view = SimpleNamespace(flags=flags, nx=ps.nx, ny=ps.ny, nz=ps.nz)
    assert gauge_free_error(got["p"].copy(), ref["p"], view, sealed_components(view)) < 1e-6


def test_fullsize_stretched_mixed_steps_converge(ctx):
    cfg = workload_c2(precision="mixed")
    cfg.room.stretch = 1.01
    ps = device.prepare(cfg, on_device=True)
    assert ps.dxv.max() / ps.dxv.min() > 3.0
    rs, f = _run(ctx, ps, _params(ps, True), 2)
    for r in rs:
        assert not r["nonfinite"] and r["pressure_residual"] < 1e-10, r
    assert all(np.isfinite(v).all() for v in f.values())


def test_fullsize_lattice_repeats_and_stays_finite(ctx):
    cfg = workload_c2(precision="mixed")
    ps = device.prepare(cfg, on_device=True)
    outs = []
    for _ in range(2):
        prm = capi.default_params(dt=ps.params.dt, precision=capi.PREC_FP64,
                                  poisson_method=capi.POISSON_CG)
        d = capi.DeviceScene(ctx, ps.nx, ps.ny, ps.nz, ps.dxv, ps.dyv, ps.dzv, ps.flags, ps.fans,
                             ps.outflows, ps.openings, prm)
        d.lbm_attach(0.8, ps.params.dt, 0.0, float(cfg.room.size[0]) / ps.nx)
        rs = [d.lbm_step() for _ in range(5)]
        outs.append((rs, d.download()))
        d.close()
    assert outs[0][0] == outs[1][0]
    for k in ("u", "v", "w", "p"):
        assert np.isfinite(outs[0][1][k]).all(), k
        assert np.array_equal(outs[0][1][k], outs[1][1][k]), k
    assert outs[0][0][-1]["velocity_change"] > 0.0
