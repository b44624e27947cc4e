"""The C ABI library and the pybind11 module load and expose their surface (CPU)."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT
from ff_room import capi

HEADER = os.path.join(ROOT, "include", "ffroom.h")


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    declared = capi.header_symbols(HEADER)
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(capi.SIGNATURES), "ctypes signatures out of sync with the header"


def test_defaults_are_the_reference_defaults():
    # FluidSolver.hpp:11-38, PoissonSolver.hpp:14-18
    p = capi.default_params()
    assert (p.rho, p.nu, p.dt, p.max_steps, p.convergence_tol) == (1.2, 1.5e-5, 0.01, 500, 1.0)
    assert (p.poisson_max_iter, p.poisson_tol) == (1000, 1e-6)
    assert p.advection == capi.UPWIND1 and not p.use_multigrid and not p.thermal
    assert (p.T_initial, p.T_target, p.g_accel, p.beta, p.k_thermal, p.cp) == (
        20.0, 1e30, 9.81, 1.0 / 293.0, 0.026, 1005.0)
    assert p.precision == capi.PREC_FP64 and p.poisson_method == capi.POISSON_AUTO


def test_status_strings_and_version():
    L = capi.lib()
    assert L.ffr_abi_version() == 1
    assert L.ffr_status_string(0) == b"ok"
    assert b"invalid" in L.ffr_status_string(-1)


def test_footprint_scales_with_precision_and_grid():
    flags = np.zeros(64 * 48 * 24, dtype=np.uint8)
    flags[-48 * 24:] = capi.OUTFLOW
    h = lambda n, L: np.full(n, L / n)
    d, keep = capi.make_desc(64, 48, 24, h(64, 10), h(48, 8), h(24, 3), flags,
                             outflows=[(0, 1, 0, 47, 0, 23)])
    L = capi.lib()
    f64 = L.ffr_scene_footprint(ctypes.byref(d), ctypes.byref(capi.default_params()))
    f32cg = L.ffr_scene_footprint(ctypes.byref(d), ctypes.byref(
        capi.default_params(precision=1, poisson_method=capi.POISSON_CG)))
    fmg = L.ffr_scene_footprint(ctypes.byref(d), ctypes.byref(
        capi.default_params(precision=1, poisson_method=capi.POISSON_MGPCG)))
    n = 64 * 48 * 24
    assert 90 * n < f64 < 140 * n   # FP64 form: ~105 B/cell
    assert 0 < f32cg < f64          # FP32 fields, same solver
    # the multigrid hierarchy adds ~26 B/cell on level 0 (six FP32 vectors:
    # e, r, two directions, q, z; split codes) plus 37 B per coarse cell
    # (operator, per-axis diagonal, BF16 z and r), the coarse levels holding
    # <= 1/7 of the cells when the cells are near-cubic (1/2 when
    # semi-coarsening halves one axis)
    assert f32cg < fmg < f32cg + 34 * n
    # invalid descriptors are rejected (footprint 0), e.g. a patch off the wall
    d2, keep2 = capi.make_desc(64, 48, 24, h(64, 10), h(48, 8), h(24, 3), flags,
                               outflows=[(0, 1, 0, 48, 0, 23)])
    assert L.ffr_scene_footprint(ctypes.byref(d2), ctypes.byref(capi.default_params())) == 0


def test_module_surface_matches_reference_bindings():
    from ff_room.scene import _load_core
    c = _load_core()
    # every class / enum of cpp/src/bindings/module.cpp except LBMSolver
    for name in ("AdvectionScheme", "CellType", "Grid", "FanBC", "OutflowBC", "OpeningBC",
                 "BoundaryManager", "PoissonSolverParams", "FluidSolverParams", "StepResult",
                 "FluidSolver"):
        assert hasattr(c, name), name
    g = c.Grid(4, 5, 6, 4.0, 5.0, 6.0)
    assert (g.Nx, g.Ny, g.Nz, g.dx) == (4, 5, 6, 1.0)
    assert g.u.size == 5 * 5 * 6 and g.v.size == 4 * 6 * 6 and g.w.size == 4 * 5 * 7
    assert g.u_shaped().shape == (5, 5, 6) and g.p_shaped().shape == (4, 5, 6)
    g.u[3] = 7.0  # views are writable and zero-copy
    assert g.u_shaped().reshape(-1)[3] == 7.0
    g.zero_velocity()
    assert not g.u.any()
    p = c.FluidSolverParams()
    assert p.poisson.max_iter == 1000 and p.advection == c.AdvectionScheme.UPWIND1
    f = c.FanBC()
    f.vel = [1.0, 2.0, 3.0]
    assert list(f.vel) == [1.0, 2.0, 3.0]


def test_no_gpu_fails_loudly():
    """Without a GPU the product refuses to run instead of falling back."""
    from ff_room.scene import _load_core
    try:
        ctx = capi.Context(0)
    except capi.FfrError as e:
        assert e.status == -2 or e.status == -1
        c = _load_core()
        g = c.Grid(4, 4, 4, 1.0, 1.0, 1.0)
        bm = c.BoundaryManager()
        with pytest.raises(RuntimeError):
            c.FluidSolver(g, bm, c.FluidSolverParams())
    else:
        ctx.close()
        pytest.skip("a GPU is present")
