"""SceneConfig -> device scene through the C ABI (used by bench and serving).

Flags and boundary lists come from the product's Scene (the host
BoundaryManager of _ffroom_core, bit-identical to the reference's scene
builder); parameters follow SolverBridge's mapping; the initial temperature
follows the FluidSolver constructor (T_initial everywhere, opening
temperatures when thermal, fan inlet temperatures on the fans' INFLOW cells).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import capi
from .config import SceneConfig
from .scene import Scene, _load_core
from .solver_bridge import make_params

_PREC = {"fp64": capi.PREC_FP64, "mixed": capi.PREC_MIXED}


@dataclass
class PreparedScene:
    nx: int
    ny: int
    nz: int
    dxv: np.ndarray
    dyv: np.ndarray
    dzv: np.ndarray
    flags: Optional[np.ndarray]  # None: the library marks the cells from `marks`
    fans: List[Tuple]
    outflows: List[Tuple]
    openings: List[Tuple]
    params: capi.Params
    T0: np.ndarray
    marks: Optional[List[Tuple]] = None

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def desc(self):
        return capi.make_desc(self.nx, self.ny, self.nz, self.dxv, self.dyv, self.dzv, self.flags,
                              self.fans, self.outflows, self.openings, self.marks or ())

    def footprint(self) -> int:
        import ctypes
        d, keep = self.desc()
        return capi.lib().ffr_scene_footprint(ctypes.byref(d), ctypes.byref(self.params))


def _params_of(config: SceneConfig, c, overrides) -> capi.Params:
    p = make_params(config, c)
    prm = capi.default_params(
        rho=p.rho, nu=p.nu, dt=p.dt, max_steps=p.max_steps, convergence_tol=p.convergence_tol,
        poisson_max_iter=p.poisson.max_iter, poisson_tol=p.poisson.tol,
        advection=int(p.advection), use_multigrid=int(p.use_multigrid), thermal=int(p.thermal),
        T_initial=p.T_initial, T_target=p.T_target, buoyancy=int(p.buoyancy),
        g_accel=p.g_accel, beta=p.beta, k_thermal=p.k_thermal, cp=p.cp,
        precision=_PREC[(config.solver.precision or "fp64").lower()],
        poisson_method=int(p.poisson_method), mg_levels=p.mg_levels, mg_bf16=int(p.mg_bf16),
        inner_tol=p.inner_tol, inner_max_iter=p.inner_max_iter)
    for k, v in overrides.items():
        setattr(prm, k, v)
    return prm, p


def prepare(config: SceneConfig, on_device: bool = False, **overrides) -> PreparedScene:
    """A scene ready for the device. The cell flags come from the host
    Grid + BoundaryManager (Scene), or with ``on_device`` from marks that
    the library applies on the GPU (no host Grid is built: see
    prepare_marks)."""
    if on_device:
        return prepare_marks(config, **overrides)
    c = _load_core()
    s = Scene(config)
    g, bm = s.grid, s.bm
    prm, p = _params_of(config, c, overrides)
    flags = g.cell_type_shaped().reshape(-1).copy()
    fans = [(f.i_min, f.i_max, f.j_min, f.j_max, f.k_min, f.k_max, *f.vel) for f in bm.fans()]
    outs = [(o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max) for o in bm.outflows()]
    opens = [(o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max, o.T_outside) for o in bm.openings()]
    nx, ny, nz = g.Nx, g.Ny, g.Nz
    # FluidSolver ctor (FluidSolver.cpp:27-31): T_initial, then opening
    # temperatures (velocities are still zero here)
    Tg = g.T_shaped()
    Tg[...] = p.T_initial
    if p.thermal and opens:
        bm.apply_opening_temperature(g)
    T0 = np.array(g.T_shaped()).reshape(-1)
    for n, fan in enumerate(config.fans):
        if getattr(fan, "T_inlet", None) is None:
            continue
        f = fans[n]
        box = np.zeros((nx, ny, nz), dtype=bool)
        box[max(0, f[0]):min(nx - 1, f[1]) + 1, max(0, f[2]):min(ny - 1, f[3]) + 1,
            max(0, f[4]):min(nz - 1, f[5]) + 1] = True
        T0[(box.reshape(-1)) & (flags == capi.INFLOW)] = float(fan.T_inlet)
    r = config.room
    if (getattr(r, "stretch", 1.0) or 1.0) > 1.0 + 1e-6:  # Grid::set_face_coords widths
        dxv, dyv, dzv = (np.diff(np.array(x)) for x in (g.xs(), g.ys(), g.zs()))
    else:  # Grid ctor: uniform L/N (Grid.cpp:8, 35-37)
        dxv, dyv, dzv = (np.full(r.grid[a], r.size[a] / r.grid[a]) for a in range(3))
    return PreparedScene(nx, ny, nz, dxv, dyv, dzv, flags, fans, outs, opens, prm, T0)


def flags_from_marks(shape, marks) -> np.ndarray:
    """Cell flags of marks applied in order on a (nx, ny, nz) grid, boxes
    clamped to it: what the device's marking kernel computes."""
    f = np.zeros(shape, dtype=np.uint8)
    for t, i0, i1, j0, j1, k0, k1 in marks:
        lo = (max(0, i0), max(0, j0), max(0, k0))
        hi = (min(shape[0] - 1, i1) + 1, min(shape[1] - 1, j1) + 1, min(shape[2] - 1, k1) + 1)
        if all(h > l for l, h in zip(lo, hi)):
            f[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = t
    return f


def prepare_marks(config: SceneConfig, **overrides) -> PreparedScene:
    """prepare() without the host Grid: the boxes of the scene (scene.py's
    cell_marks / boundary_lists, the arithmetic Scene uses) go to the library,
    which marks the cells on the device. The initial temperature is the one
    FluidSolver's constructor leaves (FluidSolver.cpp:27-31): T_initial, then
    each opening's wall cells at its outside temperature (velocities are zero,
    so every opening face counts as incoming, BoundaryManager.cpp:137-178),
    then the fans' inlet temperatures on their INFLOW cells."""
    from .scene import boundary_lists, cell_marks, patch_box
    c = _load_core()
    prm, p = _params_of(config, c, overrides)
    r = config.room
    nx, ny, nz = (int(x) for x in r.grid)
    if (getattr(r, "stretch", 1.0) or 1.0) > 1.0 + 1e-6:
        from .scene import _stretched_faces
        dxv, dyv, dzv = (np.diff(np.array(_stretched_faces(r.grid[a], r.size[a], r.stretch)))
                         for a in range(3))
    else:  # Grid ctor: uniform L/N (Grid.cpp:8, 35-37)
        dxv, dyv, dzv = (np.full(r.grid[a], r.size[a] / r.grid[a]) for a in range(3))
    fans, outs, opens = boundary_lists(config)
    marks = cell_marks(config)
    T0 = np.full((nx, ny, nz), p.T_initial)
    if p.thermal and opens:
        for o in opens:
            i0, i1, j0, j1, k0, k1 = patch_box(config, *o[:6])
            T0[i0:i1 + 1, j0:j1 + 1, k0:k1 + 1] = o[6]
    inlets = [(n, f) for n, f in enumerate(config.fans) if getattr(f, "T_inlet", None) is not None]
    if inlets:
        flags = flags_from_marks((nx, ny, nz), marks)
        for n, fan in inlets:
            f = fans[n]
            box = (slice(max(0, f[0]), min(nx - 1, f[1]) + 1), slice(max(0, f[2]), min(ny - 1, f[3]) + 1),
                   slice(max(0, f[4]), min(nz - 1, f[5]) + 1))
            sub = T0[box]
            sub[flags[box] == capi.INFLOW] = float(fan.T_inlet)
    return PreparedScene(nx, ny, nz, dxv, dyv, dzv, None, fans, outs, opens, prm, T0.reshape(-1),
                         marks)


def create(ctx: capi.Context, ps: PreparedScene, params: Optional[capi.Params] = None):
    d = capi.DeviceScene(ctx, ps.nx, ps.ny, ps.nz, ps.dxv, ps.dyv, ps.dzv, ps.flags, ps.fans,
                         ps.outflows, ps.openings, params if params is not None else ps.params,
                         marks=ps.marks or ())
    d.upload(T=ps.T0)
    return d
