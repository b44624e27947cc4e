"""Scene: SceneConfig -> Grid + BoundaryManager of the GPU core.

Cell flags are written exactly as the reference writes them
(python/ff_room/scene.py:72-231): positions map to cells with
``int(x / L * N)`` (truncation), fan and opening ranges are clamped to the
grid, and the marking order is obstacles -> fans -> openings -> automatic
outflow on the wall the first fan blows towards (only when the scene has no
openings). Later marks overwrite earlier ones.
"""
from __future__ import annotations

import importlib
import os
import sys
from typing import List

from .config import FanConfig, ObstacleConfig, OpeningConfig, SceneConfig

_WALLS = {  # wall -> (axis, side)
    "west": (0, 0), "east": (0, 1),
    "south": (1, 0), "north": (1, 1),
    "floor": (2, 0), "ceiling": (2, 1),
}


def _load_core():
    """Import the compiled ``_ffroom_core`` (built next to this file)."""
    here = os.path.dirname(os.path.abspath(__file__))
    if here not in sys.path:
        sys.path.insert(0, here)
    try:
        return importlib.import_module("_ffroom_core")
    except ImportError as e:
        raise ImportError(
            "GPU extension _ffroom_core not found or not loadable; build it with "
            "`python ff-room_amd/build.py` (needs ROCm and hipcc for gfx950)") from e


def _stretched_faces(N: int, L: float, r: float) -> List[float]:
    """N+1 face coordinates over [0, L], cells growing geometrically by r
    from both walls towards the middle (reference scene.py:34-59)."""
    if abs(r - 1.0) < 1e-6:
        return [i * L / N for i in range(N + 1)]
    half = N // 2
    first = (L / 2.0) * (r - 1.0) / (r ** half - 1.0)
    faces = [0.0]
    for i in range(half):
        faces.append(faces[-1] + first * r ** i)
    # mirror: the right half reuses the left half's realised widths
    for i in reversed(range(half)):
        faces.append(faces[-1] + (faces[i + 1] - faces[i]))
    if N % 2 == 1:  # odd N: a middle cell splitting the central gap
        faces.insert(half + 1, 0.5 * (faces[half] + faces[half + 1]))
    scale = L / faces[-1]
    return [f * scale for f in faces]


# ---- the boxes the scene marks, as pure functions of the configuration ----
# Scene (below) hands them to the host BoundaryManager; cell_marks and
# boundary_lists give the same boxes to a scene whose flags are marked on the
# device (ffr_scene_desc.marks), so both paths share this arithmetic.
FLUID, SOLID, INFLOW, OUTFLOW = 0, 1, 2, 3


def _cells(config: SceneConfig, axis: int):
    r = config.room
    return float(r.size[axis]), int(r.grid[axis])


def obstacle_box(config: SceneConfig, ob: ObstacleConfig):
    """(i0, i1, j0, j1, k0, k1) of an obstacle, unclamped (add_solid_box clamps)."""
    lo, hi = [], []
    for a in range(3):
        L, N = _cells(config, a)
        lo.append(int(ob.bbox_min[a] / L * N))
        hi.append(int(ob.bbox_max[a] / L * N) - 1)
    return lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]


def fan_box(config: SceneConfig, fan: FanConfig):
    """Cell box of a fan: one cell thick along its snapped axis, its radius
    across, clamped to the grid."""
    ax = fan.snapped_axis()
    box = []
    for a in range(3):
        L, N = _cells(config, a)
        if a == ax:  # one cell thick along the blowing axis
            c = max(0, min(N - 1, int(fan.position[a] / L * N)))
            box += [c, c]
        else:
            box += [max(0, int((fan.position[a] - fan.radius) / L * N)),
                    min(N - 1, int((fan.position[a] + fan.radius) / L * N))]
    return tuple(box)


def fan_velocity(fan: FanConfig):
    vec = fan.direction_vector()
    return (vec[0] * fan.velocity, vec[1] * fan.velocity, vec[2] * fan.velocity)


def opening_patch(config: SceneConfig, op: OpeningConfig):
    """(axis, side, a0, a1, b0, b1) of an opening on its wall."""
    ax, sd = _WALLS[op.wall.lower()]  # KeyError for an unknown wall, as in the reference
    ta, tb = [a for a in range(3) if a != ax]

    def span(c, w, a):
        L, N = _cells(config, a)
        return max(0, int((c - w / 2) / L * N)), min(N - 1, int((c + w / 2) / L * N))

    a_lo, a_hi = span(op.center[0], op.size[0], ta)
    b_lo, b_hi = span(op.center[1], op.size[1], tb)
    return ax, sd, a_lo, a_hi, b_lo, b_hi


def auto_outflow_patch(config: SceneConfig, fan: FanConfig):
    """The whole wall the first fan blows towards."""
    ax = fan.snapped_axis()
    ta, tb = [a for a in range(3) if a != ax]
    return ax, 1 if fan.snapped_sign() > 0 else 0, 0, _cells(config, ta)[1] - 1, 0, \
        _cells(config, tb)[1] - 1


def patch_box(config: SceneConfig, axis: int, side: int, a0: int, a1: int, b0: int, b1: int):
    """Cell box of a wall patch: the wall's layer of cells, a and b the two
    other axes in order (BoundaryManager.cpp:24-44)."""
    w = (_cells(config, axis)[1] - 1) if side else 0
    box = [None] * 3
    box[axis] = (w, w)
    ta, tb = [a for a in range(3) if a != axis]
    box[ta], box[tb] = (a0, a1), (b0, b1)
    return box[0] + box[1] + box[2]


def opening_T_outside(config: SceneConfig, op: OpeningConfig) -> float:
    return op.T_outside if op.T_outside is not None else config.solver.T_outside


def cell_marks(config: SceneConfig):
    """[(type, i0, i1, j0, j1, k0, k1)] in the marking order of Scene._build:
    obstacles, fans, openings, then the automatic outflow (scene.py:89-106)."""
    marks = [(SOLID,) + tuple(obstacle_box(config, ob)) for ob in config.obstacles]
    marks += [(INFLOW,) + fan_box(config, f) for f in config.fans]
    marks += [(OUTFLOW,) + tuple(patch_box(config, *opening_patch(config, op)))
              for op in config.openings]
    if config.fans and not config.openings:
        marks.append((OUTFLOW,) + tuple(patch_box(config, *auto_outflow_patch(config, config.fans[0]))))
    return marks


def boundary_lists(config: SceneConfig):
    """(fans, outflows, openings) as the BoundaryManager registers them:
    fans as (box..., vx, vy, vz); outflows hold every opening's patch, then
    the automatic outflow; openings add their outside temperature."""
    fans = [fan_box(config, f) + fan_velocity(f) for f in config.fans]
    opens = [opening_patch(config, op) + (opening_T_outside(config, op),) for op in config.openings]
    outs = [o[:6] for o in opens]
    if config.fans and not config.openings:
        outs.append(auto_outflow_patch(config, config.fans[0]))
    return fans, outs, opens


class Scene:
    """A built scene: ``grid`` and ``bm`` objects of ``_ffroom_core``."""

    def __init__(self, config: SceneConfig):
        self.config = config
        self._c = _load_core()
        self.grid = None
        self.bm = None
        self._build()

    def _cells(self, axis: int):
        r = self.config.room
        return float(r.size[axis]), int(r.grid[axis])

    def _build(self) -> None:
        cfg, r = self.config, self.config.room
        self.grid = self._c.Grid(r.grid[0], r.grid[1], r.grid[2], r.size[0], r.size[1], r.size[2])
        stretch = getattr(r, "stretch", 1.0) or 1.0
        if stretch > 1.0 + 1e-6:
            self.grid.set_face_coords(*[_stretched_faces(r.grid[a], r.size[a], stretch)
                                        for a in range(3)])
        self.bm = self._c.BoundaryManager()
        for ob in cfg.obstacles:
            self._add_obstacle(ob)
        for n, fan in enumerate(cfg.fans):
            self._add_fan(fan)
            if getattr(fan, "T_inlet", None) is not None:
                self.bm.set_fan_temperature(n, float(fan.T_inlet))
        for op in cfg.openings:
            self._add_opening(op)
        if cfg.fans and not cfg.openings:
            self._add_auto_outflow(cfg.fans[0])

    def _add_auto_outflow(self, fan: FanConfig) -> None:
        """Whole wall the first fan blows towards becomes OUTFLOW (p = 0)."""
        o = self._c.OutflowBC()
        o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max = auto_outflow_patch(self.config, fan)
        self.bm.add_outflow(self.grid, o)

    def _add_obstacle(self, ob: ObstacleConfig) -> None:
        self.bm.add_solid_box(self.grid, *obstacle_box(self.config, ob))

    def _add_fan(self, fan: FanConfig) -> None:
        bc = self._c.FanBC()
        bc.vel = list(fan_velocity(fan))
        bc.i_min, bc.i_max, bc.j_min, bc.j_max, bc.k_min, bc.k_max = fan_box(self.config, fan)
        self.bm.add_fan(self.grid, bc)

    def _add_opening(self, op: OpeningConfig) -> None:
        o = self._c.OpeningBC()
        o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max = opening_patch(self.config, op)
        o.T_outside = opening_T_outside(self.config, op)
        self.bm.add_opening(self.grid, o)

    def rebuild(self) -> None:
        self._build()
