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
        ax = fan.snapped_axis()
        o = self._c.OutflowBC()
        o.axis = ax
        o.side = 1 if fan.snapped_sign() > 0 else 0
        ta, tb = [a for a in range(3) if a != ax]
        o.a_min, o.a_max = 0, self._cells(ta)[1] - 1
        o.b_min, o.b_max = 0, self._cells(tb)[1] - 1
        self.bm.add_outflow(self.grid, o)

    def _add_obstacle(self, ob: ObstacleConfig) -> None:
        lo, hi = [], []
        for a in range(3):
            L, N = self._cells(a)
            lo.append(int(ob.bbox_min[a] / L * N))
            hi.append(int(ob.bbox_max[a] / L * N) - 1)
        self.bm.add_solid_box(self.grid, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])

    def _add_fan(self, fan: FanConfig) -> None:
        ax = fan.snapped_axis()
        vec = fan.direction_vector()
        bc = self._c.FanBC()
        bc.vel = [vec[0] * fan.velocity, vec[1] * fan.velocity, vec[2] * fan.velocity]
        box = []
        for a in range(3):
            L, N = self._cells(a)
            if a == ax:  # one cell thick along the blowing axis
                c = max(0, min(N - 1, int(fan.position[a] / L * N)))
                box += [c, c]
            else:
                box += [max(0, int((fan.position[a] - fan.radius) / L * N)),
                        min(N - 1, int((fan.position[a] + fan.radius) / L * N))]
        bc.i_min, bc.i_max, bc.j_min, bc.j_max, bc.k_min, bc.k_max = box
        self.bm.add_fan(self.grid, bc)

    def _add_opening(self, op: OpeningConfig) -> None:
        ax, sd = _WALLS[op.wall.lower()]  # KeyError for an unknown wall, as in the reference
        ta, tb = [a for a in range(3) if a != ax]

        def span(c, w, a):
            L, N = self._cells(a)
            return max(0, int((c - w / 2) / L * N)), min(N - 1, int((c + w / 2) / L * N))

        a_lo, a_hi = span(op.center[0], op.size[0], ta)
        b_lo, b_hi = span(op.center[1], op.size[1], tb)
        o = self._c.OpeningBC()
        o.axis, o.side = ax, sd
        o.a_min, o.a_max, o.b_min, o.b_max = a_lo, a_hi, b_lo, b_hi
        o.T_outside = op.T_outside if op.T_outside is not None else self.config.solver.T_outside
        self.bm.add_opening(self.grid, o)

    def rebuild(self) -> None:
        self._build()
