"""Scene configuration: the reference's YAML/JSON schema, unchanged.

Field names, defaults and nesting follow ddd3h/ff-room
python/ff_room/config.py:15-204 so that every reference scene file loads as
is. Keys added here are optional and default to the reference's behaviour:

* ``FanConfig.T_inlet``: temperature of the air a fan blows [deg C]
  (None: the fan cells keep T_initial, as in the reference).
* ``SolverConfig.beta``: thermal expansion coefficient [1/K]; None keeps the
  reference's 1/(T_initial + 273.15) (solver_bridge.py:124).
* ``SolverConfig.alpha``: thermal diffusivity [m^2/s]; None keeps
  k_thermal / (rho * cp).
* ``SolverConfig.precision``: "fp64" (reference arithmetic) or "mixed".
* ``SolverConfig.poisson_method``: also "mgpcg" (multigrid-preconditioned CG).
* ``SolverConfig.mg_levels``, ``mg_bf16``, ``inner_tol``, ``inner_max_iter``.
"""
from __future__ import annotations

import dataclasses
import json
import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import yaml


@dataclass
class RoomConfig:
    """Axis-aligned box room: size (Lx, Ly, Lz) in m, grid (Nx, Ny, Nz) cells."""
    size: Tuple[float, float, float] = (5.0, 4.0, 2.5)
    grid: Tuple[int, int, int] = (25, 20, 12)
    stretch: float = 1.0  # >1: geometric refinement towards the walls


@dataclass
class FanConfig:
    """A fan: square inflow patch of half-width ``radius`` centred at
    ``position``, blowing along (azimuth, elevation) degrees at ``velocity``."""
    position: Tuple[float, float, float] = (0.5, 2.0, 1.0)
    direction: Tuple[float, float] = (0.0, 0.0)
    velocity: float = 3.0
    radius: float = 0.15
    name: str = "fan0"
    T_inlet: Optional[float] = None

    def direction_vector(self) -> Tuple[float, float, float]:
        """Unit vector of the blowing direction (config.py:41-48)."""
        az, el = math.radians(self.direction[0]), math.radians(self.direction[1])
        ce = math.cos(el)
        return (ce * math.cos(az), ce * math.sin(az), math.sin(el))

    def snapped_axis(self) -> int:
        """Axis (0, 1, 2) of the largest |component|; ties go to the lower axis."""
        vec = self.direction_vector()
        best = 0
        for a in (1, 2):
            if abs(vec[a]) > abs(vec[best]):
                best = a
        return best

    def snapped_sign(self) -> float:
        """+1 or -1: the sign of the dominant component (0 counts as +)."""
        return 1.0 if self.direction_vector()[self.snapped_axis()] >= 0 else -1.0


@dataclass
class ObstacleConfig:
    """Solid box between bbox_min and bbox_max [m]."""
    bbox_min: Tuple[float, float, float] = (2.0, 1.5, 0.0)
    bbox_max: Tuple[float, float, float] = (2.5, 2.5, 0.8)
    name: str = "obstacle0"


@dataclass
class OpeningConfig:
    """Window or door on a wall: pressure outflow plus temperature exchange.

    wall: west/east (x=0/Lx), south/north (y=0/Ly), floor/ceiling (z=0/Lz);
    center and size are in the wall's two transverse coordinates."""
    wall: str = "east"
    center: Tuple[float, float] = (2.0, 1.2)
    size: Tuple[float, float] = (0.8, 1.0)
    T_outside: Optional[float] = None
    name: str = "opening0"


@dataclass
class SolverConfig:
    """Numerical settings (reference defaults, config.py:93-142)."""
    dt: float = 0.01
    max_steps: int = 1000
    convergence_tol: float = 1.0
    rho: float = 1.2
    nu: float = 1.5e-5
    poisson_max_iter: int = 1000
    poisson_tol: float = 1e-6
    method: str = "projection"
    tau: float = 0.8
    advection: str = "upwind"
    parallel: bool = False
    poisson_method: str = "cg"
    T_initial: float = 20.0
    T_outside: float = 20.0
    T_target: Optional[float] = None
    buoyancy: bool = False
    k_thermal: float = 0.026
    cp: float = 1005.0
    # ---- optional extensions ------------------------------------------
    beta: Optional[float] = None
    alpha: Optional[float] = None
    precision: str = "fp64"
    mg_levels: int = 0
    mg_bf16: bool = True
    inner_tol: float = 1e-4
    inner_max_iter: int = 60


@dataclass
class SceneConfig:
    """A complete, serialisable scene."""
    room: RoomConfig = field(default_factory=RoomConfig)
    fans: List[FanConfig] = field(default_factory=list)
    obstacles: List[ObstacleConfig] = field(default_factory=list)
    openings: List[OpeningConfig] = field(default_factory=list)
    solver: SolverConfig = field(default_factory=SolverConfig)
    metadata: dict = field(default_factory=dict)

    def to_dict(self) -> dict:
        # plain dicts/lists only, so that yaml.safe_load reads the file back
        def plain(x):
            if dataclasses.is_dataclass(x):
                x = dataclasses.asdict(x)
            if isinstance(x, dict):
                return {k: plain(v) for k, v in x.items()}
            if isinstance(x, (list, tuple)):
                return [plain(v) for v in x]
            return x
        return plain(self)

    @classmethod
    def from_dict(cls, d: dict) -> "SceneConfig":
        # lists become tuples; unknown keys raise TypeError as in the reference
        def tup(items):
            return {k: (tuple(v) if isinstance(v, list) else v) for k, v in items.items()}
        return cls(
            room=RoomConfig(**tup(d.get("room", {}))),
            fans=[FanConfig(**tup(f)) for f in d.get("fans", [])],
            obstacles=[ObstacleConfig(**tup(o)) for o in d.get("obstacles", [])],
            openings=[OpeningConfig(**tup(o)) for o in d.get("openings", [])],
            solver=SolverConfig(**d.get("solver", {})),
            metadata=d.get("metadata", {}),
        )

    def save_yaml(self, path: str) -> None:
        with open(path, "w") as f:
            yaml.dump(self.to_dict(), f, default_flow_style=False, sort_keys=False)

    def save_json(self, path: str) -> None:
        with open(path, "w") as f:
            json.dump(self.to_dict(), f, indent=2)

    @classmethod
    def load_yaml(cls, path: str) -> "SceneConfig":
        with open(path) as f:
            return cls.from_dict(yaml.safe_load(f))

    @classmethod
    def load_json(cls, path: str) -> "SceneConfig":
        with open(path) as f:
            return cls.from_dict(json.load(f))
