"""Scene definitions shared by the tests and the golden-fixture script.

Each entry is a SceneConfig-shaped dict plus the FluidSolverParams fields the
reference's SolverBridge would derive from it.
"""
from __future__ import annotations

import copy

BASIC_ROOM = {  # examples/basic_room.yaml of the reference: configuration c1
    "room": {"size": [5.0, 4.0, 2.5], "grid": [25, 20, 12]},
    "fans": [{"name": "fan0", "position": [0.5, 2.0, 1.2], "direction": [0.0, 0.0],
              "velocity": 3.0, "radius": 0.15}],
    "obstacles": [{"name": "sofa", "bbox_min": [2.0, 0.5, 0.0], "bbox_max": [3.0, 1.8, 0.9]}],
    "solver": {"dt": 0.01, "max_steps": 1000, "convergence_tol": 1.0, "rho": 1.2,
               "nu": 1.5e-5, "poisson_max_iter": 1000, "poisson_tol": 1e-6},
}

# 12x10x8 room with two openings (cross ventilation), an oblique-ish fan and
# a box: exercises QUICK, buoyancy, opening temperatures and outflow copies.
SMALL_THERMAL = {
    "room": {"size": [3.0, 2.5, 2.0], "grid": [12, 10, 8]},
    "fans": [{"name": "f", "position": [1.5, 2.1, 0.9], "direction": [270.0, 0.0],
              "velocity": 2.5, "radius": 0.2}],
    "obstacles": [{"name": "box", "bbox_min": [0.3, 0.3, 0.0], "bbox_max": [1.0, 1.2, 0.6]}],
    "openings": [{"name": "win", "wall": "north", "center": [1.5, 1.0], "size": [0.8, 0.8]},
                 {"name": "door", "wall": "south", "center": [2.2, 0.8], "size": [0.5, 1.5]}],
    "solver": {"dt": 0.02, "T_initial": 30.0, "T_outside": 22.0, "buoyancy": True,
               "advection": "quick"},
}

# Same room, Lax-Wendroff, fan blowing up, auto outflow on the ceiling.
SMALL_LW = {
    "room": {"size": [3.0, 2.5, 2.0], "grid": [12, 10, 8]},
    "fans": [{"name": "up", "position": [1.2, 1.2, 0.3], "direction": [30.0, 80.0],
              "velocity": 1.5, "radius": 0.3}],
    "obstacles": [{"name": "desk", "bbox_min": [2.0, 1.5, 0.0], "bbox_max": [2.8, 2.4, 0.7]}],
    "solver": {"dt": 0.01, "advection": "lax_wendroff"},
}


# examples/summer_cooling.yaml of the reference (30x22x12): two openings on
# opposite walls, a fan blowing -y, two boxes on the floor.
SUMMER_COOLING = {
    "room": {"size": [6.0, 4.5, 2.5], "grid": [30, 22, 12]},
    "fans": [{"name": "fan_at_window", "position": [3.0, 3.8, 0.9], "direction": [270.0, 0.0],
              "velocity": 2.5, "radius": 0.15}],
    "obstacles": [{"name": "bed", "bbox_min": [0.3, 0.3, 0.0], "bbox_max": [2.0, 2.5, 0.5]},
                  {"name": "desk", "bbox_min": [3.5, 0.3, 0.0], "bbox_max": [5.2, 1.5, 0.75]}],
    "openings": [{"name": "window_north", "wall": "north", "center": [3.0, 1.2],
                  "size": [0.9, 1.1]},
                 {"name": "door_south", "wall": "south", "center": [1.5, 1.0],
                  "size": [0.8, 2.0]}],
    "solver": {"dt": 0.05, "max_steps": 5000, "convergence_tol": 1.0, "T_initial": 32.0,
               "T_outside": 25.0, "T_target": 30.0, "buoyancy": True},
}

# The c1 room with fans whose boxes meet the walls: one at the east wall
# (its radius box reaches past y = 0 and z = Lz), one against the north wall
# reaching past x = 0, one placed above the ceiling blowing up.
FANS_AT_WALLS = {
    "room": {"size": [5.0, 4.0, 2.5], "grid": [25, 20, 12]},
    "fans": [{"name": "east", "position": [4.99, 0.12, 2.4], "direction": [0.0, 0.0],
              "velocity": 3.0, "radius": 0.3},
             {"name": "north", "position": [0.2, 3.95, 1.2], "direction": [90.0, 0.0],
              "velocity": 2.0, "radius": 0.25},
             {"name": "up", "position": [2.5, 2.0, 2.6], "direction": [0.0, 90.0],
              "velocity": 1.5, "radius": 0.15}],
    "solver": {"dt": 0.01},
}


def solver_params(cfg: dict) -> dict:
    """FluidSolverParams fields as the reference SolverBridge maps them
    (python/ff_room/solver_bridge.py:108-127, config.py defaults)."""
    s = cfg.get("solver", {})
    T0 = s.get("T_initial", 20.0)
    tt = s.get("T_target")
    adv = {"quick": 1, "lax_wendroff": 2}.get((s.get("advection") or "upwind").lower(), 0)
    return dict(
        rho=s.get("rho", 1.2), nu=s.get("nu", 1.5e-5), dt=s.get("dt", 0.01),
        max_steps=s.get("max_steps", 1000), convergence_tol=s.get("convergence_tol", 1.0),
        poisson_max_iter=s.get("poisson_max_iter", 1000), poisson_tol=s.get("poisson_tol", 1e-6),
        advection=adv, use_multigrid=int(s.get("poisson_method", "cg") == "multigrid"),
        thermal=int(bool(cfg.get("openings")) or bool(s.get("buoyancy", False)) or tt is not None),
        T_initial=T0, T_target=tt if tt is not None else 1e30,
        buoyancy=int(bool(s.get("buoyancy", False))), g_accel=9.81,
        beta=1.0 / (T0 + 273.15), k_thermal=s.get("k_thermal", 0.026), cp=s.get("cp", 1005.0))


def variant(cfg: dict, **solver) -> dict:
    c = copy.deepcopy(cfg)
    c.setdefault("solver", {}).update(solver)
    return c


def stretched(sc, ratio: float):
    """The scene on a stretched grid: Grid::set_face_coords faces of the given
    geometric ratio on every axis (reference scene.py:34-59, 82-87). The flags
    stay as they are: the reference maps positions to cells with
    int(x / L * N) whatever the faces."""
    import numpy as np
    from oracle.scene_flags import stretched_faces
    sc.faces = tuple(np.array(stretched_faces(n, L, ratio), dtype=np.float64)
                     for n, L in ((sc.nx, sc.lx), (sc.ny, sc.ly), (sc.nz, sc.lz)))
    return sc
