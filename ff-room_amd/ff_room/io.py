"""Result files: the reference's SimResult archive and exports.

``ResultStore.save_npz`` / ``load_npz`` keep the reference's archive keys
(io.py:24-64): ``u_field v_field w_field p_field T_field cell_type`` and the
scalars ``_steps _converged _velocity_change _divergence_max _T_mean
_wall_time_s _config_json`` (the scene as JSON text in a 0-d string array), so
archives move between the two in both directions. Loading never unpickles:
the reference's archives hold only numeric and string arrays, which
``numpy.load(allow_pickle=False)`` reads.
``save_velocity_csv`` (io.py:78-105) and ``ExperimentLog`` (io.py:108-141)
write the same columns and JSON lines as the reference. VTK export needs
pyvista, which this build does not ship.
"""
from __future__ import annotations

import hashlib
import json
import time
from pathlib import Path

import numpy as np

from .config import SceneConfig
from .solver_bridge import SimResult


class ResultStore:
    """Save and load SimResult objects."""

    @staticmethod
    def save_npz(result: SimResult, path: str) -> None:
        """Fields, diagnostics and the scene as one compressed archive (io.py:24-44)."""
        np.savez_compressed(
            path,
            u_field=result.u_field, v_field=result.v_field, w_field=result.w_field,
            p_field=result.p_field, T_field=result.T_field, cell_type=result.cell_type,
            _steps=result.steps, _converged=result.converged,
            _velocity_change=result.velocity_change, _divergence_max=result.divergence_max,
            _T_mean=result.T_mean, _wall_time_s=result.wall_time_s,
            _config_json=np.array(json.dumps(result.config.to_dict())))

    @staticmethod
    def load_npz(path: str) -> SimResult:
        """The SimResult of an archive written by either build (io.py:46-64);
        archives without T_field or _T_mean load zeros, as the reference's do."""
        with np.load(path, allow_pickle=False) as data:
            cfg = SceneConfig.from_dict(json.loads(str(data["_config_json"])))
            T = data["T_field"] if "T_field" in data else np.zeros_like(data["p_field"])
            return SimResult(
                u_field=data["u_field"], v_field=data["v_field"], w_field=data["w_field"],
                p_field=data["p_field"], T_field=T, cell_type=data["cell_type"], config=cfg,
                steps=int(data["_steps"]), converged=bool(data["_converged"]),
                velocity_change=float(data["_velocity_change"]),
                divergence_max=float(data["_divergence_max"]),
                T_mean=float(data["_T_mean"]) if "_T_mean" in data else 0.0,
                wall_time_s=float(data["_wall_time_s"]))

    @staticmethod
    def save_velocity_csv(result: SimResult, path: str, component: str = "all") -> None:
        """Cell-centred velocity as CSV rows x_m,y_m,z_m,u_ms,v_ms,w_ms,speed_ms
        over cells in (i, j, k) order at the uniform cell centres (io.py:78-105)."""
        r = result.config.room
        nx, ny, nz = r.grid
        lx, ly, lz = r.size
        vel = result.velocity_cell_centered()
        spd = result.velocity_magnitude()
        xs = np.linspace(lx / (2 * nx), lx * (1 - 1 / (2 * nx)), nx)
        ys = np.linspace(ly / (2 * ny), ly * (1 - 1 / (2 * ny)), ny)
        zs = np.linspace(lz / (2 * nz), lz * (1 - 1 / (2 * nz)), nz)
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
        rows = np.column_stack([X.ravel(), Y.ravel(), Z.ravel(), vel[..., 0].ravel(),
                                vel[..., 1].ravel(), vel[..., 2].ravel(), spd.ravel()])
        np.savetxt(path, rows, delimiter=",", header="x_m,y_m,z_m,u_ms,v_ms,w_ms,speed_ms",
                   comments="")

    @staticmethod
    def save_vtk(result: SimResult, path: str) -> None:
        raise ImportError("VTK export needs pyvista, which is not part of this build")


class ExperimentLog:
    """Append-only JSON Lines log: the scene's hash and a run's key metrics
    per line (io.py:108-141)."""

    def __init__(self, path: str):
        self.path = Path(path)

    def append(self, result: SimResult, notes: str = "") -> str:
        """Appends one entry; returns the first 8 hex digits of the SHA-256 of
        the scene's sorted JSON."""
        cfg = result.config.to_dict()
        cfg_hash = hashlib.sha256(json.dumps(cfg, sort_keys=True).encode()).hexdigest()[:8]
        entry = {"timestamp": time.strftime("%Y-%m-%dT%H:%M:%S"), "config_hash": cfg_hash,
                 "steps": result.steps, "converged": result.converged,
                 "velocity_change": result.velocity_change,
                 "divergence_max": result.divergence_max, "T_mean": result.T_mean,
                 "wall_time_s": result.wall_time_s, "notes": notes, "config": cfg}
        with open(self.path, "a") as f:
            f.write(json.dumps(entry) + "\n")
        return cfg_hash

    def load_all(self) -> list:
        if not self.path.exists():
            return []
        with open(self.path) as f:
            return [json.loads(line) for line in f if line.strip()]
