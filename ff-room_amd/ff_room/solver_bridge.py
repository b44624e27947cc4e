"""SolverBridge: run a SceneConfig on the GPU core and collect the fields.

Same surface and behaviour as the reference's python/ff_room/solver_bridge.py
(SimResult, FieldSnapshot, SolverBridge.run / run_animated): the parameter
mapping of :108-127 (thermal when there are openings, buoyancy or a target
temperature; beta = 1/(T_initial+273.15)), one callback per step, and the
fields copied out after the run. The steps themselves run on the MI355X via
``_ffroom_core.FluidSolver``.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Callable, List, Optional

import numpy as np

from .config import SceneConfig
from .scene import Scene, _load_core


@dataclass
class FieldSnapshot:
    """Fields captured during a run: speed and T are (Nx,Ny,Nz), vel (Nx,Ny,Nz,3)."""
    step: int
    time_s: float
    T_mean: float
    speed: np.ndarray
    T_field: np.ndarray
    vel: np.ndarray


def _cell_centred(u: np.ndarray, v: np.ndarray, w: np.ndarray):
    return (0.5 * (u[:-1] + u[1:]), 0.5 * (v[:, :-1] + v[:, 1:]), 0.5 * (w[:, :, :-1] + w[:, :, 1:]))


@dataclass
class SimResult:
    """Final fields (reference shapes) and run diagnostics."""
    u_field: np.ndarray
    v_field: np.ndarray
    w_field: np.ndarray
    p_field: np.ndarray
    T_field: np.ndarray
    cell_type: np.ndarray
    config: SceneConfig
    steps: int = 0
    converged: bool = False
    velocity_change: float = 0.0
    divergence_max: float = 0.0
    T_mean: float = 0.0
    wall_time_s: float = 0.0

    def velocity_magnitude(self) -> np.ndarray:
        uc, vc, wc = _cell_centred(self.u_field, self.v_field, self.w_field)
        return np.sqrt(uc ** 2 + vc ** 2 + wc ** 2)

    def velocity_cell_centered(self) -> np.ndarray:
        return np.stack(_cell_centred(self.u_field, self.v_field, self.w_field), axis=-1)


def _apply_solver_flags(sc, params, c) -> None:
    """advection name -> AdvectionScheme; parallel -> use_openmp (:74-83)."""
    scheme = (sc.advection or "upwind").lower()
    params.advection = {"quick": c.AdvectionScheme.QUICK,
                        "lax_wendroff": c.AdvectionScheme.LAX_WENDROFF}.get(
                            scheme, c.AdvectionScheme.UPWIND1)
    params.use_openmp = bool(sc.parallel)


def make_params(config: SceneConfig, c):
    """FluidSolverParams for a scene, as SolverBridge.run maps them (:108-127)."""
    sc = config.solver
    p = c.FluidSolverParams()
    p.rho, p.nu, p.dt = sc.rho, sc.nu, sc.dt
    p.max_steps = sc.max_steps
    p.convergence_tol = sc.convergence_tol
    p.poisson.max_iter = sc.poisson_max_iter
    p.poisson.tol = sc.poisson_tol
    p.thermal = bool(config.openings) or bool(sc.buoyancy) or (sc.T_target is not None)
    p.T_initial = sc.T_initial
    p.T_target = sc.T_target if sc.T_target is not None else 1e30
    p.buoyancy = bool(sc.buoyancy)
    p.cp = sc.cp
    alpha = getattr(sc, "alpha", None)
    p.k_thermal = sc.k_thermal if alpha is None else alpha * sc.rho * sc.cp
    beta = getattr(sc, "beta", None)
    p.beta = 1.0 / (sc.T_initial + 273.15) if beta is None else beta
    _apply_solver_flags(sc, p, c)
    method = (getattr(sc, "poisson_method", "cg") or "cg").lower()
    p.use_multigrid = method == "multigrid"
    if method == "mgpcg":
        p.poisson_method = c.PoissonMethod.MGPCG
    prec = (getattr(sc, "precision", "fp64") or "fp64").lower()
    if prec not in ("fp64", "mixed"):
        raise ValueError(f"unknown precision {prec!r}")
    p.precision = c.Precision.MIXED if prec == "mixed" else c.Precision.FP64
    p.mg_levels = int(getattr(sc, "mg_levels", 0) or 0)
    p.mg_bf16 = bool(getattr(sc, "mg_bf16", True))
    p.inner_tol = float(getattr(sc, "inner_tol", 1e-4))
    p.inner_max_iter = int(getattr(sc, "inner_max_iter", 60))
    return p


class SolverBridge:
    """Builds the scene, runs the GPU FluidSolver, returns a SimResult."""

    def __init__(self, config: SceneConfig):
        self.config = config
        self._c = _load_core()

    def _solver(self):
        scene = Scene(self.config)
        params = make_params(self.config, self._c)
        sc = self.config.solver
        if sc.method == "lbm":  # solver_bridge.py:129-130: the D3Q19 lattice, isothermal
            params.thermal = False
            return scene, params, self._c.LBMSolver(scene.grid, scene.bm, sc.tau, sc.max_steps,
                                                    sc.convergence_tol, sc.dt)
        return scene, params, self._c.FluidSolver(scene.grid, scene.bm, params)

    @staticmethod
    def _line(res, thermal: bool) -> str:
        s = (f"  step={res.step:4d}  vel_change={res.velocity_change:.3e}"
             f"  p_res={res.pressure_residual:.3e}  div_max={res.divergence_max:.3e}")
        if thermal:
            s += f"  T_mean={res.T_mean:.2f}°C"
        return s + ("  [CONVERGED]" if res.converged else "")

    def _result(self, g, final, wall: float) -> SimResult:
        return SimResult(
            u_field=np.array(g.u_shaped()), v_field=np.array(g.v_shaped()),
            w_field=np.array(g.w_shaped()), p_field=np.array(g.p_shaped()),
            T_field=np.array(g.T_shaped()), cell_type=g.cell_type_shaped(),
            config=self.config, steps=final.step, converged=final.converged,
            velocity_change=final.velocity_change, divergence_max=final.divergence_max,
            T_mean=final.T_mean, wall_time_s=wall)

    def run(self, progress_callback: Optional[Callable] = None,
            print_interval: int = 50) -> SimResult:
        """Run to convergence or max_steps; progress_callback(step, vel_change,
        div_max, T_mean) is called after every step."""
        scene, params, solver = self._solver()
        g = scene.grid

        def _cb(res):
            if progress_callback:
                progress_callback(res.step, res.velocity_change, res.divergence_max, res.T_mean)
            if print_interval > 0 and res.step % print_interval == 0:
                print(self._line(res, params.thermal))

        t0 = time.perf_counter()
        final = solver.run(_cb)
        wall = time.perf_counter() - t0
        return self._result(g, final, wall)

    def run_animated(self, save_every: int = 50, progress_callback: Optional[Callable] = None,
                     print_interval: int = 50):
        """Like run(), also capturing a FieldSnapshot every ``save_every``
        steps and at the end. Returns (SimResult, list[FieldSnapshot]).

        A capture is queued on the device (FluidSolver.snapshot_begin): the
        device computes its cell-centred velocity and speed in FP64 as the
        reference does on the host (solver_bridge.py:224-236), the copy to the
        host overlaps the next steps, and the arrays are collected later,
        keeping at most ``_IN_FLIGHT`` captures pending."""
        scene, params, solver = self._solver()
        g = scene.grid
        dt = self.config.solver.dt
        snaps: List[FieldSnapshot] = []
        pending: List[tuple] = []  # (slot, T_mean) in capture order

        def make(step: int, T_mean: float, u, v, w, T) -> FieldSnapshot:
            uc, vc, wc = _cell_centred(u, v, w)
            T = np.asarray(T) if params.thermal else np.zeros(uc.shape)
            return FieldSnapshot(step=step, time_s=step * dt, T_mean=T_mean,
                                 speed=np.sqrt(uc ** 2 + vc ** 2 + wc ** 2), T_field=T,
                                 vel=np.stack([uc, vc, wc], axis=-1))

        def collect():
            slot, T_mean = pending.pop(0)
            step, vel, speed, T = solver.snapshot_end(slot)
            snaps.append(FieldSnapshot(step=step, time_s=step * dt, T_mean=T_mean, speed=speed,
                                       T_field=T if params.thermal else np.zeros(speed.shape),
                                       vel=vel))

        queued = hasattr(solver, "snapshot_begin")  # the lattice captures from the Grid

        def _cb(res):
            if res.step % save_every == 0:
                if not queued:
                    snaps.append(make(res.step, res.T_mean, np.array(g.u_shaped()),
                                      np.array(g.v_shaped()), np.array(g.w_shaped()),
                                      np.array(g.T_shaped())))
                else:
                    if len(pending) >= self._IN_FLIGHT:
                        collect()
                    pending.append((solver.snapshot_begin(cells=True), res.T_mean))
            if progress_callback:
                progress_callback(res.step, res.velocity_change, res.divergence_max, res.T_mean)
            if print_interval > 0 and res.step % print_interval == 0:
                print(self._line(res, params.thermal))

        t0 = time.perf_counter()
        final = solver.run(_cb)
        wall = time.perf_counter() - t0
        while pending:
            collect()
        if not snaps or snaps[-1].step != final.step:
            snaps.append(make(final.step, final.T_mean, np.array(g.u_shaped()),
                              np.array(g.v_shaped()), np.array(g.w_shaped()),
                              np.array(g.T_shaped())))
        return self._result(g, final, wall), snaps

    _IN_FLIGHT = 3  # pending captures (the library holds up to FFR_SNAPSHOT_SLOTS)
