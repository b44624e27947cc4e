"""Writes the golden fixtures of this directory from the reference's own
solver classes, compiled from the reference sources by oracle/Makefile into
oracle/_ref/libffref.so (with -ffp-contract=off, serial).

    make -C oracle ref && python tests/golden/make_golden.py

Fixtures (numpy .npz, FP64):
  c1_upwind_cg.npz     basic_room (configuration c1): fields after steps 1, 10, 100
                       and 500 (the configured run length)
  small_thermal.npz    12x10x8 QUICK + buoyancy + two openings, 20 steps
  small_lw.npz         12x10x8 Lax-Wendroff, 10 steps
  poisson_tiny.npz     6x5x4 mixed-flag grid: rhs, CG p, multigrid p
  couette.npz          the reference's Couette test case run to convergence
  lbm_small.npz        LBMSolver (D3Q19) on small_thermal's room and on
                       basic_room: fields after 25 steps, every step's result

    python tests/golden/make_golden.py lbm     # only lbm_small.npz
    python tests/golden/make_golden.py c1      # only c1_upwind_cg.npz
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle.pyoracle import RefLBM, RefSolver, SceneArrays, ref_available  # noqa: E402
from oracle.scene_flags import build_scene  # noqa: E402
from scenes import BASIC_ROOM, SMALL_LW, SMALL_THERMAL, solver_params, variant  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def _res(r: dict) -> np.ndarray:
    return np.array([r["step"], r["velocity_change"], r["pressure_residual"],
                     r["divergence_max"], r["converged"], r["T_mean"]])


def trajectory(cfg, snap_steps, name):
    sc = build_scene(cfg)
    params = solver_params(cfg)
    params["convergence_tol"] = 0.0  # never stop early
    ref = RefSolver(sc, params)
    out = {}
    for s in range(1, max(snap_steps) + 1):
        r = ref.step()
        if s in snap_steps:
            f = ref.fields()
            for k, a in f.items():
                out[f"{k}_{s}"] = a
            out[f"result_{s}"] = _res(r)
    out["flags"] = sc.flags
    np.savez_compressed(os.path.join(OUT, name), **out)


def poisson_tiny():
    nx, ny, nz = 6, 5, 4
    rng = np.random.default_rng(7)
    flags = np.zeros((nx, ny, nz), dtype=np.uint8)
    flags[2:4, 1:3, 0:2] = 1        # SOLID block
    flags[1, 3, 2] = 2              # INFLOW cell
    flags[nx - 1, :, :] = 3         # OUTFLOW wall
    flags[0, 0, 3] = 3              # isolated OUTFLOW cell
    flags = flags.reshape(-1)
    rhs = rng.standard_normal(nx * ny * nz)
    rhs[flags == 1] = 0.0
    sc = SceneArrays(nx, ny, nz, 1.2, 1.0, 0.8, flags)
    ref = RefSolver(sc, {})
    out = {"flags": flags, "rhs": rhs}
    for mg, tag in ((0, "cg"), (1, "mg")):
        res, it = ref.poisson(rhs, bool(mg), 1000, 1e-10)
        out[f"p_{tag}"] = ref.fields()["p"]
        out[f"res_{tag}"] = np.array([res, it])
    np.savez_compressed(os.path.join(OUT, "poisson_tiny.npz"), **out)


def couette():
    # cpp/tests/test_couette.cpp:19-64: 4x4x10 unit box, top cell layer is a
    # "fan" moving at U=1 in x, dt 0.005, <=2000 steps, tol 1e-5, Poisson tol 1e-7
    nx, ny, nz = 4, 4, 10
    flags = np.zeros((nx, ny, nz), dtype=np.uint8)
    flags[:, :, nz - 1] = 2
    sc = SceneArrays(nx, ny, nz, 1.0, 1.0, 1.0, flags.reshape(-1),
                     fans=[(0, nx - 1, 0, ny - 1, nz - 1, nz - 1, 1.0, 0.0, 0.0)])
    params = dict(dt=0.005, max_steps=2000, convergence_tol=1e-5, nu=1.5e-5, poisson_tol=1e-7)
    ref = RefSolver(sc, params)
    r = ref.run()
    f = ref.fields()
    np.savez_compressed(os.path.join(OUT, "couette.npz"), result=_res(r), **f)


C1_STEPS = (1, 10, 100, 500)
LBM_CASES = (("thermal", SMALL_THERMAL, 0.8, 0.01), ("basic", BASIC_ROOM, 0.6, 0.01))
LBM_STEPS = 25


def lbm_small():
    out = {}
    for tag, cfg, tau, dt in LBM_CASES:
        sc = build_scene(cfg)
        ref = RefLBM(sc, tau, dt, convergence_tol=0.0)
        out[f"{tag}_results"] = np.array([_res(ref.step()) for _ in range(LBM_STEPS)])
        for k, a in ref.fields().items():
            if k != "T":
                out[f"{tag}_{k}"] = a
    np.savez_compressed(os.path.join(OUT, "lbm_small.npz"), **out)


def main():
    if not ref_available():
        sys.exit("oracle/_ref/libffref.so is not built (make -C oracle ref)")
    if sys.argv[1:] == ["lbm"]:
        lbm_small()
        return
    if sys.argv[1:] == ["c1"]:
        trajectory(BASIC_ROOM, C1_STEPS, "c1_upwind_cg.npz")
        return
    trajectory(BASIC_ROOM, C1_STEPS, "c1_upwind_cg.npz")
    trajectory(SMALL_THERMAL, (20,), "small_thermal.npz")
    trajectory(SMALL_LW, (10,), "small_lw.npz")
    poisson_tiny()
    couette()
    lbm_small()
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()
