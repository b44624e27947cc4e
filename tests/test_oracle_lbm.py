"""Pins the D3Q19 lattice-Boltzmann restatement (oracle.c, orc_lbm_*) - CPU.

* Golden fixture lbm_small.npz, written by the reference's own LBMSolver
  (tests/golden/make_golden.py): every step's result and the final fields,
  bit for bit, on a room with two openings and on basic_room (fan, automatic
  outflow, obstacle).
* When oracle/_ref is built: against the compiled reference on a
  Lax-Wendroff room and a seeded room with several fans and obstacles.
* The reference's bounce-back quirk, as it computes it: its OPP table
  (LBMSolver.cpp:28) reverses only one component of a diagonal direction, so
  a closed box at rest does not stay at rest - cells at the walls normal
  to y gain velocity along y after one step.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle.pyoracle import INFLOW, OUTFLOW, SOLID, OracleLBM, RefLBM, SceneArrays, ref_available
from oracle.scene_flags import build_scene
from scenes import BASIC_ROOM, SMALL_LW, SMALL_THERMAL, stretched

CASES = {"thermal": (SMALL_THERMAL, 0.8, 0.01), "basic": (BASIC_ROOM, 0.6, 0.01)}


def _res(r):
    return np.array([r["step"], r["velocity_change"], r["pressure_residual"],
                     r["divergence_max"], r["converged"], r["T_mean"]])


@pytest.mark.parametrize("tag", sorted(CASES))
def test_oracle_lbm_matches_golden(tag):
    cfg, tau, dt = CASES[tag]
    g = np.load(os.path.join(GOLDEN, "lbm_small.npz"), allow_pickle=False)
    o = OracleLBM(build_scene(cfg), tau, dt, convergence_tol=0.0)
    res = np.array([_res(o.step()) for _ in range(len(g[f"{tag}_results"]))])
    assert np.array_equal(res, g[f"{tag}_results"])
    f = o.fields()
    for k in ("u", "v", "w", "p"):
        assert np.array_equal(f[k], g[f"{tag}_{k}"]), k


def _random_room():
    from ff_room.generator import random_scene
    from ff_room.scene import boundary_lists, cell_marks
    from ff_room.device import flags_from_marks
    cfg = random_scene(77, grid=(18, 14, 10), size=(4.0, 3.0, 2.5), n_fans=3, n_obstacles=3)
    fans, outs, opens = boundary_lists(cfg)
    flags = flags_from_marks((18, 14, 10), cell_marks(cfg)).reshape(-1)
    plain = [o for o in outs if tuple(o) not in {tuple(p[:6]) for p in opens}]
    return SceneArrays(18, 14, 10, 4.0, 3.0, 2.5, flags, fans, plain, opens)


def _edge_room(fan_at_outflow: bool):
    """An outflow wall, a SOLID box, an INFLOW cell no fan holds, and a fan
    inside the room or on the cells the outflow copies from."""
    nx, ny, nz = 12, 10, 8
    f = np.zeros((nx, ny, nz), dtype=np.uint8)
    f[nx - 1] = OUTFLOW
    f[4:6, 2:4, 0:3] = SOLID
    f[2, 7, 5] = INFLOW
    i = nx - 2 if fan_at_outflow else 3
    f[i, 3:6, 2:5] = INFLOW
    fans = [(i, i, 3, 5, 2, 4, 1.5, 0.2, 0.0)]
    return SceneArrays(nx, ny, nz, 2.4, 2.0, 1.6, f.reshape(-1), fans, [(0, 1, 0, ny - 1, 0, nz - 1)])


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("which", ["lw", "random", "edge", "edge_fan_at_outflow", "stretched"])
def test_oracle_lbm_matches_reference(which):
    """(stretched: the lattice itself is uniform, but the divergence it reports
    divides by the stretched widths, LBMSolver.cpp:390-392)"""
    sc = {"lw": lambda: build_scene(SMALL_LW), "random": _random_room,
          "stretched": lambda: stretched(build_scene(SMALL_THERMAL), 1.15),
          "edge": lambda: _edge_room(False), "edge_fan_at_outflow": lambda: _edge_room(True)}[which]()
    o, r = OracleLBM(sc, 0.7, 0.005, 0.0), RefLBM(sc, 0.7, 0.005, 0.0)
    for _ in range(15):
        assert o.step() == {k: v for k, v in r.step().items()}
    fo, fr = o.fields(), r.fields()
    for k in ("u", "v", "w", "p"):
        assert np.array_equal(fo[k], fr[k]), k


def test_closed_box_gains_velocity_at_y_and_z_walls():
    """The quirk, in the oracle and (when built) in the reference alike."""
    nx, ny, nz = 7, 6, 5
    flags = np.zeros((nx, ny, nz), dtype=np.uint8).reshape(-1)
    sc = SceneArrays(nx, ny, nz, 1.4, 1.2, 1.0, flags)
    o = OracleLBM(sc, 0.9, 0.01, 0.0)
    r = o.step()
    f = o.fields()
    v = f["v"].reshape(nx, ny + 1, nz)
    assert r["velocity_change"] > 0.0
    assert np.abs(v[:, 0, :]).max() > 0.0 and np.abs(v[:, ny, :]).max() > 0.0
    if ref_available():
        ref = RefLBM(sc, 0.9, 0.01, 0.0)
        assert ref.step() == r
        for k in ("u", "v", "w", "p"):
            assert np.array_equal(ref.fields()[k], f[k]), k
