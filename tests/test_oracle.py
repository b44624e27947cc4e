"""Pins the FP64 oracle (oracle/oracle.c) - CPU only.

* Golden fixtures written by the reference's own solver classes
  (tests/golden/make_golden.py): the oracle must reproduce them bit for bit.
* Known answers: a manufactured Poisson solution with second-order
  convergence (the corrected form of cpp/tests/test_poisson.cpp:17-62) and
  the reference's Couette case (cpp/tests/test_couette.cpp:19-64) as the
  reference computes it.
* Properties the mathematics fixes for the pressure operator, checked on a
  dense matrix assembled column by column (brute force on tiny grids).
* When oracle/_ref is built: the oracle against the compiled reference on
  further scenes, including the reference multigrid.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle.pyoracle import (INFLOW, OUTFLOW, SOLID, OracleSolver, RefSolver, SceneArrays,
                             ref_available)
from oracle.scene_flags import build_scene
from scenes import BASIC_ROOM, SMALL_LW, SMALL_THERMAL, solver_params, stretched, variant


def _golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.mark.parametrize("cfg,name,steps", [
    (BASIC_ROOM, "c1_upwind_cg.npz", (1, 10, 100, 500)),
    (SMALL_THERMAL, "small_thermal.npz", (20,)),
    (SMALL_LW, "small_lw.npz", (10,)),
])
def test_oracle_reproduces_reference_bitwise(cfg, name, steps):
    g = _golden(name)
    sc = build_scene(cfg)
    assert np.array_equal(sc.flags, g["flags"])
    params = solver_params(cfg)
    params["convergence_tol"] = 0.0
    o = OracleSolver(sc, params)
    for s in range(1, max(steps) + 1):
        r = o.step()
        if s in steps:
            f = o.fields()
            for k in ("u", "v", "w", "p", "T"):
                assert np.array_equal(f[k], g[f"{k}_{s}"]), (name, s, k)
            ref_r = g[f"result_{s}"]
            assert r["step"] == ref_r[0]
            assert r["velocity_change"] == ref_r[1]
            assert r["pressure_residual"] == ref_r[2]
            assert r["divergence_max"] == ref_r[3]
            assert r["T_mean"] == ref_r[5]


def test_oracle_poisson_matches_reference_fixture():
    g = _golden("poisson_tiny.npz")
    sc = SceneArrays(6, 5, 4, 1.2, 1.0, 0.8, g["flags"])
    o = OracleSolver(sc, {}, init=False)
    res, it = o.poisson_cg(g["rhs"], 1000, 1e-10)
    assert np.array_equal(o.p, g["p_cg"])
    assert (res, it) == (g["res_cg"][0], g["res_cg"][1])
    res, it = o.poisson_mg(g["rhs"], 1000, 1e-10)
    assert np.array_equal(o.p, g["p_mg"])
    assert (res, it) == (g["res_mg"][0], g["res_mg"][1])


def test_couette_case_matches_reference_fixture():
    """The reference's Couette case (cpp/tests/test_couette.cpp:19-64).

    Its closed box makes the driven top layer a lid-driven cavity, and with
    nu = 1.5e-5 the run ends at max_steps = 2000 unconverged; the reference
    itself does not meet that test's linear-profile assertion. What is pinned
    here is the reference's actual output."""
    nx, ny, nz = 4, 4, 10
    flags = np.zeros((nx, ny, nz), dtype=np.uint8)
    flags[:, :, nz - 1] = INFLOW
    sc = SceneArrays(nx, ny, nz, 1.0, 1.0, 1.0, flags.reshape(-1),
                     fans=[(0, nx - 1, 0, ny - 1, nz - 1, nz - 1, 1.0, 0.0, 0.0)])
    params = dict(dt=0.005, max_steps=2000, convergence_tol=1e-5, nu=1.5e-5, poisson_tol=1e-7)
    o = OracleSolver(sc, params)
    for _ in range(2000):
        r = o.step()
        if r["converged"]:
            break
    g = _golden("couette.npz")
    assert r["step"] == g["result"][0] == 2000 and not r["converged"]
    for k in ("u", "v", "w", "p"):
        assert np.array_equal(getattr(o, k), g[k]), k


def manufactured(N):
    """A manufactured solution of the reference's operator L = -lap
    (PoissonSolver.cpp:14-21) on the unit cube whose boundary ring of cells is
    OUTFLOW: the unknown interior sees p = 0 on the ring's inner faces, at
    x = h and 1-h, so p = prod sin(pi (x-h)/(1-2h)) and L p = 3 (pi/(1-2h))^2 p.
    The reference's own version of this test (cpp/tests/test_poisson.cpp:17-62)
    uses the opposite sign and the cube's outer faces, and fails on the
    reference itself (max error 1.6 at N=8)."""
    h = 1.0 / N
    idx = np.indices((N, N, N))
    ring = np.any((idx == 0) | (idx == N - 1), axis=0).reshape(-1)
    flags = np.where(ring, OUTFLOW, 0).astype(np.uint8)
    k = np.pi / (1 - 2 * h)
    x, y, z = [((a + 0.5) * h - h).reshape(-1) for a in idx]
    pe = np.sin(k * x) * np.sin(k * y) * np.sin(k * z)
    rhs = np.where(ring, 0.0, 3 * k * k * pe)
    return SceneArrays(N, N, N, 1.0, 1.0, 1.0, flags), rhs, pe, ~ring


def test_poisson_manufactured_solution_second_order():
    # (at N=8 the reference CG stagnates near 3e-6 on this non-symmetric
    # system - OUTFLOW rows read their neighbours - so the test starts at 16)
    errs = []
    for N in (16, 32):
        sc, rhs, pe, inner = manufactured(N)
        o = OracleSolver(sc, {}, init=False)
        res, _ = o.poisson_cg(rhs, 1000, 1e-10)
        assert res < 1e-9
        errs.append(np.max(np.abs(o.p[inner] - pe[inner])))
    assert errs[0] < 0.01
    assert errs[0] / errs[1] > 3.5  # O(h^2)


def _mixed_grid(seed=3, nx=5, ny=4, nz=3):
    rng = np.random.default_rng(seed)
    f = rng.choice([0, 0, 0, 0, SOLID, INFLOW], size=(nx, ny, nz)).astype(np.uint8)
    f[-1, :, :] = OUTFLOW
    return SceneArrays(nx, ny, nz, 1.0, 0.8, 0.6, f.reshape(-1))


def _dense(o: OracleSolver, n):
    A = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        A[:, c] = o.laplacian(e)
    return A


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pressure_operator_structure(seed):
    sc = _mixed_grid(seed)
    n = sc.nx * sc.ny * sc.nz
    o = OracleSolver(sc, {}, init=False)
    A = _dense(o, n)
    f = sc.flags
    solid, outf = f == SOLID, f == OUTFLOW
    unk = ~solid & ~outf
    # SOLID rows are the identity (PoissonSolver.cpp:52-55)
    assert np.array_equal(A[solid], np.eye(n)[solid])
    # unknown rows never read OUTFLOW or SOLID values; OUTFLOW rows do read unknowns
    assert np.all(A[np.ix_(unk, outf | solid)] == 0.0)
    assert np.any(A[np.ix_(outf, unk)] != 0.0)
    # uniform spacing: the unknown block is symmetric with a 1/h^2 coupling
    B = A[np.ix_(unk, unk)]
    assert np.allclose(B, B.T, rtol=0, atol=1e-12)
    hx, hy, hz = sc.lx / sc.nx, sc.ly / sc.ny, sc.lz / sc.nz
    offs = set(np.round(-B[B < 0], 9))
    assert offs <= {round(1 / hx ** 2, 9), round(1 / hy ** 2, 9), round(1 / hz ** 2, 9)}
    # an unknown cell whose neighbours are all unknowns (or walls) has zero row sum
    idx = np.arange(n).reshape(sc.nx, sc.ny, sc.nz)
    for c in np.flatnonzero(unk):
        i, j, k = np.unravel_index(c, (sc.nx, sc.ny, sc.nz))
        nb = [idx[a, b, d] for a, b, d in ((i - 1, j, k), (i + 1, j, k), (i, j - 1, k),
                                            (i, j + 1, k), (i, j, k - 1), (i, j, k + 1))
              if 0 <= a < sc.nx and 0 <= b < sc.ny and 0 <= d < sc.nz]
        if all(unk[q] for q in nb) and nb:
            assert abs(A[c].sum()) < 1e-9


@pytest.mark.parametrize("seed", [1, 2])
def test_cg_solves_the_dense_system(seed):
    sc = _mixed_grid(seed, 6, 5, 4)
    n = sc.nx * sc.ny * sc.nz
    o = OracleSolver(sc, {}, init=False)
    A = _dense(o, n)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(n)
    b[sc.flags == SOLID] = 0.0
    # the reference CG runs on a non-symmetric matrix (OUTFLOW rows read their
    # neighbours) and may stagnate near 1e-8; it still solves this system
    res, it = o.poisson_cg(b, 5000, 1e-10)
    x_direct = np.linalg.solve(A, b)
    assert res < 1e-7
    assert np.max(np.abs(o.p - x_direct)) < 1e-6 * np.abs(x_direct).max()


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("cfg,extra,steps", [
    (BASIC_ROOM, dict(advection=1), 12),
    (BASIC_ROOM, dict(use_multigrid=1), 6),
    (SMALL_THERMAL, dict(advection=2), 15),
    (SMALL_THERMAL, dict(use_multigrid=1), 8),
    (SMALL_LW, dict(advection=1, thermal=1, buoyancy=1), 10),
])
def test_oracle_against_compiled_reference(cfg, extra, steps):
    sc = build_scene(cfg)
    params = solver_params(cfg)
    params.update(extra, convergence_tol=0.0)
    o, r = OracleSolver(sc, params), RefSolver(sc, params)
    for _ in range(steps):
        a, b = o.step(), r.step()
        assert a["velocity_change"] == b["velocity_change"]
        assert a["divergence_max"] == b["divergence_max"]
    fo, fr = o.fields(), r.fields()
    for k in fo:  # the reference multigrid diverges to NaN on basic_room; so must the oracle
        assert np.array_equal(fo[k], fr[k], equal_nan=True), k


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("cfg,ratio,extra,steps", [
    (BASIC_ROOM, 1.08, dict(), 10),
    (BASIC_ROOM, 1.15, dict(advection=1), 6),
    (SMALL_THERMAL, 1.2, dict(advection=2), 10),
    (SMALL_LW, 1.1, dict(advection=1, thermal=1, buoyancy=1), 8),
])
def test_oracle_against_compiled_reference_stretched(cfg, ratio, extra, steps):
    """Stretched grids (Grid::set_face_coords): every variable-spacing term of
    the step and the non-symmetric pressure operator, bit for bit."""
    sc = stretched(build_scene(cfg), ratio)
    dxv = sc.spacings()[0]
    assert dxv.max() / dxv.min() > 1.2  # the spacing really varies
    params = solver_params(cfg)
    params.update(extra, convergence_tol=0.0)
    o, r = OracleSolver(sc, params), RefSolver(sc, params)
    for _ in range(steps):
        a, b = o.step(), r.step()
        assert all(a[k] == b[k] for k in b), (a, b)
    fo, fr = o.fields(), r.fields()
    for k in fo:
        assert np.array_equal(fo[k], fr[k]), k


def test_volume_weighted_operator_is_symmetric():
    """On a stretched grid the reference operator is not symmetric (its face
    scale 1/(h_f d_c) depends on the cell's own width); scaled row by row by
    the cell volume relative to the mean cell's, its unknown block is, and a
    face's coupling is t w w: t = 1/(centre distance x mean width) of the face
    axis, w = width / mean width of the two other axes. The multigrid's inner
    solve runs on exactly this form."""
    rng = np.random.default_rng(7)
    nx, ny, nz = 9, 7, 6
    f = rng.choice([0] * 8 + [SOLID, INFLOW], size=(nx, ny, nz)).astype(np.uint8)
    f[:, -1, :] = OUTFLOW
    sc = stretched(SceneArrays(nx, ny, nz, 1.8, 1.4, 1.2, f.reshape(-1)), 1.25)
    o = OracleSolver(sc, {}, init=False)
    n = nx * ny * nz
    A = np.stack([o.laplacian(np.eye(n)[c]) for c in range(n)], axis=1)
    d = sc.spacings()
    w = [h / h.mean() for h in d]
    t = []
    for h in d:
        ta = np.zeros(len(h) + 1)
        ta[1:-1] = 1.0 / (0.5 * (h[:-1] + h[1:]) * h.mean())
        t.append(ta)
    vol = (w[0][:, None, None] * w[1][None, :, None] * w[2][None, None, :]).reshape(-1)
    unk = np.isin(f.reshape(-1), [0, INFLOW])
    Au = A[np.ix_(unk, unk)]
    assert np.abs(Au - Au.T).max() > 1.0  # the reference operator itself is not symmetric
    M = (vol[:, None] * A)[np.ix_(unk, unk)]
    assert np.abs(M - M.T).max() < 1e-14 * np.abs(M).max()
    I, J, K = np.unravel_index(np.arange(n), (nx, ny, nz))
    for c in np.flatnonzero(unk):
        i, j, k = I[c], J[c], K[c]
        for a, (nb, ok) in enumerate(((c + ny * nz, i + 1 < nx), (c + nz, j + 1 < ny),
                                      (c + 1, k + 1 < nz))):
            if not ok or not unk[nb]:
                continue
            idx = (i, j, k)
            others = [w[b][idx[b]] for b in range(3) if b != a]
            expect = t[a][idx[a] + 1] * others[0] * others[1]
            assert -vol[c] * A[c, nb] == pytest.approx(expect, rel=1e-14)


def test_planted_mistake_is_caught():
    """A sign flip planted in the oracle's output must break the fixture test."""
    g = _golden("c1_upwind_cg.npz")
    sc = build_scene(BASIC_ROOM)
    params = solver_params(BASIC_ROOM)
    o = OracleSolver(sc, params)
    o.step()
    u = o.u.copy()
    nz_faces = np.flatnonzero(u)
    u[nz_faces[len(nz_faces) // 2]] *= -1.0
    assert not np.array_equal(u, g["u_1"])
