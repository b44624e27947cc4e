"""Benchmark: ff-room projection steps on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3] [--mode mixed|fp64]

A step is one full projection time step (FluidSolver::step: predictor,
buoyancy, masks, pressure solve, projection, temperature, diagnostics) of the
named synthetic workload, with all state resident in HBM. Prints one JSON
line: whole-job cell-updates/s of the timed steps plus
  * roofline of the dominant kernel (algorithmic HBM bytes / mean kernel
    duration from dispatch begin/end HIP events on the scene's stream, against
    8 TB/s; HBM traffic from the committed rocprofv3 counter summary),
  * the plain FP64 CG form (the reference algorithm on the GPU) timed on the
    same workload, and the pressure error of each form against a tightly
    converged FP64 solve of the same right-hand side,
  * the reference's own solver (oracle/_ref, compiled from its sources) timed
    on this host's cores for a bounded sample of the workload.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ff-room_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name: str, precision: str, stretch: float = 1.0):
    cfg, desc = _workload(name, precision)
    if stretch != 1.0:  # Grid::set_face_coords faces, finer towards every wall
        cfg.room.stretch = stretch
        desc += f", stretched grid (ratio {stretch} per axis)"
    return cfg, desc


def _workload(name: str, precision: str):
    from ff_room.generator import workload_c2, workload_c3
    if name == "c2":
        cfg = workload_c2(precision=precision)
        desc = "c2: 10x8x3 m room on 256x204x80, 2 fans (seed 2), 4 boxes, QUICK + buoyancy, dt 0.005"
    elif name == "c3":
        cfg = workload_c3(precision=precision)
        desc = "c3: 10x8x3 m room on 512x408x160, 2 fans (seed 3), 4 boxes, QUICK + buoyancy, dt 0.0025"
    else:
        raise SystemExit(f"unknown workload {name}")
    return cfg, desc


def timed_steps(d, ctx, steps: int, warmup: int, prof: bool):
    for _ in range(warmup):
        d.step()
    if prof:
        ctx.prof_reset()
        ctx.prof_enable(True)
    t0 = time.perf_counter()
    res = [d.step() for _ in range(steps)]
    t1 = time.perf_counter()
    table = None
    if prof:
        table = ctx.prof_read()
        ctx.prof_enable(False)
    return t1 - t0, res, table


TIMED_KERNELS = ("predict", "correct", "temperature", "rhs", "cg_spmv", "cg_update", "cg_dir",
                 "mg_smooth_l0", "mg_resid0", "pcg_dirq", "pcg_update", "dc_residual")


def traffic_from_profiles(kernel: str, workload: str):
    """HBM-side bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, calibrated on a copy kernel), when a summary is committed."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def roofline(table: dict, d, workload: str):
    """Dominant kernel by total time in the profiled steps; its mean duration
    over 200 launches, each timed by its own dispatch begin/end events on the
    scene's stream (the duration a kernel trace reports)."""
    name, e = max(table.items(), key=lambda kv: kv[1]["total_ms"])
    mean_events_s = e["total_ms"] / e["launches"] / 1e3
    bytes_per = e["bytes_per_launch"]
    mean_s, how = mean_events_s, "per-launch events during the steps"
    if name in TIMED_KERNELS:
        ms, b = d.prof_kernel(name, 200)
        mean_s, bytes_per, how = ms / 1e3, b, "200 launches, dispatch begin/end events"
    ach = bytes_per / mean_s
    return name, {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                  "frac": ach / HBM_PEAK, "traffic": traffic_from_profiles(name, workload),
                  "kernel": name, "bytes_per_launch": bytes_per,
                  "mean_launch_us": mean_s * 1e6, "timing": how,
                  "launches_per_step": None}


# residual tolerance of the plain FP64 CG form per workload: the loosest power
# of ten whose pressure error (pressure_errors) meets 1e-6, measured on
# MI355X: c2 4.6e-7 at 1e-8 (7.6e-6 at 1e-7); c3 7.7e-8 at 1e-9 (1.01e-6 at
# 1e-8); the c4 rooms 5.2e-6 at 1e-8 (worst of the first 8). The error is
# about the condition number times the residual.
FP64_TOL = {"c2": 1e-8, "c3": 1e-9, "c4": 1e-9}


def form_params_inplace(p, mode, workload="c2"):
    """The two forms compared: the reference algorithm (unpreconditioned FP64
    CG, run tightly enough that its pressure error meets 1e-6) and the mixed
    form (FP32 stencils, FP32/BF16 MG-PCG inner solves, FP64 outer defect
    correction)."""
    from ff_room import capi
    if mode == "fp64":
        p.precision, p.poisson_method = capi.PREC_FP64, capi.POISSON_CG
        p.poisson_tol, p.poisson_max_iter = FP64_TOL.get(workload, 1e-9), 200000
    else:
        p.precision, p.poisson_method = capi.PREC_MIXED, capi.POISSON_MGPCG
        p.poisson_tol, p.poisson_max_iter = 1e-10, 50


def sealed_components(ps):
    """Labels of the components of unknown cells (FLUID, INFLOW; face
    neighbours) that touch no OUTFLOW cell, e.g. a fan sealed inside a box.
    On such a component A is singular (zero-flux walls all round), so p is
    defined there only up to a constant, which the projection never sees
    (it uses p's differences)."""
    from scipy import ndimage
    fl = ps.flags.reshape(ps.nx, ps.ny, ps.nz)
    unk = (fl == 0) | (fl == 2)
    lab, n = ndimage.label(unk)
    touch = set(np.unique(lab[ndimage.binary_dilation(fl == 3) & unk]).tolist())
    return lab.reshape(-1), [c for c in range(1, n + 1) if c not in touch]


def gauge_free_error(x, xs, ps, sealed):
    """Relative L2 of x - xs over non-SOLID cells, after removing on each
    sealed component the mean of x - xs there (the constant A cannot see)."""
    lab, comps = sealed
    d = x - xs
    for c in comps:
        m = lab == c
        d[m] -= d[m].mean()
    keep = ps.flags != 1
    return float(np.linalg.norm(d[keep]) / np.linalg.norm(xs[keep]))


def pressure_errors(ctx, ps, ps64, d_mixed, sealed):
    """Relative L2 pressure error of both forms on a realistic system: x* is
    the pressure of the mixed scene's last step, rhs = A x* with the FP64 GPU
    operator, and each form solves A x = rhs from x = 0. SOLID cells are
    excluded; on sealed components (no OUTFLOW contact, where A is singular)
    each form's constant offset is removed first."""
    from ff_room import device
    xs = d_mixed.download()["p"]
    rhs = d_mixed.laplacian(xs)
    out = {"reference": "x* = p after the timed mixed steps; rhs = A x* (FP64 operator)",
           "sealed_cells": int(sum((sealed[0] == c).sum() for c in sealed[1]))}
    for name, d in (("mixed", d_mixed), ("fp64_cg", None)):
        if d is None:
            d = device.create(ctx, ps64)
        x, info = d.poisson(rhs)
        out[name] = {"rel_l2": gauge_free_error(x, xs, ps, sealed),
                     "iterations": info["iterations"],
                     "outer": info["outer_iterations"], "residual": info["residual"]}
        if name == "fp64_cg":
            d.close()
    return out


def step_errors(ctx, ps64, d_mixed, sealed):
    """One projection step of each form from the same state (the mixed
    scene's state after the timed steps, FP32 values exact in FP64): relative
    L2 of the mixed form's p (gauge-free as above), u, v, w and T against the
    all-FP64 form with its CG at the FP64 tolerance."""
    from ff_room import device
    st = d_mixed.download()
    d64 = device.create(ctx, ps64)
    d64.upload(u=st["u"], v=st["v"], w=st["w"], p=st["p"], T=st["T"])
    r64 = d64.step()
    ref = d64.download()
    d64.close()
    d_mixed.step()
    got = d_mixed.download()
    out = {"p": gauge_free_error(got["p"].copy(), ref["p"], ps64, sealed),
           "fp64_cg_iterations": r64["poisson_iters"]}
    for k in ("u", "v", "w", "T"):
        out[k] = float(np.linalg.norm(got[k] - ref[k]) / max(np.linalg.norm(ref[k]), 1e-300))
    return out


def serve_bench(ctx, n_c4: int, steps_c4: int, n_c5: int, steps_c5: int, gangs: int = 0,
                slots: int = 6):
    """Continuous batching samples: c4 (uneven rooms all submitted at once)
    and c5 (a seeded Poisson stream at 50 requests/s), mixed MG-PCG form;
    `gangs` groups of merged launches (0: every scene on its own stream).
    c4 is one batch, admitted largest work first (measured on MI355X, 32
    rooms x 60 steps: 3.6-3.7e8 cell-updates/s against 3.0-3.4e8 in arrival
    order, whose last large rooms run alone); c5 in arrival order, and again
    with the smallest requests admitted first. `slots` scenes are resident at
    once: 6 measured best on MI355X (c4 2.0e8 cell-updates/s, p50 0.66 s
    against 1.9e8, 0.81 s with 16; c5 the same throughput, p50 3.0 s against
    3.9 s)."""
    from ff_room import capi
    from ff_room.generator import workload_c4, workload_c5
    from ff_room.serve import Request, Server, summary
    res = {}
    kw = dict(precision=capi.PREC_MIXED, poisson_method=capi.POISSON_MGPCG,
              poisson_tol=1e-10, poisson_max_iter=500)
    cfgs = workload_c4(n_scenes=n_c4, steps=steps_c4, precision="mixed")
    srv = Server(ctx, slots=slots, gangs=gangs, policy="longest", **kw)
    done = srv.run([Request(c, steps_c4) for c in cfgs], realtime=False)
    res["c4"] = dict(summary(done), gangs=gangs, gang_stats=ctx.gang_stats(),
                     sample=f"{n_c4} seeded rooms (grids 32x32x16..160x128x64), "
                     f"{steps_c4} steps each, all submitted at t=0, {slots} slots, "
                     f"largest work admitted first")
    stream = list(workload_c5(n_requests=n_c5, steps=steps_c5, precision="mixed"))
    srv = Server(ctx, slots=slots, gangs=gangs, **kw)
    done = srv.run([Request(c, steps_c5, t) for t, c in stream], realtime=True)
    res["c5"] = dict(summary(done), gangs=gangs, gang_stats=ctx.gang_stats(),
                     sample=f"first {n_c5} requests of the seeded stream "
                     f"(50 req/s, grids 24^3..256x204x80, <= {steps_c5} steps or "
                     f"max|du|/dt < 1e-4), {slots} slots")
    srv = Server(ctx, slots=slots, gangs=gangs, policy="shortest", **kw)
    done = srv.run([Request(c, steps_c5, t) for t, c in stream], realtime=True)
    res["c5_shortest_first"] = dict(summary(done), gangs=gangs,
                                    sample="the c5 sample, admitted smallest work (cells x step "
                                           "budget) first")
    ctx.serve_gangs(0)
    return res


def batched_forms(ctx, n_rooms: int, steps: int, n_err: int, slots: int = 6, repeats: int = 3):
    """The north-star comparison in its batched form: the same c4 rooms served
    with continuous batching (all submitted at t = 0) once in the mixed form
    (FP32 stencils, MG-PCG with FP64 correction to 1e-10) and once in the
    plain FP64 CG form (the reference algorithm, CG to the loosest tolerance
    whose pressure error meets 1e-6); aggregate cell-updates/s of each and
    their ratio. Pressure errors of both forms on the first `n_err` rooms:
    x* = the room's pressure after the mixed run, rhs = A x* with the FP64
    operator, each form solving from zero (gauge-free on sealed pockets)."""
    from types import SimpleNamespace
    from ff_room import capi, device
    from ff_room.generator import workload_c4
    from ff_room.serve import Request, Server, summary
    cfgs = workload_c4(n_scenes=n_rooms, steps=steps, precision="mixed")
    forms = {
        "mixed": dict(precision=capi.PREC_MIXED, poisson_method=capi.POISSON_MGPCG,
                      poisson_tol=1e-10, poisson_max_iter=500),
        "fp64_cg": dict(precision=capi.PREC_FP64, poisson_method=capi.POISSON_CG,
                        poisson_tol=FP64_TOL["c4"], poisson_max_iter=200000),
    }
    out = {"sample": f"first {n_rooms} rooms of c4, {steps} steps each, all submitted at t=0, "
                     f"{slots} slots, largest work (cells x steps) admitted first in both forms, "
                     f"after an untimed warm-up batch of one step per room; the mixed form's "
                     f"batch runs {repeats} times (its served rate varies from run to run by up "
                     f"to 25%) and the run of median rate is reported, all rates listed",
           "forms": {}}
    done_mixed = None
    for name, kw in forms.items():
        # warm-up: the serving loop's cached device blocks, pinned mirrors and
        # streams for these room sizes (as the headline's warm-up steps)
        Server(ctx, slots=slots, policy="longest", **kw).run([Request(c, 1) for c in cfgs],
                                                             realtime=False)
        runs = []
        for _ in range(repeats if name == "mixed" else 1):
            srv = Server(ctx, slots=slots, keep_fields=(name == "mixed"), policy="longest", **kw)
            done = srv.run([Request(c, steps) for c in cfgs], realtime=False)
            runs.append((summary(done), done))
        runs.sort(key=lambda r: r[0]["cell_updates_per_s"])
        sm, done = runs[len(runs) // 2]
        out["forms"][name] = {"cell_updates_per_s": sm["cell_updates_per_s"],
                              "wall_s": sm["wall_s"], "p50_latency_s": sm["p50_latency_s"],
                              "p99_latency_s": sm["p99_latency_s"],
                              "poisson_tol": kw["poisson_tol"],
                              "runs_cell_updates_per_s": [r[0]["cell_updates_per_s"] for r in runs]}
        if name == "mixed":
            done_mixed = done
    out["speedup_mixed_vs_fp64_cg"] = (out["forms"]["mixed"]["cell_updates_per_s"] /
                                       out["forms"]["fp64_cg"]["cell_updates_per_s"])
    errs = {"mixed": [], "fp64_cg": []}
    for c, cfg in zip(done_mixed[:n_err], cfgs[:n_err]):
        xs = c.fields["p"]
        scenes = {k: device.create(ctx, device.prepare(cfg, on_device=True, **kw))
                  for k, kw in forms.items()}
        rhs = scenes["fp64_cg"].laplacian(xs)
        flags = scenes["fp64_cg"].cell_type()
        g = cfg.room.grid
        view = SimpleNamespace(flags=flags, nx=g[0], ny=g[1], nz=g[2])
        sealed = sealed_components(view)
        for k, d in scenes.items():
            x, _ = d.poisson(rhs)
            errs[k].append(gauge_free_error(x, xs, view, sealed))
            d.close()
    out["p_error_max"] = {k: max(v) for k, v in errs.items()}
    out["p_error_rooms"] = n_err
    return out


def c1_reference_single(steps: int = 500):
    """The reference's own FluidSolver (oracle/_ref) on c1 (basic_room,
    25x20x12, upwind, CG tol 1e-6 capped at 1000 iterations), single-threaded
    as its default `parallel: false`, all `steps` steps (convergence off)."""
    from oracle.pyoracle import RefSolver, ref_available
    from oracle.scene_flags import build_scene
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from scenes import BASIC_ROOM, solver_params
    if not ref_available():
        return None
    params = solver_params(BASIC_ROOM)
    params["convergence_tol"] = 0.0
    sc = build_scene(BASIC_ROOM)
    ref = RefSolver(sc, params, use_openmp=False)
    t0 = time.perf_counter()
    for _ in range(steps):
        r = ref.step()
    wall = time.perf_counter() - t0
    cells = sc.nx * sc.ny * sc.nz
    return {"value": cells * steps / wall, "unit": "cell-updates/s", "cores": 1,
            "kind": "reference", "wall_s": wall, "steps": steps,
            "sample": f"c1 (basic_room 25x20x12, upwind, CG tol 1e-6, cap 1000), all {steps} "
                      f"steps, the reference's FluidSolver single-threaded (its default); "
                      f"the reference docs quote 2.25 s for this room on unstated hardware",
            "last_cg_residual": r["pressure_residual"]}


def _scene_arrays(ps):
    """The prepared scene in the oracle's layout (stretched faces included)."""
    from oracle.pyoracle import SceneArrays
    sc = SceneArrays(ps.nx, ps.ny, ps.nz, float(ps.dxv.sum()), float(ps.dyv.sum()),
                     float(ps.dzv.sum()), ps.flags, ps.fans,
                     [o for o in ps.outflows if o not in [tuple(x[:6]) for x in ps.openings]],
                     ps.openings)
    if any(np.ptp(h) > 0 for h in (ps.dxv, ps.dyv, ps.dzv)):
        sc.faces = tuple(np.concatenate([[0.0], np.cumsum(h)]) for h in (ps.dxv, ps.dyv, ps.dzv))
    return sc


def lattice_bench(ctx, cfg, desc, ps, args):
    """The D3Q19 lattice (LBMSolver, FP64) on the workload's scene: cell-updates/s
    of the timed lattice steps, the dominant lattice kernel's roofline, and the
    reference's own serial LBMSolver timed on one step of the same scene."""
    from ff_room import capi
    tau, dt = 0.8, ps.params.dt
    prm = capi.default_params(dt=dt, precision=capi.PREC_FP64, poisson_method=capi.POISSON_CG)
    d = capi.DeviceScene(ctx, ps.nx, ps.ny, ps.nz, ps.dxv, ps.dyv, ps.dzv, ps.flags, ps.fans,
                         ps.outflows, ps.openings, prm)
    d.lbm_attach(tau, dt, 0.0, float(cfg.room.size[0]) / ps.nx)
    for _ in range(args.warmup):
        d.lbm_step()
    if args.kernel_only:  # as for the projection step: the kernel, then the calibration copy
        for name in [*args.kernel_only.split(","), "copy"]:
            ms, nbytes = d.prof_kernel(name, 20)
            print(json.dumps({"kernel": name, "mean_us": ms * 1e3, "bytes": nbytes}), flush=True)
        d.close()
        return None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = d.lbm_step()
    wall = time.perf_counter() - t0
    value = ps.cells * args.steps / wall
    ctx.prof_reset()
    ctx.prof_enable(True)
    for _ in range(args.steps):
        d.lbm_step()
    table = ctx.prof_read()
    ctx.prof_enable(False)
    name, e = max(table.items(), key=lambda kv: kv[1]["total_ms"])
    ms, nbytes = d.prof_kernel(name, 200)
    ach = nbytes / (ms / 1e3)
    roof = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": ach / HBM_PEAK, "traffic": traffic_from_profiles(name, args.workload),
            "kernel": name, "bytes_per_launch": nbytes, "mean_launch_us": ms * 1e3,
            "timing": "200 launches, dispatch begin/end events",
            "launches_per_step": e["launches"] / args.steps}
    out = {
        "metric": "lattice cell-updates/s (D3Q19 LBMSolver step)", "value": value,
        "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded scene generator, lattice at rest)",
        "config": {"workload": desc + f"; lattice tau {tau}, dt {dt}", "cells": ps.cells},
        "roofline": roof,
        "kernels": {k: {"launches": v["launches"], "mean_us": 1e3 * v["total_ms"] / v["launches"],
                        "GBps": (v["bytes_per_launch"] / (v["total_ms"] / v["launches"] / 1e3) / 1e9
                                 if v["bytes_per_launch"] else None)}
                    for k, v in table.items()},
        "last_step": {"velocity_change": r["velocity_change"], "divergence_max": r["divergence_max"]},
    }
    d.close()
    log(f"lattice: {value:.4g} cell-updates/s ({out['ms_per_step']:.3f} ms/step); dominant {name} "
        f"{roof['achieved']:.0f} GB/s ({100 * roof['frac']:.1f}% of peak)")
    if not args.no_cpu:
        from oracle.pyoracle import RefLBM, ref_available
        if ref_available():
            ref = RefLBM(_scene_arrays(ps), tau, dt, 0.0)
            ref.step()  # warm-up
            n = 5 if ps.cells < 10_000_000 else 1  # about 10 s of serial work either way
            t0 = time.perf_counter()
            for _ in range(n):
                ref.step()
            tr = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": ps.cells * n / tr, "unit": "cell-updates/s", "cores": 1,
                                   "kind": "reference",
                                   "sample": f"{n} lattice step(s) after one warm-up step of the same "
                                             f"scene, the reference's LBMSolver (serial: it has no "
                                             f"threading); {tr:.2f} s",
                                   "nproc": os.cpu_count(), "cpu_model": _cpu_model()}
            log(f"cpu: {out['cpu_baseline']['value']:.4g} cell-updates/s")
    return out


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def next_rows(ctx, args):
    """Short measurements of the solver's other paths on the same room: the
    D3Q19 lattice (as --method lbm) and the stretched grid (as --stretch
    1.01: the mixed form against the plain FP64 CG form)."""
    from ff_room import autotune, device
    rows = {}
    cfg, desc = workload(args.workload, "mixed")
    lat = lattice_bench(ctx, cfg, desc, device.prepare(cfg), args)
    rows["lattice"] = {k: lat[k] for k in ("value", "unit", "ms_per_step", "dtype")}
    rows["lattice"]["roofline"] = {k: lat["roofline"][k] for k in ("kernel", "achieved", "frac", "traffic")}
    if "cpu_baseline" in lat:
        rows["lattice"]["cpu_baseline"] = lat["cpu_baseline"]
    # the stretched room on the c2 grid: the reference CG on the stretched
    # operator converges slowly, and c3's FP64 comparison step would take minutes
    sw = "c2"
    cfg, desc = workload(sw, "mixed", 1.01)
    ps = device.prepare(cfg)
    form_params_inplace(ps.params, "mixed", sw)
    ps.params.graphs = args.graphs
    tuned = autotune.lookup(ps.nx, ps.ny, ps.nz, ps.params.precision) if not args.no_tuning else None
    if tuned:
        autotune.apply(ps.params, tuned)
    d = device.create(ctx, ps)
    wall, res, _ = timed_steps(d, ctx, args.steps, args.warmup, prof=False)
    d.close()
    ps64 = device.prepare(cfg)
    form_params_inplace(ps64.params, "fp64", sw)
    d = device.create(ctx, ps64)
    w64, r64, _ = timed_steps(d, ctx, 1, 1, prof=False)
    d.close()
    v, v64 = ps.cells * args.steps / wall, ps64.cells / w64
    rows["stretched"] = {"workload": desc, "value": v, "unit": "cell-updates/s",
                         "ms_per_step": 1e3 * wall / args.steps,
                         "inner_iterations": [min(r["poisson_iters"] for r in res),
                                              max(r["poisson_iters"] for r in res)],
                         "fp64_cg": {"value": v64, "iterations": r64[0]["poisson_iters"],
                                     "residual": r64[0]["pressure_residual"]},
                         "speedup_vs_fp64_cg": v / v64}
    log(f"stretched: {v:.4g} cell-updates/s, fp64 CG {v64:.4g} -> {v / v64:.1f}x")
    return rows


def workload_row(ctx, name: str, args):
    """The mixed and the plain FP64 CG forms on another single-scene workload
    (as `--workload name`): cell-updates/s of `--row-steps` mixed steps after
    the warm-up, the FP64 CG form over `--fp64-steps`, and their ratio."""
    from ff_room import autotune, device
    cfg, desc = workload(name, "mixed")
    ps = device.prepare(cfg)
    form_params_inplace(ps.params, "mixed", name)
    ps.params.graphs = args.graphs
    tuned = autotune.lookup(ps.nx, ps.ny, ps.nz, ps.params.precision) if not args.no_tuning else None
    if tuned:
        autotune.apply(ps.params, tuned)
    d = device.create(ctx, ps)
    wall, res, _ = timed_steps(d, ctx, args.row_steps, args.warmup, prof=False)
    ms, nbytes = d.prof_kernel("mg_smooth_l0", 200)
    d.close()
    ps64 = device.prepare(cfg)
    form_params_inplace(ps64.params, "fp64", name)
    d = device.create(ctx, ps64)
    w64, r64, _ = timed_steps(d, ctx, args.fp64_steps, 1, prof=False)
    d.close()
    v, v64 = ps.cells * args.row_steps / wall, ps64.cells * args.fp64_steps / w64
    iters = [r["poisson_iters"] for r in res]
    return {"workload": desc, "value": v, "unit": "cell-updates/s", "steps": args.row_steps,
            "ms_per_step": 1e3 * wall / args.row_steps, "inner_iterations": [min(iters), max(iters)],
            "fp64_cg": {"value": v64, "steps": args.fp64_steps, "poisson_tol": FP64_TOL[name],
                        "iterations": [r["poisson_iters"] for r in r64]},
            "speedup_vs_fp64_cg": v / v64,
            "mg_smooth_l0": {"mean_launch_us": ms * 1e3, "GBps": nbytes / (ms / 1e3) / 1e9,
                             "frac": nbytes / (ms / 1e3) / HBM_PEAK,
                             "note": "working set inside the 256 MB Infinity Cache"
                             if name == "c2" else ""}}


def cpu_baseline(ps, steps: int = 2, single: bool = False):
    """The reference's own FluidSolver::step (oracle/_ref, compiled from the
    reference sources), timed on this host: with OpenMP on every thread the
    box grants, one warm-up step then `steps` timed (steady state), and with
    `single` the reference's default single thread for one step (tens of
    seconds on c2). Without the reference build, the oracle's restatement
    stands in (kind "port")."""
    from oracle.pyoracle import RefSolver, ref_available, OracleSolver
    threads = int(os.environ.get("OMP_NUM_THREADS", "16"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    p = ps.params
    params = dict(rho=p.rho, nu=p.nu, dt=p.dt, max_steps=p.max_steps, convergence_tol=0.0,
                  poisson_max_iter=1000, poisson_tol=1e-6, advection=p.advection,
                  use_multigrid=0, thermal=p.thermal, T_initial=p.T_initial,
                  T_target=p.T_target, buoyancy=p.buoyancy, g_accel=p.g_accel, beta=p.beta,
                  k_thermal=p.k_thermal, cp=p.cp)
    sc = _scene_arrays(ps)

    def make(parallel: bool):
        if ref_available():
            solver = RefSolver(sc, params, use_openmp=parallel)
            sh = sc.shapes
            solver.set_fields(np.zeros(sh["u"]), np.zeros(sh["v"]), np.zeros(sh["w"]),
                              np.zeros(sh["p"]), ps.T0)
            return solver, "reference"
        solver = OracleSolver(sc, params, nthreads=threads if parallel else 1)
        solver.set_fields(T=ps.T0)
        return solver, "port"

    solver, kind = make(True)
    solver.step()  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        r = solver.step()
    dt = time.perf_counter() - t0
    out = {"value": ps.cells * steps / dt, "unit": "cell-updates/s", "cores": threads,
           "kind": kind,
           "sample": f"{steps} steps after one warm-up step of the same workload from its initial "
                     f"state, the reference CG at its default cap (1000 iterations, tol 1e-6), "
                     f"OpenMP {threads} threads; {dt:.1f} s",
           "cg_residual_at_cap": r.get("pressure_residual"),
           "nproc": os.cpu_count(), "cpu_model": _cpu_model()}
    del solver
    if not single:
        return out
    solver, _ = make(False)
    t0 = time.perf_counter()
    solver.step()
    d1 = time.perf_counter() - t0
    out["single_thread"] = {"value": ps.cells / d1, "unit": "cell-updates/s", "cores": 1,
                            "sample": f"the first step, the reference default (parallel off); "
                                      f"{d1:.1f} s"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--mode", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--method", default="projection", choices=["projection", "lbm"],
                    help="lbm: time the D3Q19 lattice solver on the workload's scene instead")
    ap.add_argument("--stretch", type=float, default=1.0,
                    help="geometric stretch ratio of the grid towards the walls (1: uniform)")
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-single", action="store_true",
                    help="also time the reference's single-threaded step (tens of seconds on c2)")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-next", action="store_true",
                    help="skip the lattice and stretched-grid measurements")
    ap.add_argument("--no-tuning", action="store_true", help="library defaults, not the tuning table")
    ap.add_argument("--graphs", type=int, default=1,
                    help="replay inner iterations from HIP graphs (0 for tracing profilers)")
    ap.add_argument("--gangs", type=int, default=0,
                    help="serving samples: groups of merged launches (0: a stream per scene)")
    ap.add_argument("--slots", type=int, default=6, help="scenes resident at once when serving")
    ap.add_argument("--rows", default="c2",
                    help="other single-scene workloads measured beside the headline (comma list)")
    ap.add_argument("--row-steps", type=int, default=20)
    ap.add_argument("--batched", type=int, default=32,
                    help="c4 rooms timed in both forms with continuous batching (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=300)
    ap.add_argument("--c1-reference", type=int, default=500,
                    help="steps of the reference's single-threaded c1 run (0: skip)")
    ap.add_argument("--c4", type=int, default=64, help="rooms in the batched sample")
    ap.add_argument("--c4-steps", type=int, default=300)
    ap.add_argument("--c5", type=int, default=40, help="requests in the stream sample")
    ap.add_argument("--c5-steps", type=int, default=200)
    ap.add_argument("--kernel-warm", type=int, default=2,
                    help="steps before the --kernel-only launches")
    ap.add_argument("--kernel-only", default=None,
                    help="time the named kernels (comma list; for rocprofv3 --pmc runs) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported: scenes are "
              f"independent and shard one scene (or one slab of a scene) per GPU", file=sys.stderr)
        sys.exit(2)

    from ff_room import capi, device
    ctx = capi.Context(0)
    info = ctx.device_info()
    cfg, desc = workload(args.workload, args.mode, args.stretch)
    ps = device.prepare(cfg)
    log(f"device {info['name']} {info['arch']} CUs={info['compute_units']}; workload {desc}")

    if args.method == "lbm":
        out = lattice_bench(ctx, cfg, desc, ps, args)
        ctx.close()
        if out is not None:
            print(json.dumps(out), flush=True)
        return

    form_params_inplace(ps.params, args.mode, args.workload)
    ps.params.graphs = args.graphs
    tuned = None
    if args.mode == "mixed" and not args.no_tuning:  # the autotuner's entry for this grid class
        from ff_room import autotune
        tuned = autotune.lookup(ps.nx, ps.ny, ps.nz, ps.params.precision)
        if tuned:
            autotune.apply(ps.params, tuned)
            log(f"tuning: {tuned}")
    if args.kernel_only:
        # the named kernels, then the library's plain FP64 copy of nx*ny*nz
        # doubles, whose memory counters calibrate the kernel's
        d = device.create(ctx, ps)
        for _ in range(args.kernel_warm):
            d.step()
        for name in [*args.kernel_only.split(","), "copy"]:
            ms, nbytes = d.prof_kernel(name, 20)
            print(json.dumps({"kernel": name, "mean_us": ms * 1e3, "bytes": nbytes}), flush=True)
        d.close()
        ctx.close()
        return

    # ---- headline form -------------------------------------------------------
    d = device.create(ctx, ps)
    wall, res, _ = timed_steps(d, ctx, args.steps, args.warmup, prof=False)
    d_keep = d
    value = ps.cells * args.steps / wall
    iters = [r["poisson_iters"] for r in res]
    log(f"{args.mode}: {args.steps} steps in {wall:.3f} s -> {value:.4g} cell-updates/s; "
        f"solver iterations {min(iters)}..{max(iters)}")

    # profiled pass of the same steps for the per-kernel roofline
    d = device.create(ctx, ps)
    pwall, _, table = timed_steps(d, ctx, args.steps, args.warmup, prof=True)
    kname, roof = roofline(table, d, args.workload)
    roof["launches_per_step"] = table[kname]["launches"] / args.steps
    d.close()
    log(f"dominant kernel {kname}: {roof['achieved']:.0f} GB/s ({100 * roof['frac']:.1f}% of peak)")

    out = {
        "metric": "cell-updates/s + p50/p99 scene latency, mixed vs FP64 CG at L2 err<=1e-6",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+bf16 stencils/multigrid, f64 outer residual" if args.mode == "mixed" else "f64",
        "data": "synthetic (seeded scene generator, zero initial velocity)",
        "config": {"workload": desc, "cells": ps.cells, "mode": args.mode,
                   "poisson_tol": ps.params.poisson_tol, "tuning": tuned,
                   "warm_start": int(ps.params.warm_start) if args.mode == "mixed" else 0,
                   "note": "the mixed form's step solve starts from the pressure "
                           "extrapolated from the last three steps (warm_start 3); the FP64 "
                           "CG form starts from 0 as the reference does"
                           if args.mode == "mixed" else ""},
        "roofline": roof,
        "kernels": {k: {"launches": v["launches"], "mean_us": 1e3 * v["total_ms"] / v["launches"],
                        "GBps": (v["bytes_per_launch"] / (v["total_ms"] / v["launches"] / 1e3) / 1e9
                                 if v["bytes_per_launch"] else None)}
                    for k, v in table.items()},
        "solver_iterations": {"min": min(iters), "max": max(iters)},
    }

    # ---- the plain FP64 CG form on the same workload (comparison) -----------
    if args.mode == "mixed" and not args.no_compare:
        ps64 = device.prepare(cfg)
        form_params_inplace(ps64.params, "fp64", args.workload)
        sealed = sealed_components(ps)
        out["p_error"] = pressure_errors(ctx, ps, ps64, d_keep, sealed)
        log(f"pressure error vs x*: mixed {out['p_error']['mixed']['rel_l2']:.2e}, "
            f"fp64 CG {out['p_error']['fp64_cg']['rel_l2']:.2e} "
            f"({out['p_error']['sealed_cells']} cells in sealed pockets)")
        out["step_error_vs_fp64"] = step_errors(ctx, ps64, d_keep, sealed)
        log("one step from the same state, mixed vs FP64: " + ", ".join(
            f"{k} {v:.2e}" for k, v in out["step_error_vs_fp64"].items() if k in "puvwT"))
        d = device.create(ctx, ps64)
        w64, r64, _ = timed_steps(d, ctx, args.fp64_steps, 1, prof=False)
        d.close()
        v64 = ps64.cells * args.fp64_steps / w64
        out["fp64_cg"] = {"value": v64, "steps": args.fp64_steps,
                          "iterations": [r["poisson_iters"] for r in r64]}
        out["speedup_vs_fp64_cg"] = value / v64
        log(f"fp64 CG: {v64:.4g} cell-updates/s -> mixed/fp64 = {value / v64:.1f}x")

    d_keep.close()
    if args.mode == "mixed" and not args.no_compare:
        rows = [r for r in args.rows.split(",") if r and r != args.workload]
        if rows:
            out["rows"] = {}
        for r in rows:
            out["rows"][r] = workload_row(ctx, r, args)
            log(f"row {r}: {out['rows'][r]['value']:.4g} cell-updates/s, "
                f"{out['rows'][r]['speedup_vs_fp64_cg']:.1f}x the FP64 CG form")
        if args.batched > 0:
            out["batched_c4"] = batched_forms(ctx, args.batched, args.batched_steps, 8, args.slots)
            b = out["batched_c4"]
            log(f"batched c4: mixed {b['forms']['mixed']['cell_updates_per_s']:.4g}, fp64 CG "
                f"{b['forms']['fp64_cg']['cell_updates_per_s']:.4g} -> "
                f"{b['speedup_mixed_vs_fp64_cg']:.1f}x; p errors {b['p_error_max']}")
    if args.mode == "mixed" and not args.no_next:
        out["next_rows"] = next_rows(ctx, args)
    if args.mode == "mixed" and not args.no_serve:
        out["serving"] = serve_bench(ctx, args.c4, args.c4_steps, args.c5, args.c5_steps, args.gangs,
                                     args.slots)
        for k, v in out["serving"].items():
            log(f"serving {k}: {v['cell_updates_per_s']:.4g} cell-updates/s, p50 "
                f"{v['p50_latency_s']:.3f} s, p99 {v['p99_latency_s']:.3f} s over {v['requests']} requests")

    if not args.no_cpu:
        log("timing the reference CPU solver on a bounded sample ...")
        out["cpu_baseline"] = cpu_baseline(device.prepare(cfg), single=args.cpu_single)
        log(f"cpu: {out['cpu_baseline']['value']:.4g} cell-updates/s")
        if args.c1_reference > 0:
            c1 = c1_reference_single(args.c1_reference)
            if c1 is not None:
                out["cpu_baseline"]["c1_single_thread"] = c1
                log(f"cpu c1 single thread: {c1['steps']} steps in {c1['wall_s']:.2f} s")
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
