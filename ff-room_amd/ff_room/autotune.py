"""Autotuner for the mixed-precision step on one GPU.

Tunes what changes speed but not what the solve converges to: the block cap
of grid-strided launches (``ffr_set_max_blocks``), the multigrid sweeps per
level (``mg_pre_sweeps``, ``mg_level_sweeps``, ``mg_coarse_sweeps``), the
inner solve's tolerance (``inner_tol``) and the LDS budget of the
single-workgroup kernel for the smallest levels (``mg_tail_kb``). The FP64
outer correction stops every solve at the scene's ``poisson_tol`` whatever
these are, so a candidate is judged by the wall time of a few steps alone;
a candidate whose solve does not reach the tolerance is rejected.

Search: coordinate descent from the library defaults (or, with ``--refine``,
from the stored entry), one parameter at a time over its candidate values,
keeping the fastest, for ``rounds`` rounds.
Each evaluation builds a fresh scene, takes ``warmup`` steps and times
``steps`` steps. Results are keyed by grid shape class and precision and
stored as JSON (``tuning.json`` beside this module by default); ``lookup``
returns the entry for a scene, ``apply`` writes it into ffr_params.

    python -m ff_room.autotune --workload c2 [--out tuning.json] [--refine]
"""
from __future__ import annotations

import argparse
import json
import os
import time
from typing import Dict, Optional

from . import capi, device

SPACE = {
    "max_blocks": [1024, 2048, 4096, 8192],
    "mg_pre_sweeps": [1, 2, 3],
    "mg_level_sweeps": [1, 2, 3],
    "mg_coarse_sweeps": [0, 4, 8, 16],  # 0: dense inverse of the coarsest level
    # tighter inner solves trade a few inner iterations for fewer FP64 outer passes
    "inner_tol": [1e-6, 1e-5, 1e-4, 3e-4, 1e-3, 3e-3],
    "mg_tail_kb": [48, 96, 150],
}
DEFAULTS = {"max_blocks": 2048, "mg_pre_sweeps": 2, "mg_level_sweeps": 2, "mg_coarse_sweeps": 8,
            "inner_tol": 1e-4, "mg_tail_kb": 96}
DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def shape_key(nx: int, ny: int, nz: int, precision: int) -> str:
    """Grids whose cell counts share a power of two (and precision) share an
    entry: launch shapes and level counts scale with the cell count."""
    n = nx * ny * nz
    return f"{'mixed' if precision == capi.PREC_MIXED else 'fp64'}:2^{max(0, n.bit_length() - 1)}"


def apply(params: capi.Params, entry: Dict) -> capi.Params:
    """Writes a tuning entry into ffr_params (and the process-wide block cap)."""
    for k, v in entry.items():
        if k == "max_blocks":
            capi.lib().ffr_set_max_blocks(int(v))
        elif hasattr(params, k):
            setattr(params, k, type(getattr(params, k))(v))
    return params


def load(path: str = DEFAULT_PATH) -> Dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def lookup(nx: int, ny: int, nz: int, precision: int, path: str = DEFAULT_PATH) -> Optional[Dict]:
    e = load(path).get(shape_key(nx, ny, nz, precision))
    return e["params"] if e else None


def _copy(p: capi.Params) -> capi.Params:
    return capi.default_params(**{k: getattr(p, k) for k, _ in capi.Params._fields_})


def measure(ctx: capi.Context, ps: device.PreparedScene, cand: Dict, warmup: int = 2,
            steps: int = 15) -> Optional[float]:
    """Mean seconds per step of `cand`, or None if a solve misses its tolerance.
    The window matches the bench's (a few steps into the run): a solve's cost
    changes as the flow develops and the previous pressure becomes a better
    start, and 5 steps ranked c2's coarse solves the wrong way round."""
    prm = apply(_copy(ps.params), cand)
    d = device.create(ctx, ps, prm)
    try:
        for _ in range(warmup):
            d.step()
        t0 = time.perf_counter()
        res = [d.step() for _ in range(steps)]
        dt = (time.perf_counter() - t0) / steps
    finally:
        d.close()
    if any(r["pressure_residual"] > ps.params.poisson_tol or r["nonfinite"] for r in res):
        return None
    return dt


def tune(ctx: capi.Context, ps: device.PreparedScene, rounds: int = 2, log=print,
         start: Optional[Dict] = None) -> Dict:
    best = dict(DEFAULTS, **(start or {}))
    best_t = measure(ctx, ps, best)
    if best_t is None:  # a stored start that no longer converges: begin from the defaults
        best = dict(DEFAULTS)
        best_t = measure(ctx, ps, best)
    log(f"  {'start' if start else 'defaults'}: {1e3 * best_t:.3f} ms/step")
    for r in range(rounds):
        changed = False
        for key, values in SPACE.items():
            for v in values:
                if v == best[key]:
                    continue
                cand = dict(best, **{key: v})
                t = measure(ctx, ps, cand)
                log(f"  round {r} {key}={v}: " + ("rejected" if t is None else f"{1e3 * t:.3f} ms"))
                if t is not None and t < best_t * 0.99:  # keep a 1% margin against noise
                    best, best_t, changed = cand, t, True
        if not changed:
            break
    capi.lib().ffr_set_max_blocks(DEFAULTS["max_blocks"])
    return {"params": best, "ms_per_step": 1e3 * best_t}


def main():
    from .generator import workload_c2, workload_c3, workload_c4
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4"])
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--refine", action="store_true",
                    help="start from the stored entry and re-tune classes already in the table")
    args = ap.parse_args()
    ctx = capi.Context(0)
    if args.workload == "c4":  # one representative room per size class
        cfgs = workload_c4(n_scenes=64, precision="mixed")
    else:
        cfgs = [(workload_c2 if args.workload == "c2" else workload_c3)(precision="mixed")]
    table = load(args.out)
    tuned = set()  # each shape class once per run, by the first of its rooms
    for cfg in cfgs:
        ps = device.prepare(cfg, precision=capi.PREC_MIXED, poisson_method=capi.POISSON_MGPCG,
                            poisson_tol=1e-10, poisson_max_iter=500)
        key = shape_key(ps.nx, ps.ny, ps.nz, capi.PREC_MIXED)
        if key in tuned or (key in table and args.workload == "c4" and not args.refine):
            continue
        tuned.add(key)
        start = table[key]["params"] if args.refine and key in table else None
        print(f"tuning {key} ({ps.nx}x{ps.ny}x{ps.nz})", flush=True)
        table[key] = dict(tune(ctx, ps, args.rounds, start=start), grid=[ps.nx, ps.ny, ps.nz])
        with open(args.out, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    ctx.close()


if __name__ == "__main__":
    main()
