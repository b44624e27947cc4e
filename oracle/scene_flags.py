"""Scene description -> cell flags and boundary lists (test infrastructure only).

Restates ddd3h/ff-room python/ff_room/scene.py:72-231 (Scene._build and its
_add_* helpers) and config.py:41-59 (FanConfig direction helpers) over plain
dicts in the SceneConfig.to_dict() shape, writing flags with numpy in the
reference's marking order (obstacles -> fans -> openings -> auto outflow).
Used by the tests to check the product's scene builder and to feed the
oracles; the product never imports it.
"""
from __future__ import annotations

import math

import numpy as np

from .pyoracle import SceneArrays, SOLID, INFLOW, OUTFLOW


def direction_vector(az_deg: float, el_deg: float):
    # config.py:41-48
    az, el = math.radians(az_deg), math.radians(el_deg)
    return (math.cos(el) * math.cos(az), math.cos(el) * math.sin(az), math.sin(el))


def snapped_axis(vec) -> int:
    # config.py:50-53: first index of the largest |component|
    best = 0
    for i in (1, 2):
        if abs(vec[i]) > abs(vec[best]):
            best = i
    return best


def stretched_faces(n: int, length: float, ratio: float):
    """scene.py:34-59 (_stretched_faces): n+1 face coordinates. Uniform
    i*L/n when |ratio-1| < 1e-6; otherwise the first n//2 widths form the
    geometric series dx1*ratio^i with dx1 = (L/2)(ratio-1)/(ratio^(n//2)-1),
    the next n//2 widths repeat the realised ones in reverse, an odd n splits
    the first right-half cell at its midpoint, and every face is multiplied by
    L / (last face)."""
    if abs(ratio - 1.0) < 1e-6:
        return [i * length / n for i in range(n + 1)]
    m = n // 2
    dx1 = (length / 2.0) * (ratio - 1.0) / (ratio ** m - 1.0)
    xs = [0.0]
    for i in range(m):
        xs.append(xs[-1] + dx1 * ratio ** i)
    widths = [xs[i + 1] - xs[i] for i in range(m)]
    for wd in widths[::-1]:
        xs.append(xs[-1] + wd)
    if n % 2 == 1:
        xs.insert(m + 1, (xs[m] + xs[m + 1]) / 2.0)
    k = length / xs[-1]
    return [x * k for x in xs]


def build_scene(cfg: dict) -> SceneArrays:
    room = cfg["room"]
    (lx, ly, lz), (nx, ny, nz) = room["size"], room["grid"]
    L, N = (lx, ly, lz), (nx, ny, nz)
    flags = np.zeros((nx, ny, nz), dtype=np.uint8)
    fans, outflows, openings = [], [], []

    def cell(x, a):  # int(x/L*N), truncation toward zero
        return int(x / L[a] * N[a])

    # obstacles, scene.py:135-145 -> add_solid_box (BoundaryManager.cpp:6-14)
    for ob in cfg.get("obstacles", []):
        lo = [cell(ob["bbox_min"][a], a) for a in range(3)]
        hi = [cell(ob["bbox_max"][a], a) - 1 for a in range(3)]
        s = [slice(max(0, lo[a]), min(N[a] - 1, hi[a]) + 1) for a in range(3)]
        if all(x.start < x.stop for x in s):
            flags[s[0], s[1], s[2]] = SOLID

    # fans, scene.py:147-192 -> add_fan (BoundaryManager.cpp:16-22)
    fan_list = cfg.get("fans", [])
    for fan in fan_list:
        vec = direction_vector(*fan.get("direction", (0.0, 0.0)))
        ax = snapped_axis(vec)
        vel = float(fan.get("velocity", 3.0))
        rad = float(fan.get("radius", 0.15))
        pos = fan.get("position", (0.5, 2.0, 1.0))
        box = []
        for a in range(3):
            if a == ax:
                c = max(0, min(N[a] - 1, cell(pos[a], a)))
                box += [c, c]
            else:
                box += [max(0, int((pos[a] - rad) / L[a] * N[a])),
                        min(N[a] - 1, int((pos[a] + rad) / L[a] * N[a]))]
        fans.append(tuple(box) + (vec[0] * vel, vec[1] * vel, vec[2] * vel))
        s = [slice(max(0, box[2 * a]), min(N[a] - 1, box[2 * a + 1]) + 1) for a in range(3)]
        if all(x.start < x.stop for x in s):
            flags[s[0], s[1], s[2]] = INFLOW

    # openings, scene.py:194-231 -> add_opening/add_outflow (BoundaryManager.cpp:24-44)
    walls = {"west": (0, 0), "east": (0, 1), "south": (1, 0), "north": (1, 1),
             "floor": (2, 0), "ceiling": (2, 1)}
    T_out_default = float(cfg.get("solver", {}).get("T_outside", 20.0))
    for op in cfg.get("openings", []):
        ax, sd = walls[op.get("wall", "east").lower()]
        ta, tb = [a for a in range(3) if a != ax]
        (ca, cb), (wa, wb) = op.get("center", (2.0, 1.2)), op.get("size", (0.8, 1.0))
        a_lo = max(0, int((ca - wa / 2) / L[ta] * N[ta]))
        a_hi = min(N[ta] - 1, int((ca + wa / 2) / L[ta] * N[ta]))
        b_lo = max(0, int((cb - wb / 2) / L[tb] * N[tb]))
        b_hi = min(N[tb] - 1, int((cb + wb / 2) / L[tb] * N[tb]))
        T_out = op.get("T_outside")
        T_out = T_out_default if T_out is None else float(T_out)
        openings.append((ax, sd, a_lo, a_hi, b_lo, b_hi, T_out))
        _mark_outflow(flags, ax, sd, a_lo, a_hi, b_lo, b_hi)

    # auto outflow on the wall the first fan faces, only without openings (:105-133)
    if fan_list and not cfg.get("openings"):
        vec = direction_vector(*fan_list[0].get("direction", (0.0, 0.0)))
        ax = snapped_axis(vec)
        sd = 1 if vec[ax] >= 0 else 0
        ta, tb = [a for a in range(3) if a != ax]
        o = (ax, sd, 0, N[ta] - 1, 0, N[tb] - 1)
        outflows.append(o)
        _mark_outflow(flags, *o)

    # `outflows` holds only plain outflow patches; an opening is registered as an
    # outflow patch too (BoundaryManager.cpp:132-135), which the consumers add.
    return SceneArrays(nx, ny, nz, lx, ly, lz, flags.reshape(-1), fans, outflows, openings)


def _mark_outflow(flags, ax, sd, a0, a1, b0, b1):
    n = flags.shape[ax]
    idx = n - 1 if sd else 0
    if ax == 0:
        flags[idx, a0:a1 + 1, b0:b1 + 1] = OUTFLOW
    elif ax == 1:
        flags[a0:a1 + 1, idx, b0:b1 + 1] = OUTFLOW
    else:
        flags[a0:a1 + 1, b0:b1 + 1, idx] = OUTFLOW
