"""Shared helpers: scene construction for the oracles and the device."""
from __future__ import annotations

import numpy as np

from oracle.pyoracle import SceneArrays


def device_outflows(sc: SceneArrays):
    """BoundaryManager::outflows() order: every opening is also an outflow patch."""
    return [tuple(o[:6]) for o in sc.openings] + list(sc.outflows)


def device_scene(ctx, sc: SceneArrays, params):
    from ff_room.capi import DeviceScene
    dxv, dyv, dzv = sc.spacings()
    return DeviceScene(ctx, sc.nx, sc.ny, sc.nz, dxv, dyv, dzv, sc.flags, sc.fans,
                       device_outflows(sc), sc.openings, params)


def capi_params(p: dict, **extra):
    from ff_room.capi import default_params
    keys = ("rho", "nu", "dt", "max_steps", "convergence_tol", "poisson_max_iter", "poisson_tol",
            "advection", "use_multigrid", "thermal", "T_initial", "T_target", "buoyancy",
            "g_accel", "beta", "k_thermal", "cp")
    kw = {k: p[k] for k in keys if k in p}
    kw.update(extra)
    return default_params(**kw)


def rel_l2(a, b, mask=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if mask is not None:
        a, b = a[mask], b[mask]
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)
