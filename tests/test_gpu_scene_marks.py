"""Cell flags marked on the device (ffr_scene_desc.marks) and the neighbour
codes built there for every scene.

Pins: the device's flags equal the host builder's, byte for byte; a scene
created from marks steps exactly as the same scene created from host flags
(fields and step results bit for bit, mixed and FP64 forms; the neighbour
codes come from the device in both); a bad mark type is refused.
"""
import numpy as np
import pytest

from ff_room import capi, device
from ff_room.generator import random_scene

pytestmark = pytest.mark.gpu


def _cfgs(precision):
    base = dict(dt=0.005, max_steps=4, convergence_tol=0.0, advection="quick", buoyancy=True,
                beta=3.4e-3, alpha=2.1e-5, T_initial=26.0, precision=precision)
    return [random_scene(401, grid=(34, 26, 17), n_fans=2, n_obstacles=3, **base),
            random_scene(402, grid=(21, 30, 14), n_fans=3, n_obstacles=1, openings=True, **base)]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_device_marks_equal_host_flags_and_steps(ctx, precision):
    for cfg in _cfgs(precision):
        a, b = device.prepare(cfg), device.prepare(cfg, on_device=True)
        da, db = device.create(ctx, a), device.create(ctx, b)
        assert np.array_equal(db.cell_type(), a.flags) and np.array_equal(da.cell_type(), a.flags)
        for _ in range(4):
            ra, rb = da.step(), db.step()
            assert ra == rb
        fa, fb = da.download(), db.download()
        for k in ("u", "v", "w", "p", "T"):
            assert np.array_equal(fa[k], fb[k]), k
        da.close()
        db.close()


def test_bad_mark_is_refused(ctx):
    cfg = _cfgs("mixed")[0]
    ps = device.prepare(cfg, on_device=True)
    ps.marks = list(ps.marks) + [(7, 0, 1, 0, 1, 0, 1)]
    with pytest.raises(capi.FfrError):
        device.create(ctx, ps)
