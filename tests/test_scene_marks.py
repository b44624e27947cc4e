"""Scene construction from marks (the device-marked path, prepare_marks).

Pins, on the CPU, against the host Grid + BoundaryManager builder (Scene,
itself pinned to the reference's known answers in test_scene_config.py):
the flags that the marks give (flags_from_marks, what the device's marking
kernel computes), the boundary lists, the spacings (uniform and stretched)
and the initial temperature field (openings' outside temperatures, fan inlet
temperatures on INFLOW cells), over seeded rooms with overlapping boxes,
openings and fans that reach past the walls.
"""
import numpy as np
import pytest

from ff_room import device
from ff_room.config import FanConfig, ObstacleConfig, OpeningConfig, RoomConfig, SceneConfig
from ff_room.device import flags_from_marks
from ff_room.generator import random_scene
from ff_room.scene import Scene, boundary_lists, cell_marks


def _configs():
    cfgs = [random_scene(300 + i, grid=(20 + 3 * i, 14 + i, 9 + (i % 4)), n_fans=1 + i % 3,
                         n_obstacles=i % 5, openings=(i % 2 == 1), T_initial=25.0 + i,
                         buoyancy=True) for i in range(8)]
    # hand-made: a fan at the wall, overlapping boxes, two openings on one
    # wall, fan inlet temperatures, a stretched grid
    cfgs.append(SceneConfig(
        room=RoomConfig(size=(4.0, 3.0, 2.5), grid=(24, 18, 15), stretch=1.08),
        fans=[FanConfig(name="a", position=(0.05, 1.5, 1.2), direction=(0.0, 0.0), velocity=3.0,
                        radius=0.4, T_inlet=18.0),
              FanConfig(name="b", position=(2.0, 2.9, 0.4), direction=(270.0, 0.0), velocity=2.0,
                        radius=0.25, T_inlet=40.0)],
        obstacles=[ObstacleConfig(name="o1", bbox_min=(1.0, 0.5, 0.0), bbox_max=(2.5, 1.8, 1.0)),
                   ObstacleConfig(name="o2", bbox_min=(2.0, 1.5, 0.5), bbox_max=(3.5, 2.9, 2.0))],
        openings=[OpeningConfig(name="w1", wall="north", center=(1.0, 1.2), size=(0.8, 1.0)),
                  OpeningConfig(name="w2", wall="north", center=(1.4, 1.4), size=(1.0, 0.6),
                                T_outside=5.0)]))
    cfgs[-1].solver.buoyancy = True
    cfgs[-1].solver.T_initial = 27.0
    return cfgs


@pytest.mark.parametrize("cfg", _configs(), ids=lambda c: "x".join(map(str, c.room.grid)))
def test_marks_match_the_host_builder(cfg):
    s = Scene(cfg)
    g = s.grid
    host = np.array(g.cell_type_shaped())
    dev = flags_from_marks(host.shape, cell_marks(cfg))
    assert np.array_equal(dev, host)
    fans, outs, opens = boundary_lists(cfg)
    assert fans == [(f.i_min, f.i_max, f.j_min, f.j_max, f.k_min, f.k_max, *f.vel) for f in s.bm.fans()]
    assert outs == [(o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max) for o in s.bm.outflows()]
    assert opens == [(o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max, o.T_outside)
                     for o in s.bm.openings()]


@pytest.mark.parametrize("cfg", _configs(), ids=lambda c: "x".join(map(str, c.room.grid)))
def test_prepared_marks_match_prepared_flags(cfg):
    a = device.prepare(cfg)
    b = device.prepare(cfg, on_device=True)
    assert b.flags is None and b.marks == cell_marks(cfg)
    for k in ("dxv", "dyv", "dzv", "T0"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert (a.fans, a.outflows, a.openings) == (b.fans, b.outflows, b.openings)
    assert bytes(a.params) == bytes(b.params)


def test_flags_from_marks_clamps_and_orders():
    m = [(1, -3, 1, 0, 0, 0, 0),      # reaches past the low wall
         (2, 1, 9, 0, 0, 0, 0),       # past the high wall, over the first
         (3, 2, -2, 0, 0, 0, 0),      # empty (hi < lo, hi negative)
         (0, 2, 2, 0, 0, 0, 0)]       # FLUID again
    f = flags_from_marks((4, 1, 1), m)
    assert f.reshape(-1).tolist() == [1, 2, 0, 2]
