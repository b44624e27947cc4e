"""Scene construction and configuration (CPU only).

The product's Scene (ff_room.scene, writing flags through the host C++
BoundaryManager of _ffroom_core) must produce the reference's flags bit for
bit. Pins: the hand-computed known answer for configuration c1; the flags the
reference's own BoundaryManager writes for the same calls (oracle/_ref); the
oracle's restatement of scene.py on seeded random scenes; and the
configuration behaviours of the reference's python/tests/test_config.py.
"""
import math
import os
import tempfile

import numpy as np
import pytest

from ff_room import config as C
from ff_room.generator import random_scene
from ff_room.scene import Scene, _stretched_faces
from oracle.pyoracle import ref_available, ref_flags_from_calls
from oracle.scene_flags import build_scene
from scenes import BASIC_ROOM, SMALL_LW, SMALL_THERMAL

REF_EXAMPLES = "/root/reference/examples"


def product_scene(cfg_dict):
    s = Scene(C.SceneConfig.from_dict(cfg_dict))
    flags = s.grid.cell_type_shaped().reshape(-1)
    fans = [(f.i_min, f.i_max, f.j_min, f.j_max, f.k_min, f.k_max, *f.vel) for f in s.bm.fans()]
    outs = [(o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max) for o in s.bm.outflows()]
    opens = [(o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max, o.T_outside)
             for o in s.bm.openings()]
    return flags, fans, outs, opens


def test_c1_known_answer():
    # basic_room (configuration c1): fan INFLOW i=2, j 9..10, k 5..6; sofa SOLID
    # i 10..14, j 2..8, k 0..3; auto outflow on the whole east wall i=24.
    flags, fans, outs, opens = product_scene(BASIC_ROOM)
    f = flags.reshape(25, 20, 12)
    assert np.bincount(f.ravel(), minlength=4).tolist() == [5616, 140, 4, 240]
    inflow = np.argwhere(f == 2)
    assert inflow.min(0).tolist() == [2, 9, 5] and inflow.max(0).tolist() == [2, 10, 6]
    solid = np.argwhere(f == 1)
    assert solid.min(0).tolist() == [10, 2, 0] and solid.max(0).tolist() == [14, 8, 3]
    assert np.all(f[24] == 3) and np.count_nonzero(f == 3) == 240
    assert fans == [(2, 2, 9, 10, 5, 6, 3.0, 0.0, 0.0)]
    assert outs == [(0, 1, 0, 19, 0, 11)]
    assert opens == []
    sc = build_scene(BASIC_ROOM)
    assert np.array_equal(sc.flags, flags)


@pytest.mark.parametrize("seed", range(12))
def test_random_scenes_match_oracle(seed):
    cfg = random_scene(seed, grid=(24 + seed, 20 + seed % 5, 12 + seed % 3), n_fans=1 + seed % 3,
                       n_obstacles=seed % 6, openings=(seed % 4 == 0)).to_dict()
    flags, fans, outs, opens = product_scene(cfg)
    sc = build_scene(cfg)
    assert np.array_equal(flags, sc.flags)
    assert fans == [tuple(f) for f in sc.fans]
    assert outs == [tuple(o[:6]) for o in sc.openings] + [tuple(o) for o in sc.outflows]
    assert opens == [tuple(o) for o in sc.openings]


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", range(8))
def test_flags_match_reference_boundary_manager(seed):
    cfg_obj = random_scene(100 + seed, grid=(20, 16, 10), n_fans=2, n_obstacles=4,
                           openings=bool(seed % 2))
    cfg = cfg_obj.to_dict()
    flags, fans, outs, opens = product_scene(cfg)
    r = cfg_obj.room
    boxes = []
    for ob in cfg_obj.obstacles:
        lo = [int(ob.bbox_min[a] / r.size[a] * r.grid[a]) for a in range(3)]
        hi = [int(ob.bbox_max[a] / r.size[a] * r.grid[a]) - 1 for a in range(3)]
        boxes.append((lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]))
    plain = outs[len(opens):]
    ref = ref_flags_from_calls(*r.grid, *r.size, solid_boxes=boxes, fans=fans, outflows=plain,
                               openings=opens)
    assert np.array_equal(ref, flags)


def test_direction_vectors():
    # python/tests/test_config.py:27-61 of the reference
    v = C.FanConfig(direction=(0.0, 0.0)).direction_vector()
    assert abs(v[0] - 1) < 1e-9 and abs(v[1]) < 1e-9 and abs(v[2]) < 1e-9
    f = C.FanConfig(direction=(180.0, 0.0))
    assert abs(f.direction_vector()[0] + 1) < 1e-6 and f.snapped_axis() == 0 and f.snapped_sign() == -1.0
    f = C.FanConfig(direction=(90.0, 0.0))
    assert abs(f.direction_vector()[1] - 1) < 1e-6 and f.snapped_axis() == 1
    f = C.FanConfig(direction=(0.0, 90.0))
    assert abs(f.direction_vector()[2] - 1) < 1e-6 and f.snapped_axis() == 2
    for az in (0, 30, 90, 135, 270):
        for el in (-30, 0, 45, 90):
            v = C.FanConfig(direction=(az, el)).direction_vector()
            assert abs(math.sqrt(sum(x * x for x in v)) - 1) < 1e-9
    # axis-snapped angles leave ~1e-16 cross components that stay non-zero
    vx = C.FanConfig(direction=(90.0, 0.0)).direction_vector()[0]
    assert vx != 0.0 and abs(vx) < 1e-15


def test_config_round_trips(tmp_path):
    s = C.SceneConfig.from_dict(BASIC_ROOM)
    d = s.to_dict()
    s2 = C.SceneConfig.from_dict(d)
    assert s2.room.size == s.room.size and s2.fans[0].velocity == 3.0
    s.save_yaml(str(tmp_path / "a.yaml"))
    assert C.SceneConfig.load_yaml(str(tmp_path / "a.yaml")).room.grid == (25, 20, 12)
    s.save_json(str(tmp_path / "a.json"))
    assert C.SceneConfig.load_json(str(tmp_path / "a.json")).fans[0].position == (0.5, 2.0, 1.2)


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference examples absent")
@pytest.mark.parametrize("name", ["basic_room.yaml", "two_fans.yaml", "summer_cooling.yaml"])
def test_reference_yaml_loads_unchanged(name):
    cfg = C.SceneConfig.load_yaml(os.path.join(REF_EXAMPLES, name))
    s = Scene(cfg)
    assert s.grid.Nx == cfg.room.grid[0]
    assert np.array_equal(s.grid.cell_type_shaped().reshape(-1), build_scene(cfg.to_dict()).flags)


def test_errors_like_the_reference():
    with pytest.raises(TypeError):
        C.SceneConfig.from_dict({"room": {"size": [1, 1, 1], "bogus": 1}})
    bad = dict(BASIC_ROOM, openings=[{"wall": "sideways"}])
    with pytest.raises(KeyError):
        Scene(C.SceneConfig.from_dict(bad))
    from ff_room.scene import _load_core
    with pytest.raises(ValueError):
        _load_core().Grid(0, 4, 4, 1.0, 1.0, 1.0)


def test_stretched_faces():
    # no reference configuration uses stretch; properties only (parity unpinned)
    for N, r in ((10, 1.2), (11, 1.3), (25, 1.0)):
        f = np.array(_stretched_faces(N, 3.0, r))
        assert len(f) == N + 1 and f[0] == 0.0 and abs(f[-1] - 3.0) < 1e-12
        d = np.diff(f)
        assert np.all(d > 0)
        if N % 2 == 0:  # even N mirrors; odd N splits the first right-half cell
            assert np.allclose(d, d[::-1], rtol=1e-9)
        else:
            assert np.isclose(d[N // 2], d[N // 2 + 1], rtol=1e-12)
        if r > 1:
            assert np.all(np.diff(d[:N // 2]) > 0)


def test_host_boundary_helpers_match_oracle():
    """BoundaryManager.apply_* host helpers against the oracle restatement."""
    from oracle.pyoracle import OracleSolver
    cfg = SMALL_THERMAL
    s = Scene(C.SceneConfig.from_dict(cfg))
    sc = build_scene(cfg)
    rng = np.random.default_rng(5)
    g = s.grid
    u, v, w, T = g.u, g.v, g.w, g.T_shaped().reshape(-1)
    for a in (u, v, w, T):
        a[:] = rng.standard_normal(a.size)
    o = OracleSolver(sc, {}, init=False)
    o.set_fields(u=u.copy(), v=v.copy(), w=w.copy(), T=T.copy())
    for name in ("apply_noslip", "apply_inflow", "apply_outflow", "apply_opening_temperature"):
        getattr(s.bm, name)(g)
        o.call("orc_" + name)
        assert np.array_equal(g.u, o.u) and np.array_equal(g.v, o.v) and np.array_equal(g.w, o.w)
        assert np.array_equal(g.T_shaped().reshape(-1), o.T), name
