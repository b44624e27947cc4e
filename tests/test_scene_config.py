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


def _stretched_closed_form(N, L, r):
    """Widths from the closed form of scene.py:34-59: the left N//2 cells are
    d0 r^i with d0 = (L/2)(r-1)/(r^(N//2)-1), so they sum to L/2; the right
    half repeats them in reverse; odd N halves the first right-half cell into
    two cells. The widths sum to L, so the final rescale is 1 up to rounding."""
    n = N // 2
    d0 = (L / 2) * (r - 1) / (r ** n - 1)
    left = [d0 * r ** i for i in range(n)]
    right = left[::-1]
    if N % 2 == 1:
        right = [right[0] / 2, right[0] / 2] + right[1:]
    return np.array(left + right)


@pytest.mark.parametrize("N,r", [(10, 1.2), (11, 1.3), (8, 1.05), (9, 1.15), (64, 1.01),
                                 (3, 1.5), (2, 1.4)])
def test_stretched_faces_closed_form(N, r):
    """Both the product's _stretched_faces and the oracle's restatement (which
    the stretched parity tests build their grids with) against the closed
    form: consecutive left-half widths in ratio r, the mirror, the sum L, the
    odd-N split. A wrong ratio, a wrong first width or an off-by-one in the
    series fails here even though the final rescale restores the sum."""
    from oracle.scene_flags import stretched_faces
    L = 3.0
    prod = np.array(_stretched_faces(N, L, r))
    orc = np.array(stretched_faces(N, L, r))
    assert np.array_equal(prod, orc)
    d = np.diff(prod)
    assert len(prod) == N + 1 and prod[0] == 0.0 and abs(prod[-1] - L) < 1e-13
    assert np.allclose(d, _stretched_closed_form(N, L, r), rtol=1e-12, atol=0)
    n = N // 2
    for i in range(n - 1):
        assert abs(d[i + 1] / d[i] - r) < 1e-12
    if N % 2 == 0:
        assert np.allclose(d, d[::-1], rtol=1e-12, atol=0)
    else:  # the odd cell split in two equal halves, the rest mirrored
        assert abs(d[n] - d[n + 1]) < 1e-15 * L and abs(d[n] - d[n - 1] / 2) < 1e-12 * d[n]
        assert np.allclose(d[n + 2:], d[:n - 1][::-1], rtol=1e-12, atol=0)
    assert abs(d[: n].sum() - L / 2) < 1e-12


def test_stretched_faces_uniform_limit():
    # |r - 1| < 1e-6: i * L / N exactly (scene.py:40-41)
    from oracle.scene_flags import stretched_faces
    for r in (1.0, 1.0 + 5e-7):
        assert _stretched_faces(7, 2.0, r) == [i * 2.0 / 7 for i in range(8)]
        assert stretched_faces(7, 2.0, r) == [i * 2.0 / 7 for i in range(8)]


def test_summer_cooling_known_answer():
    """examples/summer_cooling.yaml (30x22x12 in 6 x 4.5 x 2.5 m), worked by
    hand from int(x / L * N) and the clamps of scene.py:135-231:
      window_north: y-layer 21, x int(2.55/6*30)=12 .. int(3.45/6*30)=17,
                    z int(0.65/2.5*12)=3 .. int(1.75/2.5*12)=8     -> 36 OUTFLOW
      door_south:   y-layer 0, x int(1.1/6*30)=5 .. int(1.9/6*30)=9,
                    z 0 .. int(2.0/2.5*12)=9                        -> 50 OUTFLOW
      fan (-y):     j int(3.8/4.5*22)=18, i int(2.85/6*30)=14 .. int(3.15/6*30)=15,
                    k int(0.75/2.5*12)=3 .. int(1.05/2.5*12)=5      -> 6 INFLOW
      bed:          i 1 .. 10-1=9, j 1 .. int(2.5/4.5*22)-1=11, k 0 .. 2-1=1 -> 198 SOLID
      desk:         i 17 .. 26-1=25, j 1 .. 7-1=6, k 0 .. 3-1=2     -> 162 SOLID
    and no automatic outflow (the scene has openings)."""
    from scenes import SUMMER_COOLING
    flags, fans, outs, opens = product_scene(SUMMER_COOLING)
    sc = build_scene(SUMMER_COOLING)
    want_opens = [(1, 1, 12, 17, 3, 8, 25.0), (1, 0, 5, 9, 0, 9, 25.0)]
    assert opens == want_opens
    assert [tuple(o) for o in sc.openings] == want_opens
    assert outs == [o[:6] for o in want_opens] and sc.outflows == []
    assert [f[:6] for f in fans] == [(14, 15, 18, 18, 3, 5)]
    assert [tuple(f[:6]) for f in sc.fans] == [(14, 15, 18, 18, 3, 5)]
    vx, vy, vz = fans[0][6:]
    assert vy == -2.5 and vz == 0.0 and vx != 0.0 and abs(vx) < 1e-15  # Q17
    for f in (flags, sc.flags):
        g = f.reshape(30, 22, 12)
        assert np.bincount(g.ravel(), minlength=4).tolist() == [7468, 360, 6, 86]
        assert np.all(g[12:18, 21, 3:9] == 3) and np.count_nonzero(g[:, 21, :] == 3) == 36
        assert np.all(g[5:10, 0, 0:10] == 3) and np.count_nonzero(g[:, 0, :] == 3) == 50
        assert np.all(g[14:16, 18, 3:6] == 2)
        assert np.all(g[1:10, 1:12, 0:2] == 1) and np.all(g[17:26, 1:7, 0:3] == 1)
        assert g[0, 1, 0] == 0 and g[10, 1, 0] == 0 and g[1, 12, 0] == 0 and g[1, 1, 2] == 0
        assert g[16, 1, 0] == 0 and g[26, 1, 0] == 0 and g[17, 7, 0] == 0 and g[17, 1, 3] == 0


def test_fans_at_walls_known_answer():
    """Fan boxes clamped at the walls (scene.py:154-192), by hand:
      east (+x, r 0.3): i int(4.99/5*25)=24; j max(0, int(-0.18/4*20)=0)=0 ..
            int(0.42/4*20)=2; k int(2.1/2.5*12)=10 .. min(11, int(2.7/2.5*12)=12)=11
      north (+y, r 0.25): j int(3.95/4*20)=19; i max(0, int(-0.05/5*25)=0)=0 ..
            int(0.45/5*25)=2; k int(0.95/2.5*12)=4 .. int(1.45/2.5*12)=6
      up (+z, r 0.15): k min(11, int(2.6/2.5*12)=12)=11; i int(2.35/5*25)=11 ..
            int(2.65/5*25)=13; j int(1.85/4*20)=9 .. int(2.15/4*20)=10
    The automatic outflow covers the east wall (the first fan blows +x), so
    the east fan's cells end OUTFLOW: 240 OUTFLOW, 9 + 6 INFLOW."""
    from scenes import FANS_AT_WALLS
    flags, fans, outs, opens = product_scene(FANS_AT_WALLS)
    sc = build_scene(FANS_AT_WALLS)
    want = [(24, 24, 0, 2, 10, 11), (0, 2, 19, 19, 4, 6), (11, 13, 9, 10, 11, 11)]
    assert [f[:6] for f in fans] == want
    assert [tuple(f[:6]) for f in sc.fans] == want
    assert outs == [(0, 1, 0, 19, 0, 11)] and opens == []
    assert fans[0][6:] == (3.0, 0.0, 0.0)
    assert fans[1][7] == 2.0 and fans[2][8] == 1.5
    for f in (flags, sc.flags):
        g = f.reshape(25, 20, 12)
        assert np.bincount(g.ravel(), minlength=4).tolist() == [5745, 0, 15, 240]
        assert np.all(g[24] == 3)
        assert np.all(g[0:3, 19, 4:7] == 2) and np.all(g[11:14, 9:11, 11] == 2)


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
