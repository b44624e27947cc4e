"""Result files (ff_room.io against the reference's io.py:24-141).

Pins: the archive holds exactly the reference's keys, with the scene as a
0-d JSON string; an archive written the reference's way (np.savez_compressed
with those keys) loads, including one without T_field/_T_mean (zeros, as the
reference's loader gives); loading never unpickles (an object array is
refused); the CSV export has the reference's header and (i, j, k) row order
at the cell centres; the experiment log's hash is that of the sorted JSON.
"""
import csv
import hashlib
import json

import numpy as np
import pytest

from ff_room.config import FanConfig, RoomConfig, SceneConfig
from ff_room.io import ExperimentLog, ResultStore
from ff_room.solver_bridge import SimResult

KEYS = {"u_field", "v_field", "w_field", "p_field", "T_field", "cell_type", "_steps",
        "_converged", "_velocity_change", "_divergence_max", "_T_mean", "_wall_time_s",
        "_config_json"}


def _result(nx=4, ny=3, nz=5, seed=0):
    rng = np.random.default_rng(seed)
    cfg = SceneConfig(room=RoomConfig(size=(2.0, 1.5, 1.0), grid=(nx, ny, nz)), fans=[FanConfig()])
    return SimResult(u_field=rng.standard_normal((nx + 1, ny, nz)),
                     v_field=rng.standard_normal((nx, ny + 1, nz)),
                     w_field=rng.standard_normal((nx, ny, nz + 1)),
                     p_field=rng.standard_normal((nx, ny, nz)),
                     T_field=20 + rng.standard_normal((nx, ny, nz)),
                     cell_type=rng.integers(0, 4, (nx, ny, nz), dtype=np.uint8), config=cfg,
                     steps=17, converged=True, velocity_change=1.5e-3, divergence_max=2e-9,
                     T_mean=21.25, wall_time_s=0.75)


def test_npz_round_trip_and_keys(tmp_path):
    r = _result()
    path = str(tmp_path / "r.npz")
    ResultStore.save_npz(r, path)
    with np.load(path, allow_pickle=False) as d:
        assert set(d.files) == KEYS
        assert d["_config_json"].shape == () and d["_config_json"].dtype.kind == "U"
        assert json.loads(str(d["_config_json"])) == r.config.to_dict()
    r2 = ResultStore.load_npz(path)
    for k in ("u_field", "v_field", "w_field", "p_field", "T_field", "cell_type"):
        a, b = getattr(r, k), getattr(r2, k)
        assert a.dtype == b.dtype and np.array_equal(a, b), k
    for k in ("steps", "converged", "velocity_change", "divergence_max", "T_mean", "wall_time_s"):
        assert getattr(r2, k) == getattr(r, k), k
    assert r2.config.room.size == r.config.room.size and r2.config.room.grid == r.config.room.grid


def test_loads_archives_without_temperature(tmp_path):
    """An archive of an older writer: no T_field, no _T_mean."""
    r = _result(seed=3)
    path = str(tmp_path / "old.npz")
    np.savez_compressed(path, u_field=r.u_field, v_field=r.v_field, w_field=r.w_field,
                        p_field=r.p_field, cell_type=r.cell_type, _steps=5, _converged=False,
                        _velocity_change=0.5, _divergence_max=0.25, _wall_time_s=1.0,
                        _config_json=np.array(json.dumps(r.config.to_dict())))
    r2 = ResultStore.load_npz(path)
    assert np.array_equal(r2.T_field, np.zeros_like(r.p_field)) and r2.T_mean == 0.0
    assert (r2.steps, r2.converged, r2.velocity_change) == (5, False, 0.5)


def test_load_refuses_pickled_objects(tmp_path):
    r = _result()
    path = str(tmp_path / "obj.npz")
    np.savez(path, u_field=np.array([{"a": 1}], dtype=object), v_field=r.v_field,
             w_field=r.w_field, p_field=r.p_field, T_field=r.T_field, cell_type=r.cell_type,
             _steps=1, _converged=False, _velocity_change=0.0, _divergence_max=0.0, _T_mean=0.0,
             _wall_time_s=0.0, _config_json=np.array(json.dumps(r.config.to_dict())))
    with pytest.raises(ValueError, match="allow_pickle"):
        ResultStore.load_npz(path)


def test_velocity_csv_columns_and_order(tmp_path):
    r = _result(nx=3, ny=2, nz=2, seed=5)
    path = str(tmp_path / "v.csv")
    ResultStore.save_velocity_csv(r, path)
    with open(path) as f:
        rows = list(csv.DictReader(f))
    assert list(rows[0].keys()) == ["x_m", "y_m", "z_m", "u_ms", "v_ms", "w_ms", "speed_ms"]
    assert len(rows) == 3 * 2 * 2
    # row of cell (i, j, k) = (1, 0, 1): index (i*ny + j)*nz + k
    row = rows[(1 * 2 + 0) * 2 + 1]
    assert float(row["x_m"]) == pytest.approx(2.0 * 1.5 / 3)
    assert float(row["y_m"]) == pytest.approx(1.5 * 0.5 / 2)
    assert float(row["z_m"]) == pytest.approx(1.0 * 1.5 / 2)
    uc = 0.5 * (r.u_field[1, 0, 1] + r.u_field[2, 0, 1])
    vc = 0.5 * (r.v_field[1, 0, 1] + r.v_field[1, 1, 1])
    wc = 0.5 * (r.w_field[1, 0, 1] + r.w_field[1, 0, 2])
    assert float(row["u_ms"]) == pytest.approx(uc)
    assert float(row["speed_ms"]) == pytest.approx(np.sqrt(uc ** 2 + vc ** 2 + wc ** 2))


def test_experiment_log(tmp_path):
    r = _result()
    log = ExperimentLog(str(tmp_path / "runs.jsonl"))
    assert log.load_all() == []
    h = log.append(r, notes="first")
    log.append(r, notes="second")
    entries = log.load_all()
    assert [e["notes"] for e in entries] == ["first", "second"]
    cfg = json.dumps(r.config.to_dict(), sort_keys=True)
    assert h == hashlib.sha256(cfg.encode()).hexdigest()[:8] == entries[0]["config_hash"]
    assert entries[0]["steps"] == 17 and entries[0]["T_mean"] == 21.25
