"""bench.py's host logic (CPU): the one-GPU contract, the workload variants
and the oracle-side view of a prepared scene."""
import os
import subprocess
import sys

import numpy as np

from conftest import ROOT


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_stretched_workload_and_its_oracle_view():
    sys.path.insert(0, ROOT)
    import bench
    from ff_room import device
    cfg, desc = bench.workload("c2", "mixed", 1.01)
    assert cfg.room.stretch == 1.01 and "stretched" in desc
    cfg.room.grid = (32, 24, 12)  # a small room of the same kind
    ps = device.prepare(cfg, on_device=True)
    assert ps.dxv.max() / ps.dxv.min() > 1.1
    sc = bench._scene_arrays(ps)
    assert sc.faces is not None
    for got, want in zip(sc.spacings(), (ps.dxv, ps.dyv, ps.dzv)):
        assert np.allclose(got, want, rtol=1e-12, atol=0)
    cfg_u, _ = bench.workload("c2", "mixed")
    cfg_u.room.grid = (32, 24, 12)
    assert bench._scene_arrays(device.prepare(cfg_u, on_device=True)).faces is None
