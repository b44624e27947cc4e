"""The autotuner's table plumbing (CPU): shape classes, lookup, and how an
entry is written into ffr_params; the committed table holds valid entries."""
import json

from ff_room import autotune, capi


def test_shape_key_groups_by_cell_count_and_precision():
    assert autotune.shape_key(256, 204, 80, capi.PREC_MIXED) == "mixed:2^21"
    assert autotune.shape_key(256, 200, 80, capi.PREC_MIXED) == "mixed:2^21"
    assert autotune.shape_key(512, 408, 160, capi.PREC_MIXED) == "mixed:2^24"
    assert autotune.shape_key(256, 204, 80, capi.PREC_FP64).startswith("fp64:")


def test_apply_writes_params_and_block_cap(tmp_path):
    p = capi.default_params()
    before = capi.lib().ffr_get_max_blocks()
    try:
        autotune.apply(p, {"mg_pre_sweeps": 3, "inner_tol": 3e-3, "mg_tail_kb": 96, "max_blocks": 4096})
        assert (p.mg_pre_sweeps, p.inner_tol, p.mg_tail_kb) == (3, 3e-3, 96)
        assert capi.lib().ffr_get_max_blocks() == 4096
        assert capi.lib().ffr_set_max_blocks(100) != 0  # out of range is refused
    finally:
        capi.lib().ffr_set_max_blocks(before)
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"mixed:2^21": {"params": {"mg_pre_sweeps": 1}, "ms_per_step": 1.0}}))
    assert autotune.lookup(256, 204, 80, capi.PREC_MIXED, str(path)) == {"mg_pre_sweeps": 1}
    assert autotune.lookup(64, 64, 64, capi.PREC_MIXED, str(path)) is None


def test_committed_table_is_within_the_search_space():
    table = autotune.load()
    assert table, "tuning table missing"
    for key, e in table.items():
        assert key.startswith(("mixed:", "fp64:"))
        for k, v in e["params"].items():
            assert k in autotune.SPACE and v in autotune.SPACE[k], (key, k, v)
