"""The solver launch-shape autotuner's table and lookup (CPU only)."""
import json
import os

from flip3d_amd import autotune as at


def test_candidates_cap_below_builtin_block_count():
    assert at.default_blocks(24) == 14 and at.default_blocks(64) == 256 and at.default_blocks(128) == 256
    c24 = at.candidates(24)
    assert all(b == 0 for b, _ in c24) and sorted(t for _, t in c24) == [4, 8, 16]
    c64 = at.candidates(64)
    assert (0, 8) in c64 and (32, 4) in c64 and len(c64) == 12
    assert all(b == 0 or b < at.default_blocks(64) for b, _ in c64)


def test_launch_for_picks_nearest_size():
    table = {"shapes": {"learned": {"32": {"pcg_blocks": 32, "pcg_batch": 8},
                                    "64": {"pcg_blocks": 128, "pcg_batch": 4}}}}
    assert at.launch_for(table, "learned", 30) == (32, 8)
    assert at.launch_for(table, "learned", 60) == (128, 4)
    assert at.launch_for(table, "learned", 48) == (32, 8)  # equidistant: the smaller size
    assert at.launch_for(table, "reference", 64) == (0, 0)
    assert at.launch_for(None, "learned", 64) == (0, 0)


def test_committed_table_is_well_formed():
    table = at.load()
    if table is None:
        return
    json.dumps(table)
    for precond, entries in table["shapes"].items():
        assert precond in ("learned", "reference")
        for n, e in entries.items():
            assert 1 <= e["pcg_blocks"] <= 256 and 1 <= e["pcg_batch"] <= 64
            assert e["pcg_blocks"] <= at.default_blocks(int(n))
            assert e["ms"] <= e["default_ms"] + 1e-9
