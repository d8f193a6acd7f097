"""Tile-shape tables (gtfock_amd.tuning): the committed tables are complete
and well-formed, and a table survives a write / read round trip."""
import glob
import os

import pytest

from gtfock_amd import tuning


def _tables():
    return sorted(glob.glob(os.path.join(tuning.TUNING_DIR, "*.json")))


def test_committed_tables_are_well_formed():
    names = [os.path.basename(p)[:-5] for p in _tables()]
    # the bench's workloads: the headline molecule and the batch, both forms
    for need in ("c10h22", "c10h22_mixed", "batch", "batch_mixed"):
        assert need in names
    for name in names:
        shapes, rec = tuning.load(name)
        assert rec["workload"] == name
        assert len(shapes) == tuning.NCLASSES
        for sh in shapes:
            assert set(sh) == set(tuning.SHAPE_KEYS)
            assert min(sh["units"], sh["unit_pq"], sh["tile_pq"], sh["grid"]) >= 1
            assert sh["exchanged"] in (0, 1) and sh["boys_lds"] in (0, 1) and sh["fp32"] in (0, 1)
        t = rec["tuned"]
        # the autotuner keeps its shapes only when the build is not slower
        assert 0 < t["tuned_ms"] <= t["start_ms"]
        assert t["candidates"] > 0 and t["tuned_on"]


def test_round_trip_and_validation(tmp_path):
    shapes = [dict(units=256 + c, unit_pq=32, tile_pq=65536, grid=1024, exchanged=c % 2, boys_lds=1 - c % 2,
                   fp32=c % 3 != 0) for c in range(21)]
    tuning.save("x", shapes, {"start_ms": 2.0, "tuned_ms": 1.0}, directory=str(tmp_path))
    back, rec = tuning.load("x", directory=str(tmp_path))
    assert back == shapes and rec["tuned"]["tuned_ms"] == 1.0
    assert tuning.flat(back)[:7] == [256, 32, 65536, 1024, 0, 1, 0]
    assert tuning.load("absent", directory=str(tmp_path)) == (None, None)
    with pytest.raises(ValueError):
        tuning.flat(shapes[:20])
