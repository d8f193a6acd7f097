"""Tile-shape tables: what the autotuner (gtf_pfock_autotune,
gtf_batch_autotune) chose for a workload, stored so that the workload runs
with those shapes later without tuning again.

One JSON file per workload under gtfock_amd/data/tuning/:
{"workload": name, "shapes": [21 x {units, unit_pq, tile_pq, grid,
exchanged, boys_lds, fp32}] in quartet-class order, "tuned": {how and where it was tuned,
and the autotuner's own before/after build times}}.
"""
import json
import os

SHAPE_KEYS = ("units", "unit_pq", "tile_pq", "grid", "exchanged", "boys_lds", "fp32")
NCLASSES = 21
TUNING_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tuning")


def path(name, directory=None):
    return os.path.join(directory or TUNING_DIR, name + ".json")


def flat(shapes):
    """21 shapes as the 105 ints of gtf_*_set_tile_shapes."""
    if len(shapes) != NCLASSES:
        raise ValueError("need %d class shapes, got %d" % (NCLASSES, len(shapes)))
    return [int(sh[k]) for sh in shapes for k in SHAPE_KEYS]


def load(name, directory=None):
    """(shapes, record) of a stored table, or (None, None) if there is none."""
    p = path(name, directory)
    if not os.path.exists(p):
        return None, None
    with open(p) as fh:
        rec = json.load(fh)
    shapes = rec["shapes"]
    flat(shapes)  # validates the count and the keys
    return shapes, rec


def save(name, shapes, tuned, directory=None):
    d = directory or TUNING_DIR
    os.makedirs(d, exist_ok=True)
    rec = {"workload": name, "shapes": [{k: int(sh[k]) for k in SHAPE_KEYS} for sh in shapes], "tuned": tuned}
    with open(path(name, d), "w") as fh:
        json.dump(rec, fh, indent=1)
    return path(name, d)
