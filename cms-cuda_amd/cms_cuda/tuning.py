"""Launch-shape table: for each batch-size bucket, the launch shapes that
``python -m cms_cuda.tune`` measured fastest on an MI355X (``tuning.json``
next to this file).  Every choice gives bit-identical results (the tuner
checks it); a missing table or bucket falls back to the library's built-in
rules (all fields 0).

Fields (see ``cms_tuning`` in include/cms_hip.h): trim_lanes, probe_count,
probe_slots, cell_sweep, search_split.
"""

import ctypes
import json
import os

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(HERE, "tuning.json")
FIELDS = ("trim_lanes", "probe_count", "probe_slots", "cell_sweep", "search_split")
# (name, smallest total atom count of the bucket); a batch falls in the last
# bucket whose bound it reaches
BUCKETS = (("single", 0), ("small", 20000), ("medium", 200000), ("large", 2000000))

_table = None


def bucket(total_atoms):
    name = BUCKETS[0][0]
    for nm, lo in BUCKETS:
        if total_atoms >= lo:
            name = nm
    return name


def table():
    global _table
    if _table is None:
        try:
            with open(TABLE_PATH) as fh:
                _table = json.load(fh)
        except (OSError, ValueError):
            _table = {}
    return _table


def choice(total_atoms):
    """The tuned fields for a batch of total_atoms atoms (zeros: rules)."""
    rec = table().get("buckets", {}).get(bucket(total_atoms), {})
    ch = rec.get("choice", {})
    return {f: int(ch.get(f, 0)) for f in FIELDS}


def default():
    return {f: 0 for f in FIELDS}


def apply(ctx, ch):
    """Set the fields of ch on a library context (_lib.Context)."""
    t = _lib.cms_tuning(*[int(ch.get(f, 0)) for f in FIELDS])
    ctx("cms_set_tuning", ctypes.byref(t))


def current(ctx):
    t = _lib.cms_tuning()
    ctx("cms_get_tuning", ctypes.byref(t))
    return {f: getattr(t, f) for f in FIELDS}
