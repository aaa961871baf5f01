"""The launch-shape table (cms_cuda/tuning.json, written by
cms_cuda.tune on an MI355X): bucket bounds, field ranges, fallback to the
built-in rules."""

import json
import os

from cms_cuda import tuning


def test_buckets():
    assert tuning.bucket(1400) == "single" and tuning.bucket(19999) == "single"
    assert tuning.bucket(20000) == "small" and tuning.bucket(90000) == "small"
    assert tuning.bucket(450000) == "medium"
    assert tuning.bucket(7891638) == "large"


def test_table_fields_valid():
    if not os.path.exists(tuning.TABLE_PATH):
        return
    with open(tuning.TABLE_PATH) as fh:
        t = json.load(fh)
    assert set(t["buckets"]) <= {nm for nm, _ in tuning.BUCKETS}
    for name, rec in t["buckets"].items():
        ch = rec["choice"]
        assert set(ch) == set(tuning.FIELDS)
        assert ch["trim_lanes"] in (0, 1, 4, 32)
        assert ch["probe_count"] in (0, 1, 2)
        assert 0 <= ch["probe_slots"] <= 8
        assert ch["cell_sweep"] in (0, 1, 2)
        assert 0 <= ch["search_split"] <= 65536
        assert rec["tuned_ms"] <= rec["default_ms"]
        assert any(tr["choice"] == ch for tr in rec["trials"])


def test_missing_bucket_falls_back_to_rules(monkeypatch):
    monkeypatch.setattr(tuning, "_table", {"buckets": {}})
    assert tuning.choice(10 ** 6) == tuning.default()
