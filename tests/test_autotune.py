"""The autotune table: the committed file is valid and covers the benchmark's class, the build
turns it into the header the library compiles, the chooser keeps the built-in shape within the
noise margin, and the library's lookup (host code, no GPU call) reads the table it was built
with."""
import json
import math
import os

import pytest

from hcorepp_amd import build as B
from hcorepp_amd import autotune as AT


def test_committed_table_valid():
    rows = B.tune_entries()
    assert rows, "hcorepp_amd/autotune.json has no table"
    with open(B.TUNE_JSON) as f:
        doc = json.load(f)
    # the benchmark's class (nb = 512 tiles, ranks log-uniform in 8..192: mean K about 110,
    # 4096 updates: 16 per CU of 256)
    hit = [r for r in rows if 512 <= r[0] and (r[1] is None or 110 <= r[1]) and (r[2] is None or 16 <= r[2])]
    assert hit and hit[0][0] == 512
    # every table row names a point that was timed, and no point changed a rank or a factor
    for cls in doc["sweep"]:
        for p in cls["points"]:
            assert p["rank_diffs"] == 0 and p["max_rel_diff"] < 1e-10, (cls["nb"], p)
    for r in doc["table"]:
        cls = [c for c in doc["sweep"] if c["nb"] == r["max_rows"]]
        assert cls
        assert any(p["panel"] == r["panel"] and p["workgroups_per_cu"] == r["workgroups_per_cu"]
                   for c in cls for p in c["points"])


def test_header_generation(tmp_path):
    doc = {"table": [
        {"max_rows": 256, "max_mean_k": 40.5, "max_per_cu": 8.0, "panel": 8, "workgroups_per_cu": 1},
        {"max_rows": 256, "max_mean_k": 40.5, "max_per_cu": None, "panel": 8, "workgroups_per_cu": 2},
        {"max_rows": 256, "max_mean_k": None, "max_per_cu": None, "panel": 16, "workgroups_per_cu": 2},
        {"max_rows": 1024, "max_mean_k": None, "max_per_cu": None, "panel": 4, "workgroups_per_cu": 2}]}
    p = tmp_path / "t.json"
    p.write_text(json.dumps(doc))
    h = open(B.write_tune_header(str(tmp_path), str(p))).read()
    assert "constexpr int kTuneRows = 4;" in h
    for row in ("{256, 40.5, 8.0, 8, 1},", "{256, 40.5, -1.0, 8, 2},", "{256, -1.0, -1.0, 16, 2},",
                "{1024, -1.0, -1.0, 4, 2},"):
        assert row in h, row
    # no file: an empty table (one placeholder row, count 0)
    h = open(B.write_tune_header(str(tmp_path / "e"), str(tmp_path / "missing.json"))).read()
    assert "constexpr int kTuneRows = 0;" in h


@pytest.mark.parametrize("row", [
    {"max_rows": 600, "max_mean_k": None, "panel": 16, "workgroups_per_cu": 2},   # 16 only to 512 rows
    {"max_rows": 512, "max_mean_k": None, "panel": 12, "workgroups_per_cu": 2},   # not an instantiated width
    {"max_rows": 512, "max_mean_k": None, "panel": 8, "workgroups_per_cu": 5},    # residency beyond 4
    {"max_rows": 512, "max_mean_k": -3.0, "panel": 8, "workgroups_per_cu": 2},    # negative K bound
    {"max_rows": 512, "max_per_cu": 0.0, "panel": 8, "workgroups_per_cu": 2},     # empty size bound
])
def test_invalid_rows_rejected(tmp_path, row):
    p = tmp_path / "bad.json"
    p.write_text(json.dumps({"table": [row]}))
    with pytest.raises(ValueError):
        B.tune_entries(str(p))


def test_unsorted_rejected(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text(json.dumps({"table": [
        {"max_rows": 512, "max_mean_k": None, "panel": 8, "workgroups_per_cu": 2},
        {"max_rows": 256, "max_mean_k": None, "panel": 8, "workgroups_per_cu": 2}]}))
    with pytest.raises(ValueError):
        B.tune_entries(str(p))


def _cls(nb, ranks, mean_k, per_cu, pts, builtin=(16, 2)):
    return {"nb": nb, "ranks": list(ranks), "mean_k": mean_k, "per_cu": per_cu, "builtin": list(builtin),
            "points": [{"panel": p, "workgroups_per_cu": w, "ms": ms} for p, w, ms in pts]}


def test_choose_and_bounds(tmp_path):
    a = _cls(512, (4, 32), 30.0, 4, [(16, 2, 10.0), (8, 2, 9.9), (16, 1, 12.0)])   # 1 % faster: keep built-in
    b = _cls(512, (8, 192), 120.0, 4, [(16, 2, 10.0), (8, 2, 9.0), (4, 2, 9.5)])   # 10 % faster: take it
    b2 = _cls(512, (8, 192), 118.0, 16, [(16, 2, 10.0), (8, 1, 9.5)])
    c = _cls(512, (64, 256), 300.0, 4, [(16, 2, 10.0), (16, 1, 8.0)])
    rows = AT.build_table([c, b2, a, b], 0.02)
    assert [(r["panel"], r["workgroups_per_cu"]) for r in rows] == [(16, 2), (8, 2), (8, 1), (16, 1)]
    # K bounds between K-classes (the smaller mean K of a class's batches), size bounds within one
    assert rows[0]["max_mean_k"] == pytest.approx(math.sqrt(30 * 118), abs=0.01)
    assert rows[1]["max_mean_k"] == rows[2]["max_mean_k"] == pytest.approx(math.sqrt(118 * 300), abs=0.01)
    assert rows[3]["max_mean_k"] is None
    assert rows[0]["max_per_cu"] is None and rows[3]["max_per_cu"] is None
    assert rows[1]["max_per_cu"] == pytest.approx(8.0) and rows[2]["max_per_cu"] is None
    assert all(r["max_rows"] == 512 for r in rows)
    # the chooser's output is a valid table for the build
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"table": rows}))
    assert [e[3:] for e in B.tune_entries(str(p))] == [(16, 2), (8, 2), (8, 1), (16, 1)]


def _py_lookup(rows, nrows, mean_k, per_cu):
    for r, k, u, pb, wg in rows:
        if nrows <= r and (k is None or mean_k <= k) and (u is None or per_cu <= u):
            return True, pb, wg
    return False, None, None


def test_library_lookup_matches_table():
    if not os.path.exists(os.path.join(os.path.dirname(B.__file__), "libhcorepp_amd.so")):
        pytest.skip("library not built")
    import hcorepp_amd as H
    rows = B.tune_entries()
    for nrows in (64, 128, 200, 256, 300, 512, 700, 1024, 1500, 2048):
        for mk in (5.0, 30.0, 60.0, 110.0, 200.0, 400.0, 900.0):
            for pc in (0.5, 4.0, 7.9, 8.1, 16.0, 100.0):
                hit, pb, wg = H.autotune_lookup(nrows, mk, pc)
                ph, ppb, pwg = _py_lookup(rows, nrows, mk, pc)
                assert hit == ph, (nrows, mk, pc)
                if hit:
                    assert (pb, wg) == (ppb, pwg), (nrows, mk, pc)
                else:  # built-in rule: widest fitting panel, two workgroups per CU
                    assert pb == (16 if nrows <= 512 else 8 if nrows <= 1024 else 4) and wg == 2
