"""The validation report (pann/report.py; reference: cuda/reporte_comparacion1.py:117-199),
pinned by hand-computed cases: ordered / set / missing-id counts, the short-query rule,
ties found by exact distances, and the best-QPS-at-recall selection."""
import csv
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NO = 0xFFFFFFFF


def test_validate_counts_by_hand():
    from pann import report
    a = np.array([[1, 2, 3],     # identical
                  [3, 1, 2],     # same set, other order
                  [1, 2, 9],     # one id differs
                  [1, NO, NO],   # short: left out
                  [7, 8, 9]], np.uint32)  # all three differ
    b = np.array([[1, 2, 3], [1, 2, 3], [1, 2, 3], [1, 2, 3], [1, 2, 3]], np.int64)
    v = report.validate(a, b)
    assert v == {"queries": 4, "short": 1, "same_order": 1, "same_set": 2, "pct_same_set": 50.0,
                 "same_distances": None, "ids_missing": 1 + 3}


def test_validate_reads_int32_as_uint32():
    """The reference's CPU files hold int32; -1 must compare equal to NO_NEIGHBOR's bits."""
    from pann import report
    a = np.array([[5, NO]], np.uint32)
    b = np.array([[5, -1]], np.int32)
    v = report.validate(a, b)
    assert v["short"] == 1 and v["queries"] == 0


def test_validate_ties_by_distance():
    """Points 1 and 2 are equidistant from the query (d = 4 each); point 3 is farther.
    A run returning {0, 2} where the other returned {0, 1} differs in ids but not in
    distances; {0, 3} differs in both."""
    from pann import report
    base = np.array([[0, 0], [2, 0], [0, 2], [3, 0]], np.int8)
    queries = np.array([[0, 0]] * 3, np.int8)
    a = np.array([[0, 2], [0, 3], [0, 1]], np.uint32)
    b = np.array([[0, 1], [0, 1], [0, 1]], np.uint32)
    v = report.validate(a, b, base, queries)
    assert v["same_set"] == 1           # row 2
    assert v["same_distances"] == 2     # rows 2 (same set) and 0 (tie)
    assert v["ids_missing"] == 2        # id 1 missing from rows 0 and 1


def test_validate_on_the_reference_fixture():
    """The reference's own neighbour file against itself: every complete row agrees; the
    rows holding NO_NEIGHBOR (130 missing ids, trailing) are the short queries."""
    from pann import formats, report
    a = formats.read_neighbors(os.path.join(HERE, "golden", "ParDiskANN_Ls_75.bin"))
    short = int(np.any(a == NO, axis=1).sum())
    v = report.validate(a, a.copy())
    assert v["short"] == short and short > 0
    assert v["queries"] == a.shape[0] - short
    assert v["same_order"] == v["same_set"] == v["queries"] and v["ids_missing"] == 0
    # one id changed in one complete row: one row fewer in each count, one id missing
    b = a.copy()
    r = int(np.flatnonzero(~np.any(a == NO, axis=1))[0])
    b[r, 4] = 999_999_999
    v2 = report.validate(a, b)
    assert v2["same_order"] == v["queries"] - 1 and v2["ids_missing"] == 1


def test_config_names():
    from pann import report
    assert report.config_of_name("ParDiskANN_Ls_75.bin") == (75, -1)
    assert report.config_of_name("ParHNSW_Ls_10_visit_3.bin") == (10, 3)
    assert report.config_of_name("ParHNSW_Ls_10.hdf5") is None


def test_best_qps_and_summary():
    from pann import report
    rows = [
        {"graph": "G", "recall_cpu": 0.89, "recall_gpu": 0.91, "qps_cpu": 900.0, "qps_gpu_total": 5e4, "qps_gpu_kernel": 9e4},
        {"graph": "G", "recall_cpu": 0.93, "recall_gpu": 0.93, "qps_cpu": 500.0, "qps_gpu_total": 4e4, "qps_gpu_kernel": 7e4},
        {"graph": "G", "recall_cpu": 0.96, "recall_gpu": 0.96, "qps_cpu": 300.0, "qps_gpu_total": 2e4, "qps_gpu_kernel": 3e4},
    ]
    # CPU: best among recall >= 0.90 is the 0.93 row (500); GPU: the 0.91 row (5e4 / 9e4)
    assert report.best_qps(rows, "recall_cpu", "qps_cpu", 0.90) == 500.0
    assert report.best_qps(rows, "recall_gpu", "qps_gpu_kernel", 0.90) == 9e4
    assert math.isnan(report.best_qps(rows, "recall_cpu", "qps_cpu", 0.98))
    s = report.summary_by_graph(rows)
    assert [r["recall_min"] for r in s] == [0.90, 0.95, 0.98]
    assert s[0]["x_total"] == 5e4 / 500.0 and s[0]["x_kernel"] == 9e4 / 500.0
    assert s[1]["qps_cpu"] == 300.0 and s[1]["qps_gpu_total"] == 2e4
    assert math.isnan(s[2]["x_total"])


def test_write_report_and_command_line(tmp_path):
    from pann import formats, report
    gdir, cdir, out = tmp_path / "g", tmp_path / "c", tmp_path / "out"
    gdir.mkdir()
    cdir.mkdir()
    ids = np.array([[0, 1], [0, 2], [1, NO]], np.uint32)
    other = np.array([[0, 1], [0, 1], [1, 2]], np.uint32)
    formats.write_neighbors(gdir / "ParDiskANN_Ls_10.bin", ids)
    formats.write_neighbors(cdir / "ParDiskANN_Ls_10.bin", other)
    formats.write_neighbors(cdir / "ParDiskANN_Ls_20.bin", other)   # no GPU file: skipped
    base = np.array([[0, 0], [2, 0], [0, 2]], np.int8)
    queries = np.zeros((3, 2), np.int8)
    formats.write_i8bin(tmp_path / "base.i8bin", base)
    formats.write_i8bin(tmp_path / "q.i8bin", queries)
    with open(tmp_path / "t.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["graph", "Ls", "visit", "recall_cpu", "recall_gpu", "qps_cpu", "qps_gpu_total",
                    "qps_gpu_kernel", "deterministic"])
        w.writerow(["ParDiskANN", 10, -1, 0.95, 0.95, 1000, 20000, 40000, "true"])
    report.main(["--gpu", str(gdir), "--cpu", str(cdir), "--base", str(tmp_path / "base.i8bin"),
                 "--queries", str(tmp_path / "q.i8bin"), "--table", str(tmp_path / "t.csv"),
                 "--out", str(out)])
    with open(out / "validation.csv", newline="") as f:
        v = list(csv.DictReader(f))
    assert len(v) == 1 and v[0]["Ls"] == "10"
    # row 0 identical, row 1 a tie (ids 1 and 2 are both at distance 4), row 2 short
    assert (v[0]["queries"], v[0]["short"], v[0]["same_set"], v[0]["same_distances"], v[0]["ids_missing"]) == \
        ("2", "1", "1", "2", "1")
    md = (out / "summary.md").read_text()
    assert "| ParDiskANN | 0.90 | 1,000 | 20,000 | 40,000 | 20.00 | 40.00 |" in md
    assert "| ParDiskANN | 1 of 1 | 50.0000 % | 100.0000 % | 1 | 1 | 1 of 1 |" in md
