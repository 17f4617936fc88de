"""Neighbour-by-neighbour validation of one search run against another on the same graphs
and queries, and the recall-QPS report over a sweep of configurations.

This is the report of cuda/reporte_comparacion1.py for this library's outputs:
    validate()         per configuration: short queries, same order, same set, ids missing,
                       tie-equal distances                           (:138-176, :117-129)
    best_qps() / summary_by_graph()
                       best QPS meeting each recall@10 floor, kernel-only and end to end,
                       and their ratios to the CPU                    (:182-199)
    write_report()     validation.csv, configurations.csv, summary.md and, when matplotlib
                       is importable, recall_qps.png and speedup.png  (:202-330)
The reference reads the CPU side from big-ann-benchmarks HDF5 files; h5py is not a
dependency here, so the command line reads both sides as ParlayANN neighbour files
(u32 nq, u32 k, then nq x k u32 ids), the format the reference's comparacion1 writes.

Command line:
    python -m pann.report --gpu DIR --cpu DIR [--base B.i8bin --queries Q.i8bin]
                          [--table configurations.csv] --out DIR
where DIR holds {graph}_Ls_{Ls}[_visit_{v}].bin files.
"""
import argparse
import csv
import math
import os
import re

import numpy as np

NO_NEIGHBOR = 0xFFFFFFFF
THRESHOLDS = (0.90, 0.95, 0.98)
GRAPHS = ("ParDiskANN", "ParHCNNG", "ParPyNNDescent", "ParHNSW")


def _distances(ids, rows, base, queries):
    q = queries[rows].astype(np.int64)[:, None, :]
    return ((base[ids[rows]].astype(np.int64) - q) ** 2).sum(-1)


def validate(ids, other, base=None, queries=None, chunk=4000):
    """Compares `ids` (nq x k, this run) with `other` (nq x k, the run checked against).

    A query whose row in `ids` holds NO_NEIGHBOR is short: the search ended with fewer
    than k results, and the reference's CPU side returns uninitialised memory there, so
    short queries are counted and left out of every comparison.  Over the other queries:
    same_order (identical rows), same_set (identical after sorting), ids_missing (ids of
    `other` absent from the row of `ids`) and, when base and queries are given,
    same_distances (rows whose sorted exact squared-L2 distances agree: ties broken
    differently still count).  Returns a dict of counts."""
    a = np.asarray(ids).astype(np.int64) & 0xFFFFFFFF
    b = np.asarray(other).astype(np.int64) & 0xFFFFFFFF
    if a.shape != b.shape or a.ndim != 2:
        raise ValueError(f"shapes differ: {a.shape} vs {b.shape}")
    short = np.any(a == NO_NEIGHBOR, axis=1)
    valid = ~short
    same_order = np.all(a == b, axis=1) & valid
    same_set = np.all(np.sort(a, axis=1) == np.sort(b, axis=1), axis=1) & valid
    missing = 0
    for s in range(0, a.shape[0], chunk):
        present = (b[s:s + chunk, :, None] == a[s:s + chunk, None, :]).any(axis=2)
        missing += int((~present)[valid[s:s + chunk]].sum())
    same_dist = None
    if base is not None and queries is not None:
        same_dist = int(same_set.sum())
        differ = np.where(valid & ~same_set)[0]
        for s in range(0, len(differ), chunk):
            rows = differ[s:s + chunk]
            da = np.sort(_distances(a, rows, base, queries), axis=1)
            db = np.sort(_distances(b, rows, base, queries), axis=1)
            same_dist += int(np.all(da == db, axis=1).sum())
    n = int(valid.sum())
    return {"queries": n, "short": int(short.sum()), "same_order": int(same_order.sum()),
            "same_set": int(same_set.sum()), "pct_same_set": 100.0 * same_set.sum() / max(n, 1),
            "same_distances": same_dist, "ids_missing": missing}


def best_qps(rows, recall_key, qps_key, threshold):
    """Largest rows[i][qps_key] among rows with rows[i][recall_key] >= threshold (NaN if none)."""
    sel = [r[qps_key] for r in rows
           if r.get(recall_key) is not None and r.get(qps_key) is not None
           and not _nan(r[recall_key]) and not _nan(r[qps_key]) and r[recall_key] >= threshold]
    return max(sel) if sel else math.nan


def summary_by_graph(rows, graphs=None, thresholds=THRESHOLDS):
    """Per graph and recall floor: best CPU QPS, best GPU end-to-end and kernel-only QPS, and
    their ratios to the CPU.  rows: dicts with graph, recall_cpu, recall_gpu, qps_cpu,
    qps_gpu_total, qps_gpu_kernel."""
    graphs = graphs or _ordered_graphs(rows)
    out = []
    for g in graphs:
        t = [r for r in rows if r["graph"] == g]
        if not t:
            continue
        for u in thresholds:
            cpu = best_qps(t, "recall_cpu", "qps_cpu", u)
            gk = best_qps(t, "recall_gpu", "qps_gpu_kernel", u)
            gt = best_qps(t, "recall_gpu", "qps_gpu_total", u)
            out.append({"graph": g, "recall_min": u, "qps_cpu": cpu, "qps_gpu_total": gt,
                        "qps_gpu_kernel": gk, "x_total": _ratio(gt, cpu), "x_kernel": _ratio(gk, cpu)})
    return out


def config_of_name(name):
    """..._Ls_75.bin -> (75, -1); ..._Ls_10_visit_3.bin -> (10, 3); None otherwise."""
    m = re.search(r"_Ls_(\d+)(?:_visit_(\d+))?\.bin$", name)
    if not m:
        return None
    return int(m.group(1)), (int(m.group(2)) if m.group(2) else -1)


def validate_dirs(gpu_dir, cpu_dir, base=None, queries=None):
    """validate() for every {graph}_Ls_{Ls}[_visit_{v}].bin present in both folders."""
    from .formats import read_neighbors
    out = []
    for name in sorted(os.listdir(cpu_dir)):
        cfg = config_of_name(name)
        if cfg is None or not os.path.exists(os.path.join(gpu_dir, name)):
            continue
        graph = name[:name.index("_Ls_")]
        v = validate(read_neighbors(os.path.join(gpu_dir, name)), read_neighbors(os.path.join(cpu_dir, name)),
                     base, queries)
        out.append({"graph": graph, "Ls": cfg[0], "visit": cfg[1], **v})
    return out


def write_report(out_dir, table, validation, title="Search: CPU vs GPU on the same graphs",
                 notes=()):
    """Writes configurations.csv (table), validation.csv and summary.md to out_dir, and the
    two plots when matplotlib is importable.  table rows: graph, Ls, visit, recall_cpu,
    recall_gpu, qps_cpu, qps_gpu_total, qps_gpu_kernel and optionally deterministic;
    validation rows: graph, Ls, visit and validate()'s counts.  Returns the file names."""
    os.makedirs(out_dir, exist_ok=True)
    files = []
    for name, rows in (("configurations.csv", table), ("validation.csv", validation)):
        if rows:
            keys = list(dict.fromkeys(k for r in rows for k in r))
            with open(os.path.join(out_dir, name), "w", newline="") as f:
                w = csv.DictWriter(f, fieldnames=keys)
                w.writeheader()
                w.writerows(rows)
            files.append(name)
    summary = summary_by_graph(table)
    lines = [f"# {title}", ""]
    lines += list(notes) + ([""] if notes else [])
    lines += ["## Best QPS meeting each recall@10 floor", "",
              "| graph | recall >= | QPS CPU | QPS GPU total | QPS GPU kernel | x total | x kernel |",
              "|---|---|---:|---:|---:|---:|---:|"]
    for r in summary:
        lines.append(f"| {r['graph']} | {r['recall_min']:.2f} | {_fmt(r['qps_cpu'])} | "
                     f"{_fmt(r['qps_gpu_total'])} | {_fmt(r['qps_gpu_kernel'])} | "
                     f"{_fmt(r['x_total'], 2)} | {_fmt(r['x_kernel'], 2)} |")
    lines += ["", "## Validation (GPU vs the other run)", "",
              "| graph | configs with equal recall | queries with the same k neighbours | "
              "queries with the same k distances | ids missing | short queries | deterministic configs |",
              "|---|---:|---:|---:|---:|---:|---:|"]
    for g in _ordered_graphs(table or validation):
        t = [r for r in table if r["graph"] == g]
        v = [r for r in validation if r["graph"] == g]
        eq = sum(1 for r in t if _equal_recall(r))
        det_rows = [r for r in t if r.get("deterministic") is not None]
        det = f"{sum(1 for r in det_rows if r['deterministic'])} of {len(t)}" if det_rows else "-"
        if v:
            nq = max(sum(r["queries"] for r in v), 1)
            pct = 100.0 * sum(r["same_set"] for r in v) / nq
            dists = [r["same_distances"] for r in v]
            pct_d = "-" if any(d is None for d in dists) else f"{100.0 * sum(dists) / nq:.4f} %"
            lines.append(f"| {g} | {eq} of {len(t)} | {pct:.4f} % | {pct_d} | "
                         f"{sum(r['ids_missing'] for r in v)} | {sum(r['short'] for r in v)} | {det} |")
        else:
            lines.append(f"| {g} | {eq} of {len(t)} | - | - | - | - | {det} |")
    lines += ["", "*Same k distances*: ids may differ, but the k neighbours lie at exactly the same",
              "distances (ties).  *Short queries*: the search ended with fewer than k results;",
              "they are counted and not compared neighbour by neighbour.", "",
              "## All configurations", "",
              "| graph | Ls | visit | recall CPU | recall GPU | QPS CPU | QPS GPU total | QPS GPU kernel | x total |",
              "|---|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for r in table:
        lines.append(f"| {r['graph']} | {r['Ls']} | {'-' if r.get('visit', -1) < 0 else r['visit']} | "
                     f"{_fmt(r.get('recall_cpu'), 5)} | {_fmt(r.get('recall_gpu'), 5)} | {_fmt(r.get('qps_cpu'))} | "
                     f"{_fmt(r.get('qps_gpu_total'))} | {_fmt(r.get('qps_gpu_kernel'))} | "
                     f"{_fmt(_ratio(r.get('qps_gpu_total'), r.get('qps_cpu')), 2)} |")
    with open(os.path.join(out_dir, "summary.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    files.append("summary.md")
    files += _plots(out_dir, table)
    return files


def _plots(out_dir, table):
    try:
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
    except ImportError:
        return []
    graphs = _ordered_graphs(table)
    if not graphs:
        return []
    cols = 2
    rows_n = (len(graphs) + cols - 1) // cols
    out = []
    for fname, ykind in (("recall_qps.png", "qps"), ("speedup.png", "x")):
        fig, axes = plt.subplots(rows_n, cols, figsize=(12, 4.5 * rows_n), squeeze=False)
        for ax, g in zip(axes.flat, graphs):
            t = [r for r in table if r["graph"] == g]
            if ykind == "qps":
                for key, rk, style, label in (("qps_cpu", "recall_cpu", "o-", "CPU"),
                                              ("qps_gpu_total", "recall_gpu", "s-", "GPU end to end"),
                                              ("qps_gpu_kernel", "recall_gpu", "^--", "GPU kernel")):
                    pts = sorted((r[rk], r[key]) for r in t if not _nan(r.get(rk)) and not _nan(r.get(key)))
                    if pts:
                        ax.plot([p[0] for p in pts], [p[1] for p in pts], style, label=label)
                ax.set_yscale("log")
                ax.set_ylabel("queries per second")
            else:
                for key, style, label in (("qps_gpu_total", "s-", "end to end"), ("qps_gpu_kernel", "^--", "kernel")):
                    pts = sorted((r["recall_cpu"], _ratio(r.get(key), r.get("qps_cpu"))) for r in t
                                 if not _nan(r.get("recall_cpu")) and not _nan(_ratio(r.get(key), r.get("qps_cpu"))))
                    if pts:
                        ax.plot([p[0] for p in pts], [p[1] for p in pts], style, label=label)
                ax.axhline(1.0, color="gray", lw=1)
                ax.set_ylabel("GPU / CPU")
            ax.set_title(g)
            ax.set_xlabel("recall@10")
            ax.grid(True, alpha=0.3)
            ax.legend(fontsize=8)
        for ax in list(axes.flat)[len(graphs):]:
            ax.set_visible(False)
        fig.tight_layout()
        fig.savefig(os.path.join(out_dir, fname), dpi=110)
        plt.close(fig)
        out.append(fname)
    return out


def _nan(x):
    return x is None or (isinstance(x, float) and math.isnan(x))


def _ratio(a, b):
    if _nan(a) or _nan(b) or not b:
        return math.nan
    return a / b


def _equal_recall(r):
    a, b = r.get("recall_cpu"), r.get("recall_gpu_sample", r.get("recall_gpu"))
    return not _nan(a) and not _nan(b) and abs(a - b) < 1e-9


def _ordered_graphs(rows):
    seen = list(dict.fromkeys(r["graph"] for r in rows))
    known = [g for g in GRAPHS if g in seen]
    return known + [g for g in seen if g not in known]


def _fmt(x, dec=0):
    return "-" if _nan(x) else f"{x:,.{dec}f}"


def main(argv=None):
    from .formats import read_i8bin
    p = argparse.ArgumentParser(description="neighbour-by-neighbour validation and recall-QPS report")
    p.add_argument("--gpu", required=True, help="folder of this run's neighbour files")
    p.add_argument("--cpu", required=True, help="folder of the other run's neighbour files")
    p.add_argument("--base", help="base vectors (.i8bin), for the tie-distance check")
    p.add_argument("--queries", help="query vectors (.i8bin), for the tie-distance check")
    p.add_argument("--table", help="CSV with graph, Ls, visit, recall_cpu, recall_gpu, qps_cpu, "
                                   "qps_gpu_total, qps_gpu_kernel[, deterministic]")
    p.add_argument("--out", required=True)
    a = p.parse_args(argv)
    base = read_i8bin(a.base) if a.base else None
    queries = read_i8bin(a.queries) if a.queries else None
    validation = validate_dirs(a.gpu, a.cpu, base, queries)
    table = []
    if a.table:
        with open(a.table, newline="") as f:
            for r in csv.DictReader(f):
                row = {"graph": r["graph"], "Ls": int(r["Ls"]), "visit": int(r.get("visit") or -1)}
                for k in ("recall_cpu", "recall_gpu", "qps_cpu", "qps_gpu_total", "qps_gpu_kernel"):
                    row[k] = float(r[k]) if r.get(k) not in (None, "") else math.nan
                if r.get("deterministic") not in (None, ""):
                    row["deterministic"] = r["deterministic"].strip().lower() in ("1", "true", "yes")
                table.append(row)
    files = write_report(a.out, table, validation)
    print(f"wrote {', '.join(files)} to {a.out}")


if __name__ == "__main__":
    main()
