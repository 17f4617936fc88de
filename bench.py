#!/usr/bin/env python3
"""QPS at recall@10 = 0.90 of ParlayANN's graph beam search on one MI355X.

Workload (default): 1,000,000 x 100 int8 points from a 1,024-component Gaussian mixture
(centers U[-60,60], sigma 20, rint, clip), 29,316 held-out queries, a Vamana graph
(R=64, L=128, alpha=1.2) built on the host, k=10, cut=1.35.  Ground truth is the exact
top-100 from the GPU brute-force kernel.  Ls is swept over the 18 standard values and
the first Ls whose tie-aware recall@10 reaches the target is timed: one step = one plain
launch over all queries (inputs already in HBM).  The roofline figure counts the
algorithmic bytes d*C + 4*E per query (C distance computations, E adjacency entries read)
over the kernel's event-timed duration.  The CPU baseline is the oracle (a C++
restatement of ParlayANN's beam search) on all host cores over a bounded query sample.

Prints one JSON line.  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "parlayann-gpu-busqueda_amd")
for _p in (PKG, REPO):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = ("QPS @ recall@10=0.90, 1M×100 int8; batched vs plain on 1 GPU; p99 ms")
LS_SWEEP = [10, 12, 15, 20, 25, 30, 35, 40, 50, 60, 75, 100, 125, 150, 200, 300, 400, 500]
# graph families under the names the reference's result files use
REPORT_NAMES = {"vamana": "ParDiskANN", "hcnng": "ParHCNNG", "pynndescent": "ParPyNNDescent", "hnsw": "ParHNSW"}
HBM_PEAK_GBS = 8000.0
TRAFFIC_SOURCES = ("beam_flat.hip", "device_common.h", "pann_internal.h")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def load_or_build_graph(base, args, cache):
    from pann import builders, formats
    name = f"vamana_n{len(base)}_c{args.ncomp}_R{args.R}_L{args.L}_a{args.alpha}_s7.graph"
    path = os.path.join(cache, name) if cache else None
    if path and os.path.exists(path):
        log(f"[bench] graph from cache {path}")
        return formats.read_graph(path)
    t = time.time()
    nthreads = min(16, host_cores())
    log(f"[bench] building Vamana R={args.R} L={args.L} alpha={args.alpha} on {nthreads} threads")
    g = builders.build_vamana(base, R=args.R, L=args.L, alpha=args.alpha, two_pass=True, seed=7,
                              nthreads=nthreads)
    log(f"[bench] graph built in {time.time() - t:.1f}s, mean degree {g.degrees.mean():.1f}")
    if path:
        os.makedirs(cache, exist_ok=True)
        formats.write_graph(path, g)
    return g


def time_launches(searcher, b, q, g, steps, warmup):
    """Returns (wall ms per launch, event ms per launch)."""
    from pann import _lib, search
    L = _lib.lib()
    for _ in range(warmup):
        searcher.launch(b, q, g)
    search.synchronize()
    import ctypes as C
    evs = []
    for _ in range(steps + 1):
        e = C.c_void_p()
        _lib.check(L.pann_event_create(C.byref(e)))
        evs.append(e.value)
    _lib.check(L.pann_device_synchronize())
    t0 = time.perf_counter()
    _lib.check(L.pann_event_record(evs[0], None))
    for i in range(steps):
        searcher.launch(b, q, g)
        _lib.check(L.pann_event_record(evs[i + 1], None))
    _lib.check(L.pann_device_synchronize())
    wall = (time.perf_counter() - t0) * 1e3 / steps
    ms = C.c_float()
    tot = 0.0
    for i in range(steps):
        _lib.check(L.pann_event_elapsed_ms(evs[i], evs[i + 1], C.byref(ms)))
        tot += ms.value
    for e in evs:
        L.pann_event_destroy(e)
    return wall, tot / steps


def end_to_end(searcher, b, q, g, queries, reps=5, launch=None):
    """The reference's end-to-end protocol (cuda/comparacion1.cu:67-75, 96-109): query rows
    from pinned host memory to the device (pann_vectors_copy_async, which also refreshes the
    norms), the launch, ids and distances back to pinned host memory, stream synchronize;
    one warm-up then best of `reps` by the host's steady clock.  Determinism: the ids of the
    warm-up run equal those of the last run."""
    from pann import _lib, search
    import ctypes as C
    L = _lib.lib()
    nq, dim = queries.shape
    k = searcher.cfg.k
    ptrs = []

    def pinned(nbytes, ctype, count):
        p = C.c_void_p()
        _lib.check(L.pann_host_alloc(nbytes, C.byref(p)))
        ptrs.append(p.value)
        return np.ctypeslib.as_array((ctype * count).from_address(p.value))
    rows = pinned(nq * dim, C.c_int8, nq * dim).reshape(nq, dim)
    rows[:] = queries
    ids = pinned(nq * k * 4, C.c_uint32, nq * k)
    dists = pinned(nq * k * 4, C.c_int32, nq * k)
    st = search.Stream()
    rows_p = rows.ctypes.data_as(C.c_void_p)
    ids_p, dists_p = ids.ctypes.data_as(C.c_void_p), dists.ctypes.data_as(C.c_void_p)

    def run():
        _lib.check(L.pann_vectors_copy_async(q.handle, rows_p, nq, st.ptr))
        if launch is None:
            searcher.launch(b, q, g, stream=st.ptr)
        else:
            launch(st.ptr)
        _lib.check(L.pann_searcher_download(searcher.handle, ids_p, dists_p, None, None, None, None,
                                            st.ptr))
        _lib.check(L.pann_synchronize(st.ptr))
    run()
    warm = ids.copy()
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        run()
        best = min(best, time.perf_counter() - t)
    deterministic = bool(np.array_equal(ids, warm))
    last = ids.copy().reshape(nq, k)
    st.free()
    for p in ptrs:
        L.pann_host_free(p)
    return {"e2e_qps": round(nq / best, 1), "e2e_ms": round(best * 1e3, 4), "reps": reps,
            "deterministic": deterministic}, last


def gather_ceiling(b, q, count=1 << 26, big_rows=8 << 20, seed=1):
    """Random 128-byte row gathers with v_dot4 (pann_gather_bench) over the 1M-row base
    (128 MB) and over a table of big_rows random rows (1 GiB at the default), GB/s counted
    as 128 bytes per distance; the chip's 256 MB last-level cache holds the first table,
    not the second."""
    from pann import search
    out = {}
    ms, _, gbs = search.gather_bench(b, q, count, seed=seed, reps=5)
    out[f"table_{b.n * 128 >> 20}MiB"] = {"gbs": round(gbs, 1), "ms": round(ms, 4), "count": count}
    rng = np.random.Generator(np.random.PCG64(seed))
    big = search.Vectors(rng.integers(-128, 128, size=(big_rows, 128), dtype=np.int8))
    try:
        ms, _, gbs = search.gather_bench(big, q, count, seed=seed, reps=5)
    finally:
        big.free()
    out[f"table_{big_rows * 128 >> 20}MiB"] = {"gbs": round(gbs, 1), "ms": round(ms, 4), "count": count}
    return out


def kernel_sources_sha():
    """sha256 over the sources of the flat search kernel: a stored traffic record applies
    only to the kernel these sources build."""
    import hashlib
    h = hashlib.sha256()
    for name in TRAFFIC_SOURCES:
        with open(os.path.join(PKG, "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def stored_traffic(path, Ls, n):
    """HBM-side bytes per launch from the committed counter record, or (None, why)."""
    if not os.path.exists(path):
        return None, "no counter record"
    tj = json.load(open(path))
    if tj.get("workload_Ls") != Ls or tj.get("n") != n:
        return None, f"counter record is for Ls={tj.get('workload_Ls')} n={tj.get('n')}"
    if tj.get("sources_sha256") != kernel_sources_sha():
        return None, "counter record predates the current kernel sources"
    return tj.get("hbm_bytes_per_launch"), (f"rocprofv3 FETCH_SIZE/WRITE_SIZE of {tj.get('kernel')} "
                                            f"scaled by the copy kernel's counters")


def write_traffic(args):
    """Builds the counter record from four rocprofv3 --pmc CSV files: FETCH_SIZE and
    WRITE_SIZE of the search kernel, then of the 1 GiB copy kernel (its true traffic is
    1 GiB read + 1 GiB written, which gives the scale factors)."""
    import csv
    import re

    def mean_counter(path, kernel_re, counter):
        vals = []
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") == counter and re.search(kernel_re, row.get("Kernel_Name", "")):
                    vals.append(float(row["Counter_Value"]))
        if not vals:
            raise SystemExit(f"no {counter} rows for /{kernel_re}/ in {path}")
        return sum(vals) / len(vals), len(vals)
    f_csv, w_csv, cf_csv, cw_csv = args.write_traffic
    kre = args.traffic_kernel
    fetch, nf = mean_counter(f_csv, kre, "FETCH_SIZE")
    write, nw = mean_counter(w_csv, kre, "WRITE_SIZE")
    cfetch, _ = mean_counter(cf_csv, "copyKernel", "FETCH_SIZE")
    cwrite, _ = mean_counter(cw_csv, "copyKernel", "WRITE_SIZE")
    gib_kib = float(1 << 20)
    fs, ws = gib_kib / cfetch, gib_kib / cwrite
    rec = {"kernel": kre, "workload_Ls": args.ls, "n": args.n, "nq": args.nq,
           "fetch_kib": round(fetch, 3), "write_kib": round(write, 3), "dispatches": [nf, nw],
           "fetch_scale": fs, "write_scale": ws,
           "hbm_bytes_per_launch": int(round((fetch * fs + write * ws) * 1024)),
           "sources_sha256": kernel_sources_sha()}
    with open(args.traffic_json, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec))


def cpu_baseline(base, graph, queries, Ls, k, seconds):
    """Oracle on all host cores: one warm-up pass over the query set, then repeated
    passes for about `seconds`, best pass reported (the reference's best-of-N protocol)."""
    from oracle import oracle as O
    cores = host_cores()
    nq = len(queries)
    params = O.make_params(nq, Ls, k)
    O.beam_flat(base, graph, queries, params, nthreads=cores)
    best, passes, t_all = 0.0, 0, time.perf_counter()
    while passes < 5 or time.perf_counter() - t_all < seconds:
        t = time.perf_counter()
        O.beam_flat(base, graph, queries, params, nthreads=cores)
        best = max(best, nq / (time.perf_counter() - t))
        passes += 1
        if passes >= 200:
            break
    spent = time.perf_counter() - t_all
    return {"value": round(best, 1), "unit": "QPS", "cores": cores, "kind": "port",
            "sample": f"all {nq} queries at Ls={Ls}, k={k}: warm-up + best of {passes} passes "
                      f"({spent:.1f}s) on {cores} threads; oracle = C++ restatement of "
                      f"ParlayANN beam_search"}


def homogeneous_batched(b, g, queries, Ls, k, steps, warmup, tune=True):
    """The same uniform workload through the batched persistent form (graph-replayed), in
    the shape the autotuner picks for it: returns (wall ms per launch, event ms per launch,
    results, shape, resident waves, autotune seconds)."""
    from pann import _lib, search
    import ctypes as C
    L = _lib.lib()
    nq = len(queries)
    params = search.make_params(nq, Ls, k)
    t = time.time()
    cfg = search.autotune(queries, params, Ls, k, base=b, graph=g)[0] if tune else None
    tune_s = time.time() - t
    st = search.Stream()
    batch = search.Batch(nq, Ls, k, base=b, graph=g, config=cfg)
    batch.set_queries(queries, params, st.ptr)
    batch.launch(st.ptr)
    batch.capture(st.ptr)
    for _ in range(warmup):
        batch.replay(st.ptr)
    st.synchronize()
    ev = []
    for _ in range(steps + 1):
        e = C.c_void_p()
        _lib.check(L.pann_event_create(C.byref(e)))
        ev.append(e.value)
    _lib.check(L.pann_device_synchronize())
    t0 = time.perf_counter()
    _lib.check(L.pann_event_record(ev[0], st.ptr))
    for i in range(steps):
        batch.replay(st.ptr)
        _lib.check(L.pann_event_record(ev[i + 1], st.ptr))
    st.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    ms, tot = C.c_float(), 0.0
    for i in range(steps):
        _lib.check(L.pann_event_elapsed_ms(ev[i], ev[i + 1], C.byref(ms)))
        tot += ms.value
    r = batch.download(st.ptr)
    for e in ev:
        L.pann_event_destroy(e)
    cfg, waves = batch.config()
    batch.free()
    st.free()
    return wall, tot / steps, r, cfg, waves, tune_s


def mixed_workload(queries, nq_total, seed=4):
    """65,536 queries resampled cyclically from the query set (each keeps its own qid):
    Ls = round(exp(U[ln 10, ln 400])), k in {1, 10, 100}, visit limit 2*Ls."""
    from pann import search
    rng = np.random.Generator(np.random.PCG64(seed))
    idx = np.arange(nq_total) % len(queries)
    Ls = np.rint(np.exp(rng.uniform(np.log(10), np.log(400), size=nq_total))).astype(np.int32)
    k = rng.choice(np.array([1, 10, 100], np.int32), size=nq_total)
    params = search.make_params(nq_total, Ls, k, visit_limit=2 * Ls, qids=idx.astype(np.uint32))
    return queries[idx], params


def batched_vs_plain(b, g, rows, params, steps, tune_sample=8192):
    """QPS of the autotuned batched form (its class launches HIP-graph replayed) against the plain form:
    one prepared searcher per distinct (Ls, k, visit), launched one after another as the
    reference's per-configuration protocol does.  Also checks every query agrees."""
    from pann import _lib, search
    import ctypes as C
    L = _lib.lib()
    nq = len(params)
    max_Ls, max_k = int(params["Ls"].max()), int(params["k"].max())
    t = time.time()
    cfg, tune_ms = search.autotune(rows[:tune_sample], params[:tune_sample], max_Ls, max_k,
                                   base=b, graph=g)
    log(f"[bench] autotune {time.time() - t:.1f}s -> waves/block {cfg.waves_per_block}, waves/query "
        f"{max(1, cfg.waves_per_query)}, LDS filter bits {cfg.lds_filter_bits}, visited slots "
        f"{cfg.visited_slots}, prefetch {cfg.prefetch}")
    st = search.Stream()
    batch = search.Batch(nq, max_Ls, max_k, base=b, graph=g, config=cfg)
    batch.set_queries(rows, params, st.ptr)
    batch.launch(st.ptr)
    batch.capture(st.ptr)
    st.synchronize()
    ev = []
    for _ in range(2):
        e = C.c_void_p()
        _lib.check(L.pann_event_create(C.byref(e)))
        ev.append(e.value)
    _lib.check(L.pann_device_synchronize())
    t0 = time.perf_counter()
    _lib.check(L.pann_event_record(ev[0], st.ptr))
    for _ in range(steps):
        batch.replay(st.ptr)
    _lib.check(L.pann_event_record(ev[1], st.ptr))
    st.synchronize()
    batched_ms = (time.perf_counter() - t0) * 1e3 / steps
    ms = C.c_float()
    _lib.check(L.pann_event_elapsed_ms(ev[0], ev[1], C.byref(ms)))
    batched_kernel_ms = ms.value / steps
    rb = batch.download(st.ptr)

    # plain form: prepared outside the timed region
    keys = np.stack([params["Ls"], params["k"], params["visit_limit"]], axis=1)
    groups = {}
    for i, key in enumerate(map(tuple, keys)):
        groups.setdefault(key, []).append(i)
    plain = []
    for (Ls, k, visit), members in groups.items():
        members = np.array(members)
        qv = search.Vectors(rows[members])
        qid = search.DeviceBuffer.from_array(params["qid"][members].astype(np.uint32))
        s = search.Searcher(g, len(members), search.SearchConfig(Ls=int(Ls), k=int(k), visit_limit=int(visit)))
        plain.append((members, qv, qid, s))
    _lib.check(L.pann_device_synchronize())
    for members, qv, qid, s in plain:  # warm-up
        s.launch(b, qv, g, qid, st.ptr)
    st.synchronize()
    t0 = time.perf_counter()
    for members, qv, qid, s in plain:
        s.launch(b, qv, g, qid, st.ptr)
    st.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3
    # the same launches spread over four streams, longest first: launch serialisation taken
    # out of the plain form (each searcher owns its scratch, so they may run side by side)
    streams = [search.Stream() for _ in range(4)]
    order = sorted(range(len(plain)), key=lambda i: -int(keys[plain[i][0][0]][0]) * len(plain[i][0]))
    _lib.check(L.pann_device_synchronize())
    t0 = time.perf_counter()
    for j, i in enumerate(order):
        members, qv, qid, s = plain[i]
        s.launch(b, qv, g, qid, streams[j % 4].ptr)
    for x in streams:
        x.synchronize()
    plain4_ms = (time.perf_counter() - t0) * 1e3
    for x in streams:
        x.free()
    mismatches = 0
    for members, qv, qid, s in plain:
        r = s.download()
        k = s.cfg.k
        mismatches += int(np.count_nonzero(
            (r.ids[:, :k] != rb.ids[members, :k]).any(axis=1) | (r.comps != rb.comps[members])))
        s.free(), qid.free(), qv.free()
    flags = int(np.count_nonzero(rb.flags & (search.FLAG_OVERFLOW | search.FLAG_CONFIG)))
    batch.free()
    st.free()
    for e in ev:
        L.pann_event_destroy(e)
    return {
        "workload": f"{nq} queries, Ls log-uniform [10,400], k in {{1,10,100}}, visit 2*Ls",
        "distinct_configs": len(groups),
        "qps_batched": round(nq / (batched_ms / 1e3), 1),
        "qps_plain": round(nq / (plain_ms / 1e3), 1),
        "ratio": round(plain_ms / batched_ms, 2),
        "batched_ms": round(batched_ms, 3), "batched_kernel_ms": round(batched_kernel_ms, 3),
        "plain_ms": round(plain_ms, 3),
        "qps_plain_4_streams": round(nq / (plain4_ms / 1e3), 1),
        "ratio_vs_plain_4_streams": round(plain4_ms / batched_ms, 2),
        "plain_4_streams_ms": round(plain4_ms, 3),
        "autotuned": {"waves_per_block": cfg.waves_per_block, "waves_per_query": max(1, cfg.waves_per_query),
                      "lds_filter_bits": cfg.lds_filter_bits, "visited_slots": cfg.visited_slots,
                      "prefetch": cfg.prefetch, "tune_ms": round(tune_ms, 3)},
        "queries_differing_from_plain": mismatches, "overflow_or_config_flags": flags,
    }


def serving(b, g, queries, gt_ids, gt_d, rates, duration, hnsw=None, max_batch=16384):
    """Poisson arrivals on the resident index (flat graph g, or the HNSW index): requests
    of 1..4096 queries, per-query Ls log-uniform in [10, 300], k in {10, 100}; p50/p99
    request latency (arrival to results in host memory), sustained QPS and recall@10 of
    the served queries.  max_batch 0: micro-batches sized by the LDS budget (one round of
    resident wavefronts)."""
    from pann import recall, search
    if hnsw is not None:
        srv = search.Server(slots=65536, max_batch=max_batch, streams=3, max_Ls=300, max_k=100, hnsw=hnsw)
    else:
        srv = search.Server(slots=65536, max_batch=max_batch, streams=3, max_Ls=300, max_k=100, base=b,
                            graph=g)
    out = []
    srv.run_poisson(queries, 50000, 0.3, seed=99)  # warm-up
    for i, rate in enumerate(rates):
        st, ids, rows = srv.run_poisson(queries, rate, duration, max_request=4096, Ls_lo=10, Ls_hi=300,
                                        ks=(10, 100), seed=7 + i, record_cap=200000)
        rec = recall.recall_at_k(gt_ids[rows], gt_d[rows], ids, 10) if len(rows) else None
        out.append({"offered_qps": rate, "sustained_qps": round(st["queries"] / max(st["busy_s"], 1e-9), 1),
                    "requests": int(st["requests"]), "queries": int(st["queries"]),
                    "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                    "max_ms": round(st["max_ms"], 3), "recall@10": round(rec, 5) if rec is not None else None})
        log(f"[bench] serving{' HNSW' if hnsw is not None else ''} {rate:,.0f} QPS offered: sustained {out[-1]['sustained_qps']:,.0f}, "
            f"p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms, recall@10 {out[-1]['recall@10']}")
    srv.free()
    return out


def calibrate_copy(nbytes, reps):
    """Runs the plain copy kernel `reps` times over `nbytes` (byte-counter calibration)."""
    from pann import _lib, search
    L = _lib.lib()
    src, dst = search.DeviceBuffer(nbytes), search.DeviceBuffer(nbytes)
    for _ in range(reps):
        _lib.check(L.pann_copy_kernel(dst.ptr, src.ptr, nbytes, None))
    search.synchronize()
    print(json.dumps({"calibrate_copy_bytes": nbytes, "reps": reps}))


def event_ms(launch):
    """HIP-event duration of one launch on the null stream."""
    from pann import _lib
    import ctypes as C
    L = _lib.lib()
    e0, e1 = C.c_void_p(), C.c_void_p()
    _lib.check(L.pann_event_create(C.byref(e0)))
    _lib.check(L.pann_event_create(C.byref(e1)))
    _lib.check(L.pann_event_record(e0.value, None))
    launch()
    _lib.check(L.pann_event_record(e1.value, None))
    ms = C.c_float()
    _lib.check(L.pann_event_elapsed_ms(e0.value, e1.value, C.byref(ms)))
    L.pann_event_destroy(e0.value)
    L.pann_event_destroy(e1.value)
    return ms.value


def sweep_graphs(args, base, queries, b, q, gt_ids, gt_df, k=10):
    """The four graph families at full size over the 18 Ls values: GPU kernel QPS (HIP
    events, plain launch over all queries), recall@10, distance computations per query,
    and the oracle's QPS on all host cores over a bounded query sample."""
    from oracle import oracle as O
    from pann import builders, recall, report, search
    cores = host_cores()
    nthreads = min(16, cores)
    sample = min(args.sweep_cpu_sample, len(queries))
    qs = queries[:sample]
    rows, table, validation = [], [], []
    for kind in args.sweep_graphs.split(","):
        t = time.time()
        if kind == "vamana":
            host = load_or_build_graph(base, args, args.cache)
        elif kind == "hcnng":
            host = builders.build_hcnng(base, num_trees=30, leaf_size=1000, mst_deg=3, seed=7,
                                        nthreads=nthreads)
        elif kind == "pynndescent":
            host = builders.build_pynndescent(base, K=40, num_trees=10, seed=7, nthreads=nthreads)
        elif kind == "hnsw":
            host = builders.build_hnsw(base, M=32, efc=128, seed=7, nthreads=nthreads)
        else:
            raise SystemExit(f"unknown graph kind {kind}")
        build_s = time.time() - t
        idx = search.HnswIndex(host, b) if kind == "hnsw" else search.FlatGraph(host)
        log(f"[sweep] {kind}: built in {build_s:.1f}s")
        for Ls in ([int(x) for x in args.sweep_ls.split(",")] if args.sweep_ls else LS_SWEEP):
            cfg = search.SearchConfig(Ls=Ls, k=k)
            if kind == "hnsw":
                s = search.HnswSearcher(idx, len(queries), cfg)
                launch = lambda: s.launch(q, idx)
            else:
                s = search.Searcher(idx, len(queries), cfg)
                launch = lambda: s.launch(b, q, idx)
            launch()
            search.synchronize()
            r = s.download()
            rec = recall.recall_at_k(gt_ids, gt_df, r.ids, k)
            ms = min(event_ms(launch) for _ in range(3))
            e2e = None
            if args.report:
                e2e_launch = (lambda st: s.launch(q, idx, stream=st)) if kind == "hnsw" else None
                e2e, _ = end_to_end(s, b, q, idx, queries, reps=3, launch=e2e_launch)
            s.free()
            cpu = 0.0
            if sample:
                params = O.make_params(sample, Ls, k)
                run = (lambda: O.search_hnsw(base, host, qs, params, nthreads=cores)) if kind == "hnsw" \
                    else (lambda: O.beam_flat(base, host, qs, params, nthreads=cores))
                ref = run()
                for _ in range(2):
                    t = time.perf_counter()
                    run()
                    cpu = max(cpu, sample / (time.perf_counter() - t))
                if args.report:
                    v = report.validate(r.ids[:sample, :k], ref.ids[:, :k], base, qs)
                    validation.append({"graph": REPORT_NAMES[kind], "Ls": Ls, "visit": -1, **v})
                    table.append({"graph": REPORT_NAMES[kind], "Ls": Ls, "visit": -1,
                                  "recall_cpu": recall.recall_at_k(gt_ids[:sample], gt_df[:sample], ref.ids, k),
                                  "recall_gpu": rec,
                                  "recall_gpu_sample": recall.recall_at_k(gt_ids[:sample], gt_df[:sample],
                                                                          r.ids[:sample], k),
                                  "qps_cpu": cpu, "qps_gpu_total": e2e["e2e_qps"],
                                  "qps_gpu_kernel": len(queries) / ms * 1e3,
                                  "deterministic": e2e["deterministic"]})
            row = {"graph": kind, "Ls": Ls, "recall": round(rec, 5), "gpu_kernel_ms": round(ms, 4),
                   "gpu_qps": round(len(queries) / ms * 1e3, 1), "cpu_qps": round(cpu, 1) if sample else None,
                   "e2e_qps": e2e["e2e_qps"] if e2e else None,
                   "comps": round(float(r.comps.mean()), 1), "hops": round(float(r.hops.mean()), 1),
                   "flagged": int(np.count_nonzero(r.flags & ~search.FLAG_SHORT)),
                   "alg_gbs": round(float(100 * r.comps.astype(np.float64).sum()
                                          + 4 * r.edges.astype(np.float64).sum()) / ms / 1e6, 1),
                   "build_s": round(build_s, 1)}
            rows.append(row)
            log(f"[sweep] {kind} Ls={Ls}: recall {rec:.5f}, {row['gpu_qps']:,.0f} QPS GPU "
                f"({ms:.3f} ms), {cpu:,.0f} QPS CPU, comps {row['comps']}")
        idx.free()
    if args.report:
        notes = [f"{len(base):,} x {base.shape[1]} int8 base, {len(queries):,} queries, k = {k}.",
                 f"CPU: the oracle on {cores} host threads over the first {sample:,} queries, best of 2 "
                 "passes; recall CPU and the validation are over that sample, recall GPU over all queries.",
                 "GPU kernel: HIP events, best of 3 launches, queries resident.  GPU total: query rows "
                 "host to device + search + ids and distances back, best of 3 (steady clock)."]
        files = report.write_report(args.report, table, validation, notes=notes)
        log(f"[sweep] report: {', '.join(files)} in {args.report}")
    return {"sweep": rows, "cpu_cores": cores, "cpu_sample_queries": sample,
            "n": len(base), "nq": len(queries)}


def hnsw_large(args):
    """HNSW at 10M x 128 int8 (GMM with 4,096 components), M=32, efC=200, mL=1/ln 32:
    100,000 queries at ef in {16, 64, 256}, k=10.  Throughput and recall@10 on the full
    query set; bit-exact check (ids, distances, counts, flags) against the oracle on the
    first `--w4-check` queries, whose oracle time gives the CPU QPS."""
    from oracle import oracle as O
    from pann import builders, datagen, recall, search
    n, nq, dim = args.w4_n, args.w4_nq, 128
    cores = host_cores()
    t = time.time()
    base, queries = datagen.gmm_dataset(n, nq, dim=dim, ncomp=args.w4_ncomp)
    log(f"[w4] data {base.shape} + {queries.shape} in {time.time() - t:.1f}s")
    b, q = search.Vectors(base), search.Vectors(queries)
    t = time.time()
    if args.w4_host_build:
        os.environ["PANN_BUILD_PROGRESS"] = "1"
        h = builders.build_hnsw(base, M=32, efc=200, seed=7, nthreads=min(16, cores))
    else:
        # the device build writes the host builder's file byte for byte
        h = builders.build_hnsw_gpu(b, M=32, efc=200, seed=7)
    build_s = time.time() - t
    log(f"[w4] HNSW built on the {'host' if args.w4_host_build else 'GPU'} in {build_s:.1f}s, "
        f"top level {int(h.level.max())}")
    idx = search.HnswIndex(h, b)
    t = time.time()
    gt_ids, gt_d = search.bruteforce_knn(b, q, 100)
    gt_df = gt_d.astype(np.float32)
    log(f"[w4] GPU ground truth top-100 in {time.time() - t:.2f}s")
    sub = min(args.w4_check, nq)
    rows = []
    for ef in (16, 64, 256):
        s = search.HnswSearcher(idx, nq, search.SearchConfig(Ls=ef, k=10))
        s.launch(q, idx)
        search.synchronize()
        r = s.download()
        ms = min(event_ms(lambda: s.launch(q, idx)) for _ in range(3))
        s.free()
        rec = recall.recall_at_k(gt_ids, gt_df, r.ids, 10)
        params = O.make_params(sub, ef, 10)
        O.search_hnsw(base, h, queries[:sub], params, nthreads=cores)
        t = time.perf_counter()
        o = O.search_hnsw(base, h, queries[:sub], params, nthreads=cores)
        cpu_qps = sub / (time.perf_counter() - t)
        same = ((r.ids[:sub] == o.ids[:, :10]).all(axis=1) & (r.dists[:sub] == o.dists[:, :10]).all(axis=1)
                & (r.comps[:sub] == o.comps) & (r.flags[:sub] == o.flags) & (r.edges[:sub] == o.edges)
                & (r.hops[:sub] == o.hops))
        row = {"ef": ef, "recall": round(rec, 5), "gpu_kernel_ms": round(ms, 3),
               "gpu_qps": round(nq / ms * 1e3, 1), "cpu_qps": round(cpu_qps, 1),
               "comps": round(float(r.comps.mean()), 1), "hops": round(float(r.hops.mean()), 1),
               "alg_gbs": round(float(dim * r.comps.astype(np.float64).sum()
                                      + 4 * r.edges.astype(np.float64).sum()) / ms / 1e6, 1),
               "checked": sub, "identical_to_oracle": int(same.sum()),
               "flagged": int(np.count_nonzero(r.flags & ~search.FLAG_SHORT))}
        rows.append(row)
        log(f"[w4] ef={ef}: recall {rec:.5f}, {row['gpu_qps']:,.0f} QPS GPU ({ms:.3f} ms), "
            f"{cpu_qps:,.0f} QPS CPU, {row['identical_to_oracle']}/{sub} identical to the oracle")
    idx.free(), q.free(), b.free()
    return {"workload": f"HNSW {n}x{dim} int8 GMM({args.w4_ncomp}), M=32 efC=200, {nq} queries, k=10",
            "build_s": round(build_s, 1), "built_on": "host" if args.w4_host_build else "gpu",
            "cpu_cores": cores, "rows": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=29_316)
    ap.add_argument("--ncomp", type=int, default=1024)
    ap.add_argument("--R", type=int, default=64)
    ap.add_argument("--L", type=int, default=128)
    ap.add_argument("--alpha", type=float, default=1.2)
    ap.add_argument("--target", type=float, default=0.90)
    ap.add_argument("--ls", type=int, default=0, help="time this Ls instead of sweeping")
    ap.add_argument("--cache", default=os.environ.get("PANN_BENCH_CACHE", ""))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the batched-vs-plain workload")
    ap.add_argument("--sweep", action="store_true",
                    help="instead of the headline line: the 18-Ls sweep over the four graph families")
    ap.add_argument("--sweep-graphs", default="vamana,hcnng,pynndescent,hnsw")
    ap.add_argument("--sweep-cpu-sample", type=int, default=2000, help="0: no CPU leg")
    ap.add_argument("--sweep-ls", default="", help="comma-separated Ls values (default: the 18)")
    ap.add_argument("--report", default="",
                    help="with --sweep: write the validation / recall-QPS report to this folder "
                         "(adds an end-to-end leg per configuration)")
    ap.add_argument("--hnsw-large", action="store_true",
                    help="instead of the headline line: HNSW at 10M x 128 (ef 16/64/256)")
    ap.add_argument("--w4-n", type=int, default=10_000_000)
    ap.add_argument("--w4-nq", type=int, default=100_000)
    ap.add_argument("--w4-ncomp", type=int, default=4096)
    ap.add_argument("--w4-check", type=int, default=5000)
    ap.add_argument("--w4-host-build", action="store_true",
                    help="build the 10M HNSW with the host builder instead of the device build")
    ap.add_argument("--no-tune", action="store_true",
                    help="time the batched form in its default shape instead of the autotuned one")
    ap.add_argument("--mixed-nq", type=int, default=65_536)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-rates", default="50000,200000,500000,1000000,2000000")
    ap.add_argument("--serving-seconds", type=float, default=2.0)
    ap.add_argument("--serving-max-batch", type=int, default=16384,
                    help="queries per serving micro-batch (0: the LDS budget, one round of resident wavefronts)")
    ap.add_argument("--no-serving-hnsw", action="store_true")
    ap.add_argument("--serving-hnsw-rates", default="50000,200000,500000,1000000,2000000")
    ap.add_argument("--traffic-json", default=os.path.join(REPO, "profiles", "traffic.json"))
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run the copy kernel over this many bytes (counter calibration)")
    ap.add_argument("--write-traffic", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV",
                                                         "COPY_WRITE_CSV"),
                    help="only write the counter record (--traffic-json) from rocprofv3 CSVs")
    ap.add_argument("--traffic-kernel", default="beamFlatFastKernel",
                    help="kernel-name regex of the counter record")
    ap.add_argument("--no-batched", action="store_true",
                    help="headline: time the plain launches only (plain-only kernel traces)")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end (H2D+search+D2H) leg")
    ap.add_argument("--no-gather", action="store_true", help="skip the random-gather ceiling")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy, 3)
        return
    if args.write_traffic:
        write_traffic(args)
        return

    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); the search shards by "
              f"query with no exchange step, so N GPUs would run N independent copies.",
              file=sys.stderr)
        sys.exit(2)

    from pann import datagen, recall, search

    if search.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if args.hnsw_large:
        print(json.dumps(hnsw_large(args)))
        return
    k = 10
    t = time.time()
    base, queries = datagen.gmm_dataset(args.n, args.nq, dim=100, ncomp=args.ncomp)
    log(f"[bench] data {base.shape} + {queries.shape} in {time.time() - t:.1f}s")
    graph = load_or_build_graph(base, args, args.cache)

    b, q, g = search.Vectors(base), search.Vectors(queries), search.FlatGraph(graph)
    t = time.time()
    gt_ids, gt_d = search.bruteforce_knn(b, q, 100)
    log(f"[bench] GPU ground truth top-100 in {time.time() - t:.2f}s")
    gt_ms = []
    for _ in range(3):
        t = time.perf_counter()
        again_ids, again_d = search.bruteforce_knn(b, q, 100)
        gt_ms.append((time.perf_counter() - t) * 1e3)
    gt_info = {"ms_best_of_3": round(min(gt_ms), 3), "k": 100,
               "int8_tops": round(2.0 * args.nq * args.n * 128 / (min(gt_ms) * 1e-3) / 1e12, 1),
               "peak_int8_tops": 5000.0,
               "deterministic": bool(np.array_equal(again_ids, gt_ids) and np.array_equal(again_d, gt_d))}
    log(f"[bench] brute-force top-100: {gt_info['ms_best_of_3']:.2f} ms (blocking call), "
        f"{gt_info['int8_tops']:,.0f} TOP/s over 128-byte rows")
    gt_df = gt_d.astype(np.float32)
    if args.sweep:
        print(json.dumps(sweep_graphs(args, base, queries, b, q, gt_ids, gt_df)))
        return

    sweep = []
    chosen = None
    ls_list = [args.ls] if args.ls else LS_SWEEP
    for Ls in ls_list:
        s = search.Searcher(g, args.nq, search.SearchConfig(Ls=Ls, k=k))
        s.launch(b, q, g)
        search.synchronize()
        r = s.download()
        rec = recall.recall_at_k(gt_ids, gt_df, r.ids, k)
        sweep.append({"Ls": Ls, "recall": round(rec, 5), "comps": float(r.comps.mean()),
                      "flags": int(np.count_nonzero(r.flags & ~search.FLAG_SHORT))})
        log(f"[bench] Ls={Ls} recall@10={rec:.5f} comps/q={r.comps.mean():.1f} hops/q={r.hops.mean():.1f}")
        if rec >= args.target or args.ls:
            chosen = (Ls, rec, s, r)
            break
        s.free()
    if chosen is None:
        raise SystemExit("target recall not reached by the sweep")
    Ls, rec, s, r = chosen
    plain_wall, plain_ev = time_launches(s, b, q, g, args.steps, args.warmup)
    log(f"[bench] plain Ls={Ls}: {plain_wall:.3f} ms/launch wall, {plain_ev:.3f} ms kernel, "
        f"{args.nq / plain_wall * 1e3:,.0f} QPS")
    e2e = None
    if not args.no_e2e:
        e2e, last_ids = end_to_end(s, b, q, g, queries)
        e2e["identical_to_first_launch"] = bool(np.array_equal(last_ids, r.ids))
        log(f"[bench] end to end Ls={Ls}: {e2e['e2e_ms']:.3f} ms best of {e2e['reps']}, "
            f"{e2e['e2e_qps']:,.0f} QPS, deterministic {e2e['deterministic']}, "
            f"same as the first launch {e2e['identical_to_first_launch']}")
    if args.no_batched:
        bat_wall, bat_ev, rbat, bcfg, bwaves, btune = None, None, None, None, None, 0.0
        same = None
    else:
        bat_wall, bat_ev, rbat, bcfg, bwaves, btune = homogeneous_batched(b, g, queries, Ls, k, args.steps,
                                                                           args.warmup, not args.no_tune)
        same = bool(np.array_equal(rbat.ids[:, :k], r.ids) and np.array_equal(rbat.comps, r.comps))
        log(f"[bench] batched Ls={Ls}: {bat_wall:.3f} ms/launch wall, {bat_ev:.3f} ms kernel, "
            f"{args.nq / bat_wall * 1e3:,.0f} QPS, identical to plain: {same}")
    form = ("batched persistent (HIP graph replay)"
            if (bat_wall is not None and bat_wall < plain_wall and same) else "plain")
    wall_ms, ev_ms = (bat_wall, bat_ev) if form != "plain" else (plain_wall, plain_ev)
    qps = args.nq / (wall_ms / 1e3)
    alg_bytes = float(100 * r.comps.astype(np.float64).sum() + 4 * r.edges.astype(np.float64).sum())
    achieved = alg_bytes / (ev_ms / 1e3) / 1e9
    traffic, traffic_note = stored_traffic(args.traffic_json, Ls, args.n)
    log(f"[bench] Ls={Ls}: {wall_ms:.3f} ms/launch wall, {ev_ms:.3f} ms kernel, {qps:,.0f} QPS, "
        f"{achieved:.1f} GB/s algorithmic; traffic: {traffic_note}")
    gather = None
    if not args.no_gather:
        gather = gather_ceiling(b, q)
        log(f"[bench] random-gather ceiling (128 B rows, v_dot4): "
            + ", ".join(f"{key} {v['gbs']:,.0f} GB/s" for key, v in gather.items()))

    mixed = None
    if not args.no_mixed:
        rows, params = mixed_workload(queries, args.mixed_nq)
        mixed = batched_vs_plain(b, g, rows, params, max(3, args.steps // 2))
        log(f"[bench] mixed: batched {mixed['qps_batched']:,.0f} QPS vs plain {mixed['qps_plain']:,.0f} "
            f"QPS ({mixed['ratio']}x over {mixed['distinct_configs']} configs; "
            f"{mixed['ratio_vs_plain_4_streams']}x over them on 4 streams), "
            f"{mixed['queries_differing_from_plain']} queries differ")

    serve = None
    if not args.no_serving:
        serve = serving(b, g, queries, gt_ids, gt_df, [float(x) for x in args.serving_rates.split(",")],
                        args.serving_seconds, max_batch=args.serving_max_batch)
    serve_hnsw = None
    if not args.no_serving and not args.no_serving_hnsw:
        from pann import builders
        t = time.time()
        hg = builders.build_hnsw_gpu(b, M=32, efc=128, seed=7)
        log(f"[bench] HNSW M=32 efC=128 built on the GPU in {time.time() - t:.1f}s for the serving loop")
        hidx = search.HnswIndex(hg, b)
        serve_hnsw = serving(b, None, queries, gt_ids, gt_df,
                             [float(x) for x in args.serving_hnsw_rates.split(",")], args.serving_seconds,
                             hnsw=hidx, max_batch=args.serving_max_batch)
        hidx.free()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(base, graph, queries, Ls, k, args.cpu_seconds)
        log(f"[bench] CPU baseline {cpu['value']:,.0f} QPS on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": round(qps, 1),
        "unit": "QPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 (exact int32 distances)",
        "data": "synthetic: Gaussian mixture, int8",
        "config": {
            "workload": (f"{args.n}x100 int8 GMM({args.ncomp}), {args.nq} queries, "
                         f"Vamana R={args.R} L={args.L} alpha={args.alpha}, k=10, cut=1.35, "
                         f"smallest Ls with recall@10 >= {args.target}; one step = one {form} "
                         f"launch over all queries"),
            "form": form, "graph": "vamana", "n": args.n, "nq": args.nq, "Ls": Ls,
            "recall@10": round(rec, 5),
            "plain_qps": round(args.nq / plain_wall * 1e3, 1), "plain_kernel_ms": round(plain_ev, 4),
            "end_to_end": e2e,
            "batched_qps": round(args.nq / bat_wall * 1e3, 1) if bat_wall else None,
            "batched_kernel_ms": round(bat_ev, 4) if bat_ev else None,
            "batched_identical_to_plain": same,
            "batched_shape": None if bcfg is None else {
                "waves_per_block": bcfg["waves_per_block"], "lds_filter_bits": bcfg["lds_filter_bits"],
                "visited_slots": bcfg["visited_slots"], "prefetch": bcfg["prefetch"],
                "waves_per_query": max(1, bcfg.get("waves_per_query", 1)),
                "resident_waves": bwaves, "autotuned": not args.no_tune, "autotune_s": round(btune, 2)},
            "kernel_ms": round(ev_ms, 4), "comps_per_query": round(float(r.comps.mean()), 1),
            "edges_per_query": round(float(r.edges.mean()), 1),
            "hops_per_query": round(float(r.hops.mean()), 1),
            "ls_sweep": sweep,
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_note,
                     "algorithmic_bytes_per_launch": int(alg_bytes),
                     "gather_ceiling": gather},
        "cpu_baseline": cpu,
        "ground_truth": gt_info,
        "batched_vs_plain": mixed,
        "serving": serve,
        "serving_hnsw": serve_hnsw,
    }
    s.free(), g.free(), q.free(), b.free()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
