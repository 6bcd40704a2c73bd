"""Benchmark: exact kd-tree kNN on one MI355X, LiDAR-like 1M x 1M, k=16.

Workload: 1,000,000 reference points on a synthetic 100 m scene (20 planes +
10 cylinders, sigma 0.01 m), 1,000,000 queries = the cloud under a random
rigid motion (<= 2 deg, <= 0.2 m) plus noise; k=16, epsilon=0,
maxRadius=1.0, bucketSize=8. One step = one knn call over all queries on the
resident tree (query ordering + search kernel), inputs already in HBM.

Prints one JSON line (rank 0). Roofline of the search kernel: algorithmic
bytes per launch = sum over queries of 12 + 8k + 16*P_q + 8*V_q, where P_q
(bucket entries scanned) and V_q (nodes visited) come from the oracle run on
the same inputs and tree (they equal the GPU's, whose traversal is the
same), divided by the kernel's mean duration measured with HIP events on its
launch stream. cpu_baseline: the oracle's OpenMP kd-tree on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per search launch on the default workload (lidar 1M x 1M, k=16,
# bucket 8), from two rocprofv3 counter passes scaled by a 256 MiB copy:
# FETCH_SIZE 55,681 KiB x 2.000 + WRITE_SIZE 130,497 KiB x 1.000
# (profiles/kd_search_traffic_lidar1M_k16.md).
MEASURED_TRAFFIC_BYTES = (55_681 * 2.000 + 130_497 * 1.000) * 1024
TRAFFIC_SOURCE = ("not measured in this run: rocprofv3 FETCH_SIZE / WRITE_SIZE passes of the "
                  "same command, profiles/kd_search_traffic_lidar1M_k16.md")
METRIC = "exact kNN queries/s, 1M 3D pts k=16; batched+served vs plain GPU; p99 ms"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--eps", type=float, default=0.0)
    ap.add_argument("--radius", type=float, default=1.0)
    ap.add_argument("--bucket", type=int, default=8)
    ap.add_argument("--no-query-sort", action="store_true")
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 256 MiB device-to-device copy (counter calibration)")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline line (no brute-force / batched / served sections)")
    ap.add_argument("--bf-forms", default="fp32,fp32_stl,bf16_m8,fp16_m8,bf16_m32,fp16_m32",
                    help="brute-force forms to run (fp32 is always run: the others are "
                         "compared with it)")
    ap.add_argument("--extras", default="brute_force,batched,served",
                    help="comma list of the sections to run after the headline")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    """Cores this process may run on (its affinity set), not the machine's."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cloud, query, k, eps, radius, bucket, budget_s):
    """Oracle kd-tree (OpenMP guided,32 over queries, all host cores) on a
    bounded sample of the same workload, and on one core over a 200k-query
    sample; also returns per-query P_q, V_q over all queries for the
    roofline."""
    import oracle as orc

    o = orc.default()
    threads = host_cores()
    t0 = time.perf_counter()
    tree = o.kdtree(cloud, bucket_size=bucket)
    build_s = time.perf_counter() - t0
    # Full pass once (gives P_q and V_q for every query), then repeat
    # passes while the time budget allows.
    t0 = time.perf_counter()
    oi, od, touched, visits = tree.knn(query, k, eps, 0, radius, stats=True, threads=threads)
    times = [time.perf_counter() - t0]
    while sum(times) < budget_s and len(times) < 5:
        t0 = time.perf_counter()
        tree.knn(query, k, eps, 0, radius, threads=threads)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    # one core, as libnabo without OpenMP: a 200k-query sample, median of 3
    sample = query[: min(len(query), 200_000)]
    one = []
    for _ in range(3):
        t0 = time.perf_counter()
        tree.knn(sample, k, eps, 0, radius, threads=1)
        one.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    o.kdtree(cloud, bucket_size=bucket)
    build_1 = time.perf_counter() - t0
    return {
        "value": len(query) / med,
        "unit": "queries/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle kd-tree, all {len(query)} queries x {len(times)} passes "
                  f"(median), build {build_s:.3f}s excluded",
        "build_s": build_s,
        "cpu_model": cpu_model(),
        "single_core": {"value": round(len(sample) / float(np.median(one)), 1),
                        "unit": "queries/s", "cores": 1,
                        "sample": f"{len(sample)} queries x 3 passes (median)",
                        "build_s": round(build_1, 3)},
    }, (oi, od, touched, visits)


def bench_lidar_variants(nns, cloud, query, d_query, radius, bucket, use_oracle: bool):
    """The other k / epsilon pairs of the LiDAR workload (k in {1, 16},
    epsilon in {0, 0.1}), and k = 16 with KDTREE_TREE_HEAP (IndexHeapSTL):
    time per knn call, search kernel time, parity with the oracle on all
    queries, and the kernel's HBM roofline fraction by the same
    algorithmic-byte count as the headline."""
    import torch

    import libnabo_amd as nabo
    import oracle as orc

    m = len(query)
    tree = orc.default().kdtree(cloud, bucket_size=bucket) if use_oracle else None
    out = {}
    stl = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_TREE_HEAP, 3,
                                      {"bucketSize": bucket})
    for k, eps, heap in ((1, 0.0, "linear"), (1, 0.1, "linear"), (16, 0.1, "linear"),
                         (16, 0.0, "stl")):
        index = nns if heap == "linear" else stl
        di = torch.empty((m, k), dtype=torch.int32, device="cuda")
        dd = torch.empty((m, k), dtype=torch.float32, device="cuda")
        for _ in range(2):
            index.knn_device(d_query, k, di, dd, eps, 0, radius)
        torch.cuda.synchronize()
        nabo.kernel_timing(True)
        nabo.kernel_timing_collect()
        reps = 10
        t0 = time.perf_counter()
        for _ in range(reps):
            index.knn_device(d_query, k, di, dd, eps, 0, radius)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        kms, launches = nabo.kernel_timing_collect()
        nabo.kernel_timing(False)
        kernel_ms = kms / max(1, launches)
        entry = {"ms_per_call": round(ms, 4), "queries_per_s": round(m / (ms / 1e3), 1),
                 "kernel_ms": round(kernel_ms, 4)}
        if tree is not None:
            oi, od, touched, visits = tree.knn(query, k, eps, 0, radius, heap=heap, stats=True,
                                               threads=os.cpu_count() or 1)
            gi, gd = di.cpu().numpy(), dd.cpu().numpy()
            entry["parity_vs_oracle"] = bool(np.array_equal(gi, oi) and
                                             np.array_equal(gd.view(np.uint32), od.view(np.uint32)))
            algo = m * (12 + 8 * k) + 16 * int(touched.sum()) + 8 * int(visits.sum())
            achieved = algo / (kernel_ms / 1e3) / 1e9
            entry["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(achieved / HBM_PEAK_GBS, 4)}
        name = f"k{k}_eps{eps}" + ("_tree_heap" if heap == "stl" else "")
        out[name] = entry
        log(f"lidar {name}: {entry}")
    stl.close()
    return out


def bench_brute_force(log_cpu: bool, forms=None):
    """Brute force, 200k x 33-D Dirichlet descriptors, 50k queries, k=8: exact
    fp32 path and bf16 / fp16 MFMA shortlists (m = 8, 32) with exact
    rescoring. Miss rate = share of queries whose k-best differs from the
    exact fp32 k-best; error = max relative |dists2 - FP64| over a sample."""
    import torch

    import libnabo_amd as nabo
    import oracle as orc
    from libnabo_amd import datasets

    cloud, query = datasets.dirichlet_descriptors(200_000, 50_000, seed=3)
    m, dim, k = len(query), cloud.shape[1], 8
    pairs = float(len(cloud)) * m
    dq = torch.from_numpy(query).cuda()
    out = {"workload": "dirichlet_200k_x_50k_d33_k8"}
    exact = None
    stl = None
    for prec, extra in (("fp32", 0), ("fp32_stl", 0), ("bf16", 8), ("fp16", 8), ("bf16", 32),
                        ("fp16", 32)):
        key = prec if prec.startswith("fp32") else f"{prec}_m{extra}"
        if forms is not None and key != "fp32" and key not in forms:
            continue
        # fp32_stl: BRUTE_FORCE, the reference's own IndexHeapSTL order
        nns = nabo.NearestNeighbourSearch(
            cloud, nabo.SearchType.BRUTE_FORCE if prec == "fp32_stl" else nabo.SearchType.BRUTE_FORCE_HIP,
            dim, {"bfPrecision": "fp32" if prec == "fp32_stl" else prec, "bfExtra": extra})
        di = torch.empty((m, k), dtype=torch.int32, device="cuda")
        dd = torch.empty((m, k), dtype=torch.float32, device="cuda")
        nns.knn_device(dq, k, di, dd)
        torch.cuda.synchronize()
        nabo.kernel_timing(True)
        nabo.kernel_timing_collect()
        t0 = time.perf_counter()
        reps = 2
        for _ in range(reps):
            nns.knn_device(dq, k, di, dd)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        kms, launches = nabo.kernel_timing_collect()
        nabo.kernel_timing(False)
        gi, gd = di.cpu().numpy(), dd.cpu().numpy()
        entry = {"ms": round(ms, 3), "queries_per_s": round(m / (ms / 1e3), 1)}
        if prec == "fp32_stl":
            stl = (gi, gd)
            entry["slowdown_vs_fp32"] = round(ms / out["fp32"]["ms"], 2)
        elif prec == "fp32":
            exact = (gi, gd)
            flops = 3.0 * dim * pairs
            entry["roofline"] = {"bound": "valu_fp32", "achieved": round(flops / (ms / 1e3) / 1e12, 2),
                                 "peak": 157.3, "unit": "TFLOP/s",
                                 "frac": round(flops / (ms / 1e3) / 1e12 / 157.3, 4)}
        else:
            # 2 * d flops per (query, point) pair: the distance's dot product
            # (padding columns and the rescoring are not counted)
            flops = 2.0 * dim * pairs
            entry["mfma_tflops"] = round(flops / (ms / 1e3) / 1e12, 1)
            entry["roofline"] = {"bound": "mfma", "achieved": entry["mfma_tflops"], "peak": 2500.0,
                                 "unit": "TFLOP/s",
                                 "frac": round(flops / (ms / 1e3) / 1e12 / 2500.0, 4)}
            miss = np.any(gd.view(np.uint32) != exact[1].view(np.uint32), axis=1)
            entry["miss_rate"] = float(miss.mean())
            entry["speedup_vs_fp32"] = round(out["fp32"]["ms"] / ms, 2)
        sample = slice(0, 2000)
        f64 = orc.default().dist2_f64(cloud, query[sample], gi[sample])
        entry["max_rel_err_vs_fp64"] = float(np.max(np.abs(gd[sample] - f64) / np.maximum(f64, 1e-30)))
        out[key] = entry
        nns.close()
        log(f"brute force {key}: {entry}")
    if log_cpu:
        o = orc.default()
        threads = host_cores()
        nq = 500
        t0 = time.perf_counter()
        oi, od = o.bf_knn(cloud, query[:nq], k, flags=2, threads=threads)
        dt = time.perf_counter() - t0
        # serial, as the reference's brute force (brute_force_cpu.cpp:81-98)
        ns = 60
        t0 = time.perf_counter()
        o.bf_knn(cloud, query[:ns], k, flags=0, threads=1)
        dt1 = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(ns / dt1, 1), "unit": "queries/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle brute force, serial as the reference, {ns} queries",
                               "cpu_model": cpu_model(),
                               "all_cores": {"value": round(nq / dt, 1), "cores": threads,
                                             "sample": f"{nq} queries, OpenMP over queries"}}
        out["fp32_parity_sample"] = bool(np.array_equal(exact[1][:nq].view(np.uint32),
                                                        od.view(np.uint32)))
        if stl is not None:
            # the STL-heap form returns the reference's own order: bit-identical rows
            si, sd = o.bf_knn(cloud, query[:nq], k, flags=0, threads=threads)
            out["fp32_stl_parity_sample"] = bool(
                np.array_equal(stl[0][:nq], si) and
                np.array_equal(stl[1][:nq].view(np.uint32), sd.view(np.uint32)))
    return out


def bench_batched():
    """512 uneven clouds (2k..200k points, k=8, eps 0): one batched build and
    one search launch (autotuned bucket size / query order) against
    per-cloud plain create + knn calls; both on device-resident data."""
    import ctypes as C

    import torch

    import libnabo_amd as nabo
    from libnabo_amd import autotune, datasets

    pairs = datasets.uneven_clouds(512, seed=4)
    clouds = [c for c, _ in pairs]
    queries = [q for _, q in pairs]
    off = np.concatenate([[0], np.cumsum([len(c) for c in clouds])]).astype(np.int64)
    qoff = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.int64)
    pts = torch.from_numpy(np.concatenate(clouds)).cuda()
    qs = torch.from_numpy(np.concatenate(queries)).cuda()
    dqoff = torch.from_numpy(qoff).cuda()
    k = 8
    best, table = autotune.tune_batch(pts, off, qs, qoff, k)
    log(f"autotune best {best}")
    params = autotune.params_of(best)
    # Host CLOCK_MONOTONIC at the start and end of each timed phase, so that
    # a kernel trace of the run (rocprofv3 stamps dispatches on the same
    # clock) can be split by phase.
    marks = {}
    # batched: build + search
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m0 = time.monotonic_ns()
    b = nabo.BatchSearch(pts, off, 3, 0, params)
    torch.cuda.synchronize()
    marks["batched_build"] = [m0, time.monotonic_ns()]
    build_ms = (time.perf_counter() - t0) * 1e3
    bi, bd = b.knn(qs, dqoff, k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m0 = time.monotonic_ns()
    bi, bd = b.knn(qs, dqoff, k, indices=bi, dists2=bd)
    torch.cuda.synchronize()
    marks["batched_knn"] = [m0, time.monotonic_ns()]
    knn_ms = (time.perf_counter() - t0) * 1e3
    # plain: one index and one knn call per cloud, same parameters
    lib = nabo.load_library()
    p = nabo.make_params(bucket_size=best["bucket_size"], query_sort=bool(best["query_sort"]),
                         wave_queries=best["wave_queries"], stack_place=best["stack_place"])
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    pi = torch.empty_like(bi)
    pd = torch.empty_like(bd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m0 = time.monotonic_ns()
    handles = []
    for c in range(len(clouds)):
        h = C.c_void_p()
        rc = lib.nabo_hip_create_device(C.c_void_p(pts[off[c]:].data_ptr()), int(off[c + 1] - off[c]),
                                        3, 3, 3, int(nabo.SearchType.KDTREE_HIP), 0, C.byref(p),
                                        stream, C.byref(h))
        assert rc == 0, lib.nabo_hip_last_error()
        handles.append(h)
    torch.cuda.synchronize()
    marks["plain_build"] = [m0, time.monotonic_ns()]
    plain_build_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    m0 = time.monotonic_ns()
    for c, h in enumerate(handles):
        q0, q1 = int(qoff[c]), int(qoff[c + 1])
        rc = lib.nabo_hip_knn_device(h, C.c_void_p(qs[q0:].data_ptr()), 3, 3, q1 - q0, k, 0.0, 0,
                                     float("inf"), None, C.c_void_p(pi[q0:].data_ptr()),
                                     C.c_void_p(pd[q0:].data_ptr()), None, stream)
        assert rc == 0, lib.nabo_hip_last_error()
    torch.cuda.synchronize()
    marks["plain_knn"] = [m0, time.monotonic_ns()]
    plain_knn_ms = (time.perf_counter() - t0) * 1e3
    for h in handles:
        lib.nabo_hip_destroy(h)
    same = bool(torch.equal(bi, pi) and torch.equal(bd, pd))
    total_q = int(qoff[-1])
    batched_total = build_ms + knn_ms
    plain_total = plain_build_ms + plain_knn_ms
    # One graph: every build and the search in one launch (the plan's stacks
    # are in LDS; its level budget is the data's depth + 1 + 4).
    plan_params = dict(params, stackPlace="auto")
    plan = nabo.BatchPlan(pts, off, qs, qoff, k, params=plan_params)
    plan.launch()
    complete, levels = plan.status()
    reps = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m0 = time.monotonic_ns()
    for _ in range(reps):
        plan.launch()
    torch.cuda.synchronize()
    marks["graph_launches"] = [m0, time.monotonic_ns()]
    graph_ms = (time.perf_counter() - t0) * 1e3 / reps
    complete, _ = plan.status()
    graph_same = bool(complete and torch.equal(plan.indices, bi) and torch.equal(plan.dists2, bd))
    plan.close()
    return {"workload": "512_uneven_clouds_2k_200k_k8", "clouds": len(clouds),
            "points": int(off[-1]), "queries": total_q,
            "autotune": {"best": best, "table": table},
            "batched_ms": {"build": round(build_ms, 2), "knn": round(knn_ms, 2),
                           "total": round(batched_total, 2)},
            "graph": {"ms_per_launch": round(graph_ms, 2), "levels": levels,
                      "note": "build + search of all clouds, one HIP graph launch",
                      "results_equal_batched": graph_same,
                      "speedup_vs_plain": round(plain_total / graph_ms, 2)},
            "plain_ms": {"build": round(plain_build_ms, 2), "knn": round(plain_knn_ms, 2),
                         "total": round(plain_total, 2)},
            "speedup_total": round(plain_total / batched_total, 2),
            "speedup_knn": round(plain_knn_ms / knn_ms, 2),
            "batched_queries_per_s": round(total_q / (knn_ms / 1e3), 1),
            "results_equal_plain": same,
            "phase_marks_monotonic_ns": marks}


def _serial_queue(arrivals, service):
    """Latencies (ms) and makespan (s) of a one-at-a-time server."""
    done = 0.0
    lat = []
    for a, s in zip(arrivals, service):
        start = max(a, done)
        done = start + s
        lat.append((done - a) * 1e3)
    return lat, done - arrivals[0]


def _served_run(nabo, pool, reqs, arrivals, budget, slots, warm, host_cache=0):
    import threading

    srv = nabo.Server(budget_bytes=budget, slots=slots, host_cache_bytes=host_cache)
    ids = [srv.add_cloud(c, 3) for c in pool]
    if warm:  # every tree resident before the stream starts
        for cid, c in zip(ids, pool):
            srv.submit(cid, c[:1], 1, 0.0, 1).wait()
    base = srv.stats()
    handles = [None] * len(reqs)
    submitted = [0.0] * len(reqs)

    def submitter():
        t_start = time.perf_counter()
        for i, (c, q, k, r) in enumerate(reqs):
            delay = arrivals[i] - arrivals[0] - (time.perf_counter() - t_start)
            if delay > 0:
                time.sleep(delay)
            submitted[i] = time.perf_counter()
            handles[i] = srv.submit(ids[c], q, k, 0.0, 0, r)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    lat = []
    for h in handles:
        h.wait()
        lat.append(h.latency_ms)
    # makespan: first submit to last completion on the device (the callers'
    # copy-out of results is not part of the service)
    makespan = max(s + l / 1e3 for s, l in zip(submitted, lat)) - submitted[0]
    st = srv.stats()
    st = {key: st[key] - (base[key] if key in ("builds", "evictions", "submitted", "completed",
                                                "launches", "merged_requests", "readmits")
                          else 0)
          for key in st}
    srv.close()
    return lat, makespan, st, handles


def bench_served(requests: int = 400, slots: int = 4):
    """Served stream over 64 clouds (10k..2M points): Poisson arrivals,
    batches of 1k..500k queries, k in {1, 8, 32}, maxRadius 0.5.

    warm: every tree resident; the server (4 stream slots, continuous
    admission, same-cloud merging) against the same requests served one at
    a time by plain knn calls on pre-built trees (a serial queue with the
    same arrival times, fed with each call's measured service time).
    budget: HBM budget of half the trees' total bytes, cold start: the server
    with LRU eviction against serial plain calls under the same budget and
    LRU policy (trees built on a miss, evicted least recently used first).
    saturated: all requests at once on the warm server, makespan against
    the sum of the plain calls' service times (throughput at full load).
    budget_host_cache: the same budget, with the server keeping evicted trees
    serialized in pinned host memory (room for all of them), so a re-admitted
    tree costs one copy instead of a build."""
    import libnabo_amd as nabo
    from libnabo_amd import datasets

    pool = datasets.served_pool(64, seed=5)
    reqs = datasets.served_requests(pool, requests, seed=6)
    total_q = sum(len(q) for _, q, _, _ in reqs)
    plain = {c: nabo.NearestNeighbourSearch(pool[c], nabo.SearchType.KDTREE_HIP, 3)
             for c in range(len(pool))}
    tree_bytes = {c: plain[c].info().device_bytes for c in plain}
    service, ref = [], []
    for c, q, k, r in reqs:
        t0 = time.perf_counter()
        ref.append(plain[c].knn(q, k, 0.0, 0, r))
        service.append(time.perf_counter() - t0)
    for v in plain.values():
        v.close()
    mean_s = float(np.mean(service))
    rate = 0.9 / mean_s  # offered load: 90% of what the warm serial server takes
    rng = np.random.default_rng(7)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(reqs)))
    serial_lat, serial_span = _serial_queue(arrivals, service)
    lat, span, st, handles = _served_run(nabo, pool, reqs, arrivals, 0, slots, True)
    same = all(np.array_equal(h.indices, ri) and np.array_equal(h.dists2, rd)
               for h, (ri, rd) in zip(handles, ref))
    pct = lambda v, p: round(float(np.percentile(v, p)), 2)  # noqa: E731
    out = {"workload": "served_64_clouds_10k_2M_poisson", "requests": len(reqs),
           "queries": int(total_q), "offered_load_req_per_s": round(rate, 2), "slots": slots,
           "warm": {"served": {"p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                               "queries_per_s": round(total_q / span, 1), "stats": st},
                    "serial_plain": {"p50_ms": pct(serial_lat, 50), "p99_ms": pct(serial_lat, 99),
                                     "queries_per_s": round(total_q / serial_span, 1),
                                     "mean_service_ms": round(mean_s * 1e3, 2)},
                    "results_equal_plain": bool(same)}}
    # budget: half of the trees' bytes, cold start, LRU on both sides
    budget = sum(tree_bytes.values()) // 2
    lat_b, span_b, st_b, handles_b = _served_run(nabo, pool, reqs, arrivals, budget, slots, False)
    cache, order, used, service_b = {}, [], 0, []
    for c, q, k, r in reqs:
        t0 = time.perf_counter()
        if c not in cache:
            while used + tree_bytes[c] > budget and order:
                victim = order.pop(0)
                used -= tree_bytes[victim]
                cache.pop(victim).close()
            cache[c] = nabo.NearestNeighbourSearch(pool[c], nabo.SearchType.KDTREE_HIP, 3)
            used += tree_bytes[c]
        else:
            order.remove(c)
        order.append(c)
        cache[c].knn(q, k, 0.0, 0, r)
        service_b.append(time.perf_counter() - t0)
    for v in cache.values():
        v.close()
    serial_lat_b, serial_span_b = _serial_queue(arrivals, service_b)
    same_b = all(np.array_equal(h.indices, ri) and np.array_equal(h.dists2, rd)
                 for h, (ri, rd) in zip(handles_b, ref))
    out["budget"] = {"budget_mb": round(budget / 2**20, 1),
                     "served": {"p50_ms": pct(lat_b, 50), "p99_ms": pct(lat_b, 99),
                                "queries_per_s": round(total_q / span_b, 1), "stats": st_b},
                     "serial_plain_lru": {"p50_ms": pct(serial_lat_b, 50),
                                          "p99_ms": pct(serial_lat_b, 99),
                                          "queries_per_s": round(total_q / serial_span_b, 1)},
                     "results_equal_plain": bool(same_b)}
    # saturated: every request submitted at once to the warm server; its
    # makespan against the serial sum of the plain calls' service times
    sat_arrivals = np.zeros(len(reqs))
    lat_s, span_s, st_s, handles_s = _served_run(nabo, pool, reqs, sat_arrivals, 0, slots, True)
    same_s = all(np.array_equal(h.indices, ri) and np.array_equal(h.dists2, rd)
                 for h, (ri, rd) in zip(handles_s, ref))
    serial_span_s = float(np.sum(service))
    out["saturated"] = {"served": {"queries_per_s": round(total_q / span_s, 1),
                                   "makespan_ms": round(span_s * 1e3, 2),
                                   "p50_ms": pct(lat_s, 50), "p99_ms": pct(lat_s, 99),
                                   "stats": st_s},
                        "serial_plain": {"queries_per_s": round(total_q / serial_span_s, 1),
                                         "makespan_ms": round(serial_span_s * 1e3, 2)},
                        "speedup": round(serial_span_s / span_s, 2),
                        "results_equal_plain": bool(same_s)}
    host_cache = 2 * sum(tree_bytes.values())
    lat_h, span_h, st_h, handles_h = _served_run(nabo, pool, reqs, arrivals, budget, slots, False,
                                                 host_cache)
    same_h = all(np.array_equal(h.indices, ri) and np.array_equal(h.dists2, rd)
                 for h, (ri, rd) in zip(handles_h, ref))
    out["budget_host_cache"] = {"budget_mb": round(budget / 2**20, 1),
                                "host_cache_mb": round(host_cache / 2**20, 1),
                                "served": {"p50_ms": pct(lat_h, 50), "p99_ms": pct(lat_h, 99),
                                           "queries_per_s": round(total_q / span_h, 1),
                                           "stats": st_h},
                                "results_equal_plain": bool(same_h)}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(every box here has one MI355X)", file=sys.stderr)
        sys.exit(2)
    import torch

    import libnabo_amd as nabo
    from libnabo_amd import datasets

    dev = torch.device("cuda:0")
    log(f"generating lidar scene n={args.n}")
    cloud, query = datasets.lidar_scene(args.n, seed=2)
    m = len(query)
    k = args.k

    d_cloud = torch.from_numpy(cloud).to(dev)
    d_query = torch.from_numpy(query).to(dev)
    d_idx = torch.empty((m, k), dtype=torch.int32, device=dev)
    d_d2 = torch.empty((m, k), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    log("building tree on the GPU")
    params = {"bucketSize": args.bucket, "querySort": 0 if args.no_query_sort else 1}
    # The process's first build also loads code objects and fills the
    # scratch pools; the second is the steady-state figure.
    t0 = time.perf_counter()
    nns = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_HIP, 3, params)
    build_ms_first = (time.perf_counter() - t0) * 1e3
    nns.close()
    t0 = time.perf_counter()
    nns = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_HIP, 3, params)
    build_ms_host = (time.perf_counter() - t0) * 1e3
    info = nns.info()
    # Build from device-resident points as well (excludes the host copy).
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    build_ms_dev = None
    try:
        import ctypes as C

        lib = nabo.load_library()
        h = C.c_void_p()
        p = nabo.make_params(bucket_size=args.bucket)
        rc = lib.nabo_hip_create_device(C.c_void_p(d_cloud.data_ptr()), len(cloud), 3, 3, 3,
                                        int(nabo.SearchType.KDTREE_HIP), 0, C.byref(p),
                                        C.c_void_p(torch.cuda.current_stream().cuda_stream),
                                        C.byref(h))
        torch.cuda.synchronize()
        build_ms_dev = (time.perf_counter() - t0) * 1e3
        if rc == 0:
            lib.nabo_hip_destroy(h)
    except Exception as e:  # measurement only
        log(f"device build timing failed: {e}")

    def step():
        nns.knn_device(d_query, k, d_idx, d_d2, args.eps, 0, args.radius)

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    nabo.kernel_timing(True)
    nabo.kernel_timing_collect()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        step()
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nabo.kernel_timing(False)
    kernel_ms_total, launches = nabo.kernel_timing_collect()
    ev_ms = start.elapsed_time(stop)
    ms_per_step = ev_ms / args.steps
    kernel_ms = kernel_ms_total / max(1, launches)
    value = m * args.steps / (ev_ms / 1e3)
    log(f"ms/step {ms_per_step:.3f} (wall {wall * 1e3 / args.steps:.3f}), search kernel "
        f"{kernel_ms:.3f} ms over {launches} launches")

    if args.calibrate_copy:
        # A plain copy of known size: 256 MiB read + 256 MiB written per
        # launch, for scaling FETCH_SIZE / WRITE_SIZE counters.
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    gi = d_idx.cpu().numpy()
    gd = d_d2.cpu().numpy()

    # The drop-in caller's rate: host query array in, host result arrays out
    # (nabo_hip_knn: PCIe copies through pinned staging, plus the search).
    host_ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        hi_, hd_ = nns.knn(query, k, args.eps, 0, args.radius)
        host_ms.append((time.perf_counter() - t0) * 1e3)
    host_path = {"ms_per_call": round(float(np.median(host_ms)), 3),
                 "queries_per_s": round(m / (float(np.median(host_ms)) / 1e3), 1),
                 "equal_device_path": bool(np.array_equal(hi_, gi) and np.array_equal(hd_, gd)),
                 "note": "nabo_hip_knn with numpy arrays in and out (median of 3 calls)"}
    del hi_, hd_
    log(f"host-buffer path: {host_path}")

    cpu = None
    p_sum = v_sum = None
    parity = None
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        cpu, (oi, od, touched, visits) = cpu_baseline(cloud, query, k, args.eps, args.radius,
                                                      args.bucket, args.cpu_seconds)
        p_sum = int(touched.sum())
        v_sum = int(visits.sum())
        parity = bool(np.array_equal(gi, oi) and np.array_equal(gd.view(np.uint32),
                                                                  od.view(np.uint32)))
        log(f"parity vs oracle: {parity}; P/q {p_sum / m:.2f} V/q {v_sum / m:.2f}")

    roofline = None
    if p_sum is not None:
        algo_bytes = m * (12 + 8 * k) + 16 * p_sum + 8 * v_sum
        achieved = algo_bytes / (kernel_ms / 1e3) / 1e9
        default_cfg = (args.n == 1_000_000 and k == 16 and args.bucket == 8 and args.eps == 0.0
                       and args.radius == 1.0 and not args.no_query_sort)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": int(MEASURED_TRAFFIC_BYTES) if default_cfg else None,
                    "traffic_source": (TRAFFIC_SOURCE if default_cfg else None),
                    "algorithmic_bytes": int(algo_bytes),
                    "kernel": "k_kd_search<3,16>", "kernel_ms": round(kernel_ms, 4)}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "lidar_1M_x_1M_k16", "n_points": args.n, "n_queries": m,
                   "k": k, "epsilon": args.eps, "max_radius": args.radius,
                   "bucket_size": args.bucket, "query_sort": not args.no_query_sort,
                   "tree_nodes": info.node_count, "tree_depth": info.depth},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "build_ms": {"first_in_process": round(build_ms_first, 2),
                     "host_cloud": round(build_ms_host, 2),
                     "device_cloud": None if build_ms_dev is None else round(build_ms_dev, 2)},
        "parity_vs_oracle": parity,
        "host_buffer_path": host_path,
    }
    if not args.no_extras:
        log("lidar: the other k / epsilon pairs")
        out["lidar_variants"] = bench_lidar_variants(nns, cloud, query, d_query, args.radius,
                                                     args.bucket, not args.no_cpu)
    extras = [] if args.no_extras else [e for e in args.extras.split(",") if e]
    unknown = set(extras) - {"brute_force", "batched", "served"}
    if unknown:
        raise SystemExit(f"unknown --extras sections: {sorted(unknown)}")
    if extras:
        nns.close()
        del d_query, d_idx, d_d2, d_cloud
        torch.cuda.empty_cache()
    if "brute_force" in extras:
        log("brute force: 200k x 33-D points, 50k queries, k=8")
        out["brute_force"] = bench_brute_force(not args.no_cpu, set(args.bf_forms.split(",")))
        torch.cuda.empty_cache()
    if "batched" in extras:
        log("batched: 512 uneven clouds, k=8")
        out["batched"] = bench_batched()
        torch.cuda.empty_cache()
    if "served" in extras:
        log("served: 64 resident clouds, Poisson request stream")
        out["served"] = bench_served()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
