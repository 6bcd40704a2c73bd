"""Autotuner for batched kd-trees: picks the bucket size, the query order and
the search's schedule.

For a batch of clouds, candidate configurations are built and searched on a
sample of the clouds, timed with device synchronisation:
  1. bucketSize x querySort: the fastest build + search total wins. The
     bucket size is the reference's own parameter (nabo/kdtree_cpu.cpp:230):
     it changes the tree, so results equal those of the reference configured
     with the chosen bucket size (for epsilon 0 the k-best set and distances
     do not depend on it; only the order among equal distances can);
  2. with that bucket size and order, the search's schedule: queries a
     wavefront claims at a time (waveQueries) x where the per-lane stacks
     live (stackPlace: LDS or HBM); the fastest search wins. These change
     neither the tree nor the results.
"""
from __future__ import annotations

import time

import numpy as np

DEFAULT_BUCKETS = (4, 6, 8, 12, 16)
DEFAULT_WAVE_QUERIES = (64, 128, 256)
DEFAULT_STACKS = ("auto", "hbm")


def _sample(offsets, qoffsets, clouds: int, rng):
    n = len(offsets) - 1
    pick = np.sort(rng.choice(n, size=min(clouds, n), replace=False))
    return pick


def _time_config(BatchSearch, sp, so, sq, sqo, k, params, reps):
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b = BatchSearch(sp, so, 3, 0, params)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    b.knn(sq, sqo, k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.knn(sq, sqo, k)
    torch.cuda.synchronize()
    knn_ms = (time.perf_counter() - t0) * 1e3 / reps
    b.close()
    return build_ms, knn_ms


def tune_batch(points, offsets, queries, qoffsets, k: int, buckets=DEFAULT_BUCKETS,
               wave_queries=DEFAULT_WAVE_QUERIES, stacks=DEFAULT_STACKS,
               sample_clouds: int = 64, reps: int = 2, seed: int = 0):
    """points/queries: torch CUDA tensors (rows = points); offsets/qoffsets:
    numpy int64 (clouds + 1). Returns (best, table): best is a dict with
    bucket_size, query_sort, wave_queries, stack_place, build_ms and knn_ms
    of the winner; table lists every configuration timed."""
    import torch

    from . import BatchSearch

    rng = np.random.default_rng(seed)
    pick = _sample(offsets, qoffsets, sample_clouds, rng)
    p_parts = [points[offsets[c]:offsets[c + 1]] for c in pick]
    q_parts = [queries[qoffsets[c]:qoffsets[c + 1]] for c in pick]
    sp = torch.cat(p_parts).contiguous()
    sq = torch.cat(q_parts).contiguous()
    so = np.concatenate([[0], np.cumsum([len(p) for p in p_parts])]).astype(np.int64)
    sqo = torch.from_numpy(
        np.concatenate([[0], np.cumsum([len(q) for q in q_parts])]).astype(np.int64)).cuda()
    table = []
    for bucket in buckets:
        for qsort in (1, 0):
            build_ms, knn_ms = _time_config(BatchSearch, sp, so, sq, sqo, k,
                                            {"bucketSize": bucket, "querySort": qsort}, reps)
            table.append({"bucket_size": bucket, "query_sort": qsort, "wave_queries": 64,
                          "stack_place": "auto", "build_ms": round(build_ms, 3),
                          "knn_ms": round(knn_ms, 3)})
    best = min(table, key=lambda r: r["build_ms"] + r["knn_ms"])
    for wq in wave_queries:
        for stack in stacks:
            if wq == 64 and stack == "auto":
                continue  # timed in stage 1
            params = {"bucketSize": best["bucket_size"], "querySort": best["query_sort"],
                      "waveQueries": wq, "stackPlace": stack}
            build_ms, knn_ms = _time_config(BatchSearch, sp, so, sq, sqo, k, params, reps)
            table.append({"bucket_size": best["bucket_size"], "query_sort": best["query_sort"],
                          "wave_queries": wq, "stack_place": stack,
                          "build_ms": round(build_ms, 3), "knn_ms": round(knn_ms, 3)})
    schedule = [r for r in table if r["bucket_size"] == best["bucket_size"]
                and r["query_sort"] == best["query_sort"]]
    winner = min(schedule, key=lambda r: r["knn_ms"])
    best = dict(best, wave_queries=winner["wave_queries"], stack_place=winner["stack_place"],
                knn_ms=winner["knn_ms"])
    return best, table


def params_of(best) -> dict:
    """The NearestNeighbourSearch / BatchSearch params of a tune_batch result."""
    return {"bucketSize": best["bucket_size"], "querySort": best["query_sort"],
            "waveQueries": best["wave_queries"], "stackPlace": best["stack_place"]}
