"""Autotuner for batched kd-trees: picks bucket size and query ordering.

For a batch of clouds, each candidate configuration (bucketSize x
querySort) is built and searched on a sample of the clouds, timed with
device synchronisation, and the fastest build + search total wins. The
bucket size is the reference's own parameter (nabo/kdtree_cpu.cpp:230): it
changes the tree, so results equal those of the reference configured with
the chosen bucket size (for epsilon 0 the k-best set and distances do not
depend on it; only the order among equal distances can).
"""
from __future__ import annotations

import time

import numpy as np

DEFAULT_BUCKETS = (4, 6, 8, 12, 16)


def _sample(offsets, qoffsets, clouds: int, rng):
    n = len(offsets) - 1
    pick = np.sort(rng.choice(n, size=min(clouds, n), replace=False))
    return pick


def tune_batch(points, offsets, queries, qoffsets, k: int, buckets=DEFAULT_BUCKETS,
               sample_clouds: int = 64, reps: int = 2, seed: int = 0):
    """points/queries: torch CUDA tensors (rows = points); offsets/qoffsets:
    numpy int64 (clouds + 1). Returns (best, table): best is a dict with
    bucket_size, query_sort, build_ms and knn_ms of the winner."""
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
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b = BatchSearch(sp, so, 3, 0, {"bucketSize": bucket, "querySort": qsort})
            torch.cuda.synchronize()
            build_ms = (time.perf_counter() - t0) * 1e3
            b.knn(sq, sqo, k)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                b.knn(sq, sqo, k)
            torch.cuda.synchronize()
            knn_ms = (time.perf_counter() - t0) * 1e3 / reps
            table.append({"bucket_size": bucket, "query_sort": qsort,
                          "build_ms": round(build_ms, 3), "knn_ms": round(knn_ms, 3)})
            b.close()
    best = min(table, key=lambda r: r["build_ms"] + r["knn_ms"])
    return best, table
