"""Shared checks: the parity bar and the reference's query generators."""
from __future__ import annotations

import numpy as np

FLT_EPS = np.finfo(np.float32).eps


def exhaustive_queries(cloud: np.ndarray) -> np.ndarray:
    """createQuery with method -1 (tests/helpers.h:82-105): 2N queries, point
    i % N shifted by +sum(max-min) for the first N and by -sum(max-min) after,
    with the bounds of a BRUTE_FORCE search (true min/max per dimension)."""
    n = len(cloud)
    span = np.float32(0)
    lo = cloud.min(axis=0)
    hi = cloud.max(axis=0)
    for d in range(cloud.shape[1]):
        span = np.float32(span + np.float32(hi[d] - lo[d]))
    q = np.concatenate([cloud + span, cloud - span]).astype(np.float32)
    return q


def random_box_queries(cloud: np.ndarray, count: int, seed: int = 0) -> np.ndarray:
    """createQuery with method >= 0 (tests/helpers.h:106-113): uniform in the
    bounding box. The reference draws from libc rand(); a seeded numpy
    generator draws here."""
    rng = np.random.default_rng(seed)
    lo = cloud.min(axis=0)
    hi = cloud.max(axis=0)
    return (lo + rng.random((count, cloud.shape[1])) * (hi - lo)).astype(np.float32)


def assert_bit_exact(gi, gd, oi, od, what=""):
    """Indices and dists2 identical, bit for bit."""
    gi, gd, oi, od = map(np.asarray, (gi, gd, oi, od))
    assert gi.shape == oi.shape and gd.shape == od.shape, what
    bad = np.nonzero((gi != oi).any(axis=1) | (gd.view(np.uint32) != od.view(np.uint32)).any(axis=1))[0]
    if bad.size:
        r = bad[0]
        raise AssertionError(f"{what}: {bad.size} queries differ; first {r}: gpu {gi[r]} {gd[r]} "
                             f"vs oracle {oi[r]} {od[r]}")


def assert_same_up_to_ties(gi, gd, oi, od, what=""):
    """dists2 identical; indices identical except among equal distances."""
    gi, gd, oi, od = map(np.asarray, (gi, gd, oi, od))
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32)), f"{what}: dists2 differ"
    for r in np.nonzero((gi != oi).any(axis=1))[0]:
        for v in np.unique(gd[r]):
            sel = gd[r] == v
            assert sorted(gi[r][sel]) == sorted(oi[r][sel]) or v == gd[r][-1], \
                f"{what}: query {r} differs beyond ties"


def check_against_exact(cloud, query, idx, d2, exact_d2_sorted, flags=0, max_radius=np.inf,
                        rel=1e-6):
    """dists2 within `rel` of the FP64 exact k-th distances, and each returned
    index really at its reported distance (FP64 recheck)."""
    k = idx.shape[1]
    found = idx >= 0
    ex = exact_d2_sorted
    r2 = np.float64(np.float32(max_radius) * np.float32(max_radius))
    ex = np.where(ex <= r2 * (1 + 1e-6), ex, np.inf)
    both = np.isfinite(ex) & found
    denom = np.maximum(ex, 1e-30)
    err = np.abs(d2.astype(np.float64) - ex) / denom
    assert np.all(err[both] <= rel), f"max rel err {err[both].max()}"
    return err[both].max() if both.any() else 0.0
