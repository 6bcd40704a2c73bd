"""Seeded sweep of the search parameters against the oracle.

Each case draws a cloud shape (uniform, Gaussian clusters, an integer grid
whose points and distances tie, points on a line or a plane inside a larger
space, heavy-tailed coordinates), a size, a dimension (sometimes below the
cloud's row count: nabo.cpp:71 keeps the first dim rows), a bucket size, k,
epsilon, the option flags, a scalar or per-query radius and a query mix
(cloud points for the self-match rule, jittered points, points in and
outside the box), then checks one search type:
- KDTREE_LINEAR_HEAP / KDTREE_TREE_HEAP: the tree node for node, indices and
  dists2 bit for bit (heap order for the tree heap without SORT_RESULTS),
  and the TOUCH_STATISTICS count;
- BRUTE_FORCE (IndexHeapSTL, brute_force_cpu.cpp:45-103): bit for bit;
- BRUTE_FORCE_HIP: dists2 bit for bit, indices equal up to ties;
- batched trees (epsilon and radius drawn too): every cloud's results bit
  for bit;
- large k (49-200) and high dimensions (up to 100), for the kd-tree's
  larger k-best forms and the brute force's index-order scan.
The case's parameters are printed before it runs.
"""
import numpy as np
import pytest

from tests.helpers import assert_bit_exact, assert_same_up_to_ties
from tests.test_gpu_kdtree import assert_same_tree

pytestmark = pytest.mark.gpu

SELF, SORT = 1, 2
SHAPES = ("uniform", "clusters", "grid", "line", "plane", "heavy")


def _cloud(rng, shape, n, rows):
    if shape == "uniform":
        c = rng.random((n, rows)) * rng.uniform(0.01, 100)
    elif shape == "clusters":
        centers = rng.random((max(1, n // 500), rows)) * 50
        c = centers[rng.integers(0, len(centers), n)] + rng.normal(scale=0.3, size=(n, rows))
    elif shape == "grid":
        c = rng.integers(0, 5, (n, rows)).astype(np.float64)
    elif shape == "line":
        t = rng.random((n, 1))
        c = t * rng.normal(size=(1, rows)) + rng.normal(size=(1, rows))
    elif shape == "plane":
        c = rng.random((n, 2)) @ rng.normal(size=(2, rows))
    else:
        c = rng.standard_cauchy((n, rows)).clip(-1e4, 1e4)
    return np.ascontiguousarray(c, dtype=np.float32)


def _queries(rng, cloud, m):
    n, rows = cloud.shape
    lo, hi = cloud.min(axis=0), cloud.max(axis=0)
    span = np.maximum(hi - lo, 1e-3)
    parts = [
        cloud[rng.integers(0, n, m // 4)],                                   # self matches
        cloud[rng.integers(0, n, m // 4)] + rng.normal(scale=0.01, size=(m // 4, rows)) * span,
        lo + rng.random((m // 4, rows)) * span,                              # in the box
        lo + (rng.random((m - 3 * (m // 4), rows)) * 3 - 1) * span,          # around it
    ]
    return np.ascontiguousarray(np.concatenate(parts), dtype=np.float32)


def _draw(seed):
    rng = np.random.default_rng(1000 + seed)
    shape = SHAPES[seed % len(SHAPES)]
    if rng.random() < 0.3:
        n = int(np.exp(rng.uniform(0, np.log(2_000))))
    else:
        n = int(rng.integers(2_000, 40_000))
    rows = int(rng.choice([1, 2, 3, 3, 3, 3, 4, 5, 7, 9, 12, 40, 70]))
    if rows > 12:
        n = min(n, 5_000)  # above 64 dimensions the kd search keeps offsets in HBM
    dim = rows if rng.random() < 0.8 else int(rng.integers(1, rows + 1))
    bucket = int(rng.choice([2, 2, 3, 5, 8, 8, 8, 13, 32]))
    flags = int(rng.integers(0, 4))
    if n == 1:
        flags |= SELF
    # without ALLOW_SELF_MATCH, k is at most n - 1 (nabo.cpp:106-113)
    k = int(min(n if flags & SELF else n - 1, rng.choice([1, 2, 4, 7, 8, 10, 16, 24, 33, 48])))
    eps = float(rng.choice([0.0, 0.0, 0.0, 0.1, 0.5]))
    cloud = _cloud(rng, shape, n, rows)
    m = int(rng.integers(4, 3000))
    query = _queries(rng, cloud, m)
    # a radius around the median neighbour distance of a sample, or none
    radius, radii = np.inf, None
    pick = rng.random()
    if pick < 0.45:
        s = cloud[rng.integers(0, n, min(n, 64)), :dim]
        d = np.sqrt(((s[:, None, :] - s[None, :, :]) ** 2).sum(-1))
        radius = float(np.float32(np.median(d) * rng.uniform(0.05, 1.0) + 1e-6))
    elif pick < 0.6:
        radii = (rng.random(len(query)) * 2.0).astype(np.float32)
        radii[rng.random(len(query)) < 0.1] = np.inf
    return dict(shape=shape, n=n, rows=rows, dim=dim, bucket=bucket, k=k, eps=eps, flags=flags,
                m=len(query), radius=radius, radii=radii is not None), cloud, query, radii


def _coin(seed, p):
    """A draw independent of the case's own (the shape follows the seed)."""
    return np.random.default_rng(5000 + seed).random() < p


def _label(p):
    return " ".join(f"{key}={val}" for key, val in p.items())


@pytest.mark.parametrize("seed", range(120))
def test_kdtree_sweep(nabo, oracle, seed):
    p, cloud, query, radii = _draw(seed)
    heap = "stl" if _coin(seed, 0.35) else "linear"
    print(f"case sweep kd seed={seed} heap={heap} {_label(p)}", flush=True)
    st = nabo.SearchType.KDTREE_TREE_HEAP if heap == "stl" else nabo.SearchType.KDTREE_LINEAR_HEAP
    nns = nabo.NearestNeighbourSearch(cloud, st, p["dim"], {"bucketSize": p["bucket"]},
                                      creationOptionFlags=nabo.CreationOptionFlags.TOUCH_STATISTICS)
    otree = oracle.kdtree(cloud, dim=p["dim"], bucket_size=p["bucket"])
    assert_same_tree(nns, otree)
    gi, gd = nns.knn(query, p["k"], p["eps"], p["flags"], p["radius"], maxRadii=radii)
    oi, od, otouch, _ = otree.knn(query, p["k"], p["eps"], p["flags"], p["radius"], radii=radii,
                                  heap=heap, threads=8, stats=True)
    assert_bit_exact(gi, gd, oi, od, f"kd sweep {seed}")
    assert nns.last_touched == int(otouch.sum()), "touched count differs"
    nns.close()


@pytest.mark.parametrize("seed", range(120, 150))
def test_brute_force_sweep(nabo, oracle, seed):
    p, cloud, query, radii = _draw(seed)
    exact = _coin(seed, 0.5)
    print(f"case sweep bf seed={seed} stl={exact} {_label(p)}", flush=True)
    st = nabo.SearchType.BRUTE_FORCE if exact else nabo.SearchType.BRUTE_FORCE_HIP
    bf = nabo.NearestNeighbourSearch(cloud, st, p["dim"])
    flags = p["flags"] if exact else p["flags"] | SORT
    gi, gd = bf.knn(query, p["k"], 0.0, flags, p["radius"], maxRadii=radii)
    oi, od = oracle.bf_knn(cloud, query, p["k"], flags=flags, max_radius=p["radius"], radii=radii,
                           dim=p["dim"], threads=8)
    if exact:
        assert_bit_exact(gi, gd, oi, od, f"bf sweep {seed}")
    else:
        assert_same_up_to_ties(gi, gd, oi, od, f"bf sweep {seed}")
    bf.close()


@pytest.mark.parametrize("seed", range(150, 166))
def test_batched_sweep(nabo, oracle, seed):
    import torch

    rng = np.random.default_rng(2000 + seed)
    rows = int(rng.choice([2, 3, 3, 4]))
    bucket = int(rng.choice([2, 8, 13]))
    k = int(rng.choice([1, 4, 8, 16]))
    sizes = [int(max(k + 1, np.exp(rng.uniform(0, np.log(30_000))))) for _ in range(int(rng.integers(2, 12)))]
    clouds = [_cloud(rng, SHAPES[(seed + i) % len(SHAPES)], s, rows) for i, s in enumerate(sizes)]
    queries = [_queries(rng, c, int(rng.integers(1, 800)) + 4) for c in clouds]
    flags = int(rng.integers(0, 4))
    eps = float(rng.choice([0.0, 0.0, 0.1]))
    radius = float(rng.choice([np.inf, np.inf, 0.3, 2.0]))
    print(f"case sweep batch seed={seed} rows={rows} bucket={bucket} k={k} flags={flags} "
          f"eps={eps} radius={radius} sizes={sizes}", flush=True)
    off = np.concatenate([[0], np.cumsum([len(c) for c in clouds])]).astype(np.int64)
    qoff = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.int64)
    b = nabo.BatchSearch(torch.from_numpy(np.concatenate(clouds)).cuda(), off,
                         params={"bucketSize": bucket})
    gi, gd = b.knn(torch.from_numpy(np.concatenate(queries)).cuda(), torch.from_numpy(qoff).cuda(),
                   k, epsilon=eps, optionFlags=flags, maxRadius=radius)
    gi, gd = gi.cpu().numpy(), gd.cpu().numpy()
    for c, (cl, q) in enumerate(zip(clouds, queries)):
        oi, od = oracle.kdtree(cl, bucket_size=bucket).knn(q, k, eps, flags, radius, threads=8)
        assert_bit_exact(gi[qoff[c]:qoff[c + 1]], gd[qoff[c]:qoff[c + 1]], oi, od, f"cloud {c}")
    b.close()


def _draw_large(seed):
    """Large k (the 64/96/128-entry, output-row and sorted-row k-best forms)
    and high dimensions (offsets in HBM above 64)."""
    rng = np.random.default_rng(3000 + seed)
    shape = SHAPES[seed % len(SHAPES)]
    rows = int(rng.choice([2, 3, 3, 8, 70, 100]))
    n = int(rng.integers(300, 6_000 if rows > 8 else 20_000))
    flags = int(rng.integers(0, 4))
    k = int(min(n - 1, rng.choice([49, 64, 65, 96, 97, 128, 129, 200])))
    eps = float(rng.choice([0.0, 0.0, 0.2]))
    cloud = _cloud(rng, shape, n, rows)
    query = _queries(rng, cloud, int(rng.integers(8, 600)))
    radius = np.inf
    if rng.random() < 0.3:
        s = cloud[rng.integers(0, n, 64)]
        radius = float(np.float32(np.median(np.sqrt(((s[:, None] - s[None]) ** 2).sum(-1)))))
    return dict(shape=shape, n=n, rows=rows, k=k, eps=eps, flags=flags, m=len(query),
                radius=radius), cloud, query


@pytest.mark.parametrize("seed", range(24))
def test_large_k_sweep(nabo, oracle, seed):
    p, cloud, query = _draw_large(seed)
    heap = "stl" if _coin(7000 + seed, 0.5) else "linear"
    print(f"case sweep large kd seed={seed} heap={heap} {_label(p)}", flush=True)
    st = nabo.SearchType.KDTREE_TREE_HEAP if heap == "stl" else nabo.SearchType.KDTREE_LINEAR_HEAP
    nns = nabo.NearestNeighbourSearch(cloud, st, p["rows"], {"bucketSize": 8},
                                      creationOptionFlags=nabo.CreationOptionFlags.TOUCH_STATISTICS)
    otree = oracle.kdtree(cloud, bucket_size=8)
    gi, gd = nns.knn(query, p["k"], p["eps"], p["flags"], p["radius"])
    oi, od, otouch, _ = otree.knn(query, p["k"], p["eps"], p["flags"], p["radius"], heap=heap,
                                  threads=8, stats=True)
    assert_bit_exact(gi, gd, oi, od, f"large kd sweep {seed}")
    assert nns.last_touched == int(otouch.sum()), "touched count differs"
    nns.close()


@pytest.mark.parametrize("seed", range(24, 36))
def test_large_k_brute_force_sweep(nabo, oracle, seed):
    p, cloud, query = _draw_large(seed)
    exact = _coin(7000 + seed, 0.5)
    print(f"case sweep large bf seed={seed} stl={exact} {_label(p)}", flush=True)
    st = nabo.SearchType.BRUTE_FORCE if exact else nabo.SearchType.BRUTE_FORCE_HIP
    bf = nabo.NearestNeighbourSearch(cloud, st, p["rows"])
    flags = p["flags"] if exact else p["flags"] | SORT
    gi, gd = bf.knn(query, p["k"], 0.0, flags, p["radius"])
    oi, od = oracle.bf_knn(cloud, query, p["k"], flags=flags, max_radius=p["radius"], threads=8)
    if exact:
        assert_bit_exact(gi, gd, oi, od, f"large bf sweep {seed}")
    else:
        assert_same_up_to_ties(gi, gd, oi, od, f"large bf sweep {seed}")
    bf.close()
