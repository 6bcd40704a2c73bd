"""Served kNN (nabo_hip_server_*): results in the mix equal results alone.

Clouds registered with a small HBM budget (so trees are evicted and rebuilt),
requests of several sizes, k and radii submitted from two threads while
others run; each served result must be bit-identical to the oracle's kd-tree
result on the same cloud and query set, and to the same request served
alone. test_served_c5_matches_oracle runs the served workload's own shape
(resident clouds of 10k to 2M points, k in {1, 8, 32}, maxRadius 0.5, an
evicting budget, with and without the host cache).
"""
import threading

import numpy as np
import pytest

from libnabo_amd import datasets
from tests.helpers import assert_bit_exact

pytestmark = pytest.mark.gpu


def clouds_and_requests(seed=9, count=6, requests=40):
    rng = np.random.default_rng(seed)
    clouds = []
    for i in range(count):
        n = int(rng.integers(2_000, 60_000))
        clouds.append((rng.random((n, 3)) * rng.uniform(1, 20)).astype(np.float32))
    reqs = []
    for _ in range(requests):
        c = int(rng.integers(0, count))
        m = int(rng.integers(1, 20_000))
        base = clouds[c][rng.integers(0, len(clouds[c]), m)]
        q = (base + rng.normal(scale=0.05, size=base.shape)).astype(np.float32)
        k = int(rng.choice([1, 8, 32]))
        r = float(rng.choice([0.5, np.inf]))
        reqs.append((c, q, k, r))
    return clouds, reqs


def test_served_equals_alone(nabo, oracle):
    clouds, reqs = clouds_and_requests()
    print("case served mix", flush=True)
    srv = nabo.Server(budget_bytes=3_000_000, slots=4)
    ids = [srv.add_cloud(c, 3, {"bucketSize": 8}) for c in clouds]
    handles = [None] * len(reqs)

    def submit(part):
        for i in part:
            c, q, k, r = reqs[i]
            handles[i] = srv.submit(ids[c], q, k, 0.0, 0, r)

    threads = [threading.Thread(target=submit, args=(range(j, len(reqs), 2),)) for j in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    results = [h.wait() for h in handles]
    st = srv.stats()
    assert st["completed"] == st["submitted"] == len(reqs)
    assert st["evictions"] > 0 and st["builds"] >= len(clouds)
    trees = [oracle.kdtree(c, bucket_size=8) for c in clouds]
    for i, ((c, q, k, r), (gi, gd)) in enumerate(zip(reqs, results)):
        oi, od = trees[c].knn(q, k, 0.0, 0, r)
        assert_bit_exact(gi, gd, oi, od, f"served request {i} (cloud {c}, k={k}, r={r})")
    # the first request alone on a fresh server
    alone = nabo.Server(slots=1)
    cid = alone.add_cloud(clouds[reqs[0][0]], 3, {"bucketSize": 8})
    ai, ad = alone.submit(cid, reqs[0][1], reqs[0][2], 0.0, 0, reqs[0][3]).wait()
    assert np.array_equal(ai, results[0][0]) and np.array_equal(ad, results[0][1])
    srv.close()
    alone.close()


def test_served_merges_same_cloud(nabo):
    cloud, _ = datasets.uniform_cube(30_000, 1, seed=2)
    srv = nabo.Server(slots=1)
    cid = srv.add_cloud(cloud, 3)
    rng = np.random.default_rng(3)
    qs = [rng.random((int(rng.integers(100, 3000)), 3)).astype(np.float32) for _ in range(30)]
    hs = [srv.submit(cid, q, 4) for q in qs]
    outs = [h.wait() for h in hs]
    plain = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_HIP, 3)
    for q, (gi, gd) in zip(qs, outs):
        oi, od = plain.knn(q, 4)
        assert np.array_equal(gi, oi) and np.array_equal(gd, od)
    assert srv.stats()["launches"] <= 30
    srv.close()


def test_served_argument_errors(nabo):
    cloud, _ = datasets.uniform_cube(100, 1, seed=2)
    srv = nabo.Server()
    cid = srv.add_cloud(cloud, 3)
    with pytest.raises(nabo.NaboError, match="minus 1"):
        srv.submit(cid, cloud[:5], 100)
    with pytest.raises(nabo.NaboError, match="unknown cloud"):
        srv.submit(cid + 1, cloud[:5], 1)
    srv.close()


def _submit_from_threads(srv, ids, reqs, threads=2):
    handles = [None] * len(reqs)

    def submit(part):
        for i in part:
            c, q, k, r = reqs[i]
            handles[i] = srv.submit(ids[c], q, k, 0.0, 0, r)

    ts = [threading.Thread(target=submit, args=(range(j, len(reqs), threads),))
          for j in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return [h.wait() for h in handles]


@pytest.mark.parametrize("host_cache", [False, True])
def test_served_c5_matches_oracle(nabo, oracle, host_cache):
    """The served workload's shape: 16 resident clouds of 10k to 1.87M points
    (5.0M in all), 48 request batches of 1k to 60k perturbed points, k in
    {1, 8, 32}, maxRadius 0.5, an HBM budget of half the trees' bytes (LRU
    eviction; with host_cache, evicted trees are re-admitted from pinned host
    copies). Every served result equals the oracle's bit for bit."""
    pool = datasets.served_pool(16, seed=8)
    reqs = datasets.served_requests(pool, 48, lo=1_000, hi=60_000, seed=9)
    assert max(len(c) for c in pool) > 1_000_000
    assert {k for _, _, k, _ in reqs} == {1, 8, 32}
    tree_bytes = 0
    for c in pool:  # bytes a resident tree takes (nabo_hip_get_info)
        nns = nabo.NearestNeighbourSearch(c[:20_000], nabo.SearchType.KDTREE_HIP, 3)
        tree_bytes += nns.info().device_bytes * len(c) / 20_000
        nns.close()
    print(f"case served C5 host_cache={host_cache}", flush=True)
    srv = nabo.Server(budget_bytes=int(tree_bytes // 2), slots=4,
                      host_cache_bytes=int(2 * tree_bytes) if host_cache else 0)
    ids = [srv.add_cloud(c, 3, {"bucketSize": 8}) for c in pool]
    results = _submit_from_threads(srv, ids, reqs)
    # a second wave on the clouds used first, which the budget has evicted
    # since: they are rebuilt, or re-admitted from the host cache
    again = reqs[:16]
    results += _submit_from_threads(srv, ids, again)
    reqs = reqs + again
    st = srv.stats()
    print(st, flush=True)
    assert st["completed"] == len(reqs) and st["evictions"] > 0
    assert st["builds"] + st["readmits"] > len({c for c, _, _, _ in reqs})
    if host_cache:
        assert st["readmits"] > 0
    srv.close()
    trees = {}
    for i, ((c, q, k, r), (gi, gd)) in enumerate(zip(reqs, results)):
        if c not in trees:
            trees[c] = oracle.kdtree(pool[c], bucket_size=8)
        oi, od = trees[c].knn(q, k, 0.0, 0, r)
        assert_bit_exact(gi, gd, oi, od, f"served request {i} (cloud {c}, n={len(pool[c])}, k={k})")
