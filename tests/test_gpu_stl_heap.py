"""IndexHeapSTL results on the GPU (nabo/index_heap.h:51-165): KDTREE_TREE_HEAP
and BRUTE_FORCE against the oracle, whose heap is libstdc++'s own
std::push_heap / std::pop_heap / std::sort_heap, as the reference's is.

Bar: indices and dists2 identical, element for element, in the returned order
(heap order without SORT_RESULTS, sort_heap order with it), including which
of equal distances are kept and the (-1, +inf) sentinel of a heap that never
filled up; the TOUCH_STATISTICS count equals the linear heap's (both heaps
admit the same distances, so the descent is the same).
"""
import numpy as np
import pytest

from libnabo_amd import datasets

pytestmark = pytest.mark.gpu

SORT = 2
SELF = 1


def _ties(n, m, seed):
    # integer grid: many equal coordinates, duplicate points and equal distances
    rng = np.random.default_rng(seed)
    cloud = rng.integers(0, 12, (n, 3)).astype(np.float32)
    query = (rng.integers(0, 12, (m, 3)) + rng.choice([0.0, 0.5], (m, 3))).astype(np.float32)
    return cloud, query


KD_CASES = {
    "uniform3d": lambda: datasets.uniform_cube(20_000, 3_000, seed=21),
    "lidar": lambda: (lambda c, q: (c, q[::20]))(*datasets.lidar_scene(60_000, seed=22)),
    "ties": lambda: _ties(6_000, 2_000, 23),
    "scan2d": lambda: (datasets.load_reference_scan("scan.2d"),
                       datasets.load_reference_scan("scan.2d")[::2] + np.float32(0.01)),
    "uniform5d": lambda: datasets.uniform_cube(8_000, 1_000, dim=5, seed=24),
}


def _exact(gi, gd, oi, od):
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32)), "dists2 differ"
    assert np.array_equal(gi, oi), "indices differ"


@pytest.mark.parametrize("name", list(KD_CASES))
@pytest.mark.parametrize("k,eps,flags,radius", [
    (16, 0.0, 0, np.inf), (16, 0.0, SORT, np.inf), (5, 0.1, 0, np.inf), (1, 0.0, 0, np.inf),
    (40, 0.0, SORT, np.inf), (10, 0.0, 0, 0.6), (7, 0.0, SELF, np.inf),
    (80, 0.0, SORT, np.inf),
    (128, 0.0, 0, np.inf),  # k above 96: the heap is kept in the output rows
])
def test_tree_heap_matches_oracle(nabo, oracle, name, k, eps, flags, radius):
    cloud, query = KD_CASES[name]()
    print(f"case kd stl {name} k={k} eps={eps} flags={flags} r={radius}", flush=True)
    nns = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_TREE_HEAP, cloud.shape[1],
                                      creationOptionFlags=nabo.CreationOptionFlags.TOUCH_STATISTICS)
    gi, gd = nns.knn(query, k, eps, flags, radius)
    touched = nns.last_touched
    ot = oracle.kdtree(cloud)
    oi, od, otouch, _ = ot.knn(query, k, eps, flags, radius, heap="stl", threads=8, stats=True)
    _exact(gi, gd, oi, od)
    assert touched == int(otouch.sum())


def test_tree_heap_per_query_radii(nabo, oracle):
    cloud, query = KD_CASES["ties"]()
    radii = np.random.default_rng(5).uniform(0.3, 2.5, len(query)).astype(np.float32)
    print("case kd stl radii", flush=True)
    nns = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_TREE_HEAP)
    gi, gd = nns.knn(query, 9, 0.0, 0, maxRadii=radii)
    oi, od = oracle.kdtree(cloud).knn(query, 9, 0.0, 0, radii=radii, heap="stl", threads=8)
    _exact(gi, gd, oi, od)


def test_tree_heap_same_distances_as_linear(nabo):
    cloud, query = KD_CASES["ties"]()
    print("case kd stl vs linear", flush=True)
    stl = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_TREE_HEAP)
    lin = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_HIP)
    si, sd = stl.knn(query, 12, 0.0, SORT)
    li, ld = lin.knn(query, 12, 0.0, SORT)
    assert np.array_equal(sd.view(np.uint32), ld.view(np.uint32))
    # Below a row's largest distance both keep every entry; sort_heap is not
    # stable, so equal distances may come in another order.
    for r in np.nonzero((si != li).any(axis=1))[0]:
        for v in np.unique(sd[r][sd[r] < sd[r][-1]]):
            assert sorted(si[r][sd[r] == v]) == sorted(li[r][ld[r] == v]), f"query {r}"


BF_CASES = {
    "uniform3d": lambda: datasets.uniform_cube(20_000, 2_000, seed=31),
    "dirichlet33": lambda: datasets.dirichlet_descriptors(20_000, 700, seed=32),
    "ties": lambda: _ties(5_000, 1_500, 33),
}


@pytest.mark.parametrize("name", list(BF_CASES))
@pytest.mark.parametrize("k,flags,radius", [(8, 0, np.inf), (8, SORT, np.inf), (1, SELF, np.inf),
                                             (10, 0, 0.5), (33, SORT, np.inf), (64, 0, np.inf)])
def test_brute_force_stl_matches_oracle(nabo, oracle, name, k, flags, radius):
    cloud, query = BF_CASES[name]()
    print(f"case bf stl {name} k={k} flags={flags} r={radius}", flush=True)
    bf = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.BRUTE_FORCE)
    gi, gd = bf.knn(query, k, 0.0, flags, radius)
    oi, od = oracle.bf_knn(cloud, query, k, flags=flags, max_radius=radius, threads=8)
    _exact(gi, gd, oi, od)


def test_stl_heap_refuses_shortlist(nabo):
    cloud, _ = BF_CASES["uniform3d"]()
    with pytest.raises(nabo.NaboError) as e:
        nabo.NearestNeighbourSearch(cloud, nabo.SearchType.BRUTE_FORCE_HIP, 3,
                                    {"heap": "stl", "bfPrecision": "bf16"})
    assert e.value.code == 4


def test_brute_force_stl_log_overflow(nabo, oracle):
    # Points met in order of decreasing distance from the first queries:
    # every point enters a chunk's k-best, which overflows its admission log;
    # those queries are redone by a full scan, the others replayed.
    rng = np.random.default_rng(34)
    n = 5_000
    line = np.zeros((n, 3), np.float32)
    line[:, 0] = 1000.0 - 0.1 * np.arange(n, dtype=np.float32)
    cloud = np.concatenate([line, (rng.random((3_000, 3)) * 50).astype(np.float32)])
    query = np.concatenate([
        np.array([[-1.0, 0.0, 0.0], [-2.0, 0.01, 0.0]], np.float32),
        (rng.random((300, 3)) * 50).astype(np.float32)])
    print("case bf stl log overflow", flush=True)
    bf = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.BRUTE_FORCE)
    for k, flags in ((1, 0), (4, SORT), (8, 0)):
        gi, gd = bf.knn(query, k, 0.0, flags)
        oi, od = oracle.bf_knn(cloud, query, k, flags=flags, threads=8)
        _exact(gi, gd, oi, od)


@pytest.mark.parametrize("k,flags", [(1, 0), (5, SORT), (16, 0), (40, SORT), (64, 0), (96, SORT)])
def test_tree_heap_admission_log_refills(nabo, oracle, k, flags):
    # One bucket (n <= bucketSize) scanned in index order, with the points
    # met in order of decreasing distance from the first queries: every
    # point is admitted, so each lane's admission log (at most 256 entries)
    # fills and is replayed into the heap many times during the query, at
    # different moments for different lanes of a wave.
    rng = np.random.default_rng(35)
    n = 3_000
    cloud = np.zeros((n, 3), np.float32)
    cloud[:, 0] = 1000.0 - 0.1 * np.arange(n, dtype=np.float32)
    cloud[:, 1] = rng.random(n).astype(np.float32) * 1e-3
    query = np.concatenate([
        np.array([[-1.0, 0.0, 0.0], [-2.0, 0.01, 0.0], [-50.0, 3.0, 1.0]], np.float32),
        (rng.random((200, 3)) * [1000.0, 1.0, 1.0]).astype(np.float32)])
    print(f"case kd stl log refills k={k} flags={flags}", flush=True)
    nns = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_TREE_HEAP, 3, {"bucketSize": 4_000},
                                      creationOptionFlags=nabo.CreationOptionFlags.TOUCH_STATISTICS)
    gi, gd = nns.knn(query, k, 0.0, flags)
    oi, od, otouch, _ = oracle.kdtree(cloud, bucket_size=4_000).knn(
        query, k, 0.0, flags, heap="stl", threads=8, stats=True)
    _exact(gi, gd, oi, od)
    assert nns.last_touched == int(otouch.sum())
