"""k and dim beyond the register-resident and tiled kernels' sizes.

The reference bounds k only by the cloud size (nabo/nabo.cpp:105-116) and
recommends KDTREE_TREE_HEAP for large k (nabo/nabo.h:293-294). The kd-tree
keeps its k-best in the query's output row above k = 128 (linear heap) and
96 (IndexHeapSTL), and its per-query offsets in HBM above 64 dimensions; the
brute force scans in index order with the k-best in the output row above
k = 64 or dim = 128. Each is checked against the oracle.
"""
import numpy as np
import pytest

from libnabo_amd import datasets
from tests.helpers import assert_bit_exact, assert_same_up_to_ties

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [129, 200, 333])
@pytest.mark.parametrize("heap", ["linear", "stl"])
def test_kdtree_large_k(nabo, oracle, k, heap):
    cloud, query = datasets.uniform_cube(6_000, 700, seed=51)
    st = nabo.SearchType.KDTREE_HIP if heap == "linear" else nabo.SearchType.KDTREE_TREE_HEAP
    nns = nabo.NearestNeighbourSearch(cloud, st, 3, {"bucketSize": 8},
                                      creationOptionFlags=nabo.CreationOptionFlags.TOUCH_STATISTICS)
    otree = oracle.kdtree(cloud, bucket_size=8)
    print(f"case kd large k={k} heap={heap}", flush=True)
    for flags, radius, eps in ((0, np.inf, 0.0), (2, 0.2, 0.0), (0, np.inf, 0.1)):
        gi, gd = nns.knn(query, k, eps, flags, radius)
        oi, od, touched, _ = otree.knn(query, k, eps, flags, radius, heap=heap, stats=True)
        assert_bit_exact(gi, gd, oi, od, f"k={k} {heap} flags={flags} r={radius} eps={eps}")
        assert nns.last_touched == int(touched.sum())


def test_kdtree_large_k_lidar_radius(nabo, oracle):
    cloud, query = datasets.lidar_scene(50_000, seed=52)
    query = query[:3_000]
    nns = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_HIP, 3, {"bucketSize": 8})
    print("case kd k=200 lidar r=1", flush=True)
    gi, gd = nns.knn(query, 200, 0.0, 0, 1.0)
    oi, od = oracle.kdtree(cloud, bucket_size=8).knn(query, 200, 0.0, 0, 1.0)
    assert_bit_exact(gi, gd, oi, od, "lidar k=200")


@pytest.mark.parametrize("k", [65, 200])
def test_brute_force_large_k(nabo, oracle, k):
    cloud, query = datasets.dirichlet_descriptors(5_000, 300, seed=53)
    print(f"case bf large k={k}", flush=True)
    stl = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.BRUTE_FORCE, 33)
    gi, gd = stl.knn(query, k, 0.0, 0)
    oi, od = oracle.bf_knn(cloud, query, k, 0)
    assert_bit_exact(gi, gd, oi, od, f"BRUTE_FORCE k={k} (IndexHeapSTL order)")
    gi, gd = stl.knn(query, k, 0.0, 2)
    oi, od = oracle.bf_knn(cloud, query, k, 2)
    assert_bit_exact(gi, gd, oi, od, f"BRUTE_FORCE k={k} sorted")
    lin = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.BRUTE_FORCE_HIP, 33)
    gi, gd = lin.knn(query, k, 0.0, 0)
    assert_same_up_to_ties(gi, gd, oi, od, f"BRUTE_FORCE_HIP k={k}")
    # a shortlist longer than 64 takes the exact path
    bf16 = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.BRUTE_FORCE_HIP, 33,
                                       {"bfPrecision": "bf16", "bfExtra": 8})
    gi, gd = bf16.knn(query, k, 0.0, 0)
    assert_same_up_to_ties(gi, gd, oi, od, f"bf16 shortlist k={k}")


def test_brute_force_high_dim(nabo, oracle):
    rng = np.random.default_rng(54)
    cloud = rng.random((3_000, 150)).astype(np.float32)
    query = rng.random((200, 150)).astype(np.float32)
    print("case bf dim 150", flush=True)
    for st, flags in ((nabo.SearchType.BRUTE_FORCE, 0), (nabo.SearchType.BRUTE_FORCE_HIP, 2)):
        nns = nabo.NearestNeighbourSearch(cloud, st, 150)
        gi, gd = nns.knn(query, 5, 0.0, flags)
        oi, od = oracle.bf_knn(cloud, query, 5, flags)
        if st == nabo.SearchType.BRUTE_FORCE:
            assert_bit_exact(gi, gd, oi, od, "dim 150 BRUTE_FORCE")
        else:
            assert_same_up_to_ties(gi, gd, oi, od, "dim 150 BRUTE_FORCE_HIP")


@pytest.mark.parametrize("heap", ["linear", "stl"])
def test_kdtree_high_dim(nabo, oracle, heap):
    """kd-tree above 64 dimensions (the search's offsets then live in HBM):
    tree and results bit-exact against the oracle."""
    rng = np.random.default_rng(55)
    dim = 100
    cloud = rng.random((4_000, dim)).astype(np.float32)
    query = np.concatenate([cloud[:50], rng.random((250, dim)).astype(np.float32)])
    st = nabo.SearchType.KDTREE_HIP if heap == "linear" else nabo.SearchType.KDTREE_TREE_HEAP
    nns = nabo.NearestNeighbourSearch(cloud, st, dim, {"bucketSize": 8},
                                      creationOptionFlags=nabo.CreationOptionFlags.TOUCH_STATISTICS)
    otree = oracle.kdtree(cloud, bucket_size=8)
    print(f"case kd dim {dim} heap={heap}", flush=True)
    for k, flags, radius, eps in ((5, 0, np.inf, 0.0), (12, 3, 1.4, 0.0), (5, 0, np.inf, 0.2)):
        gi, gd = nns.knn(query, k, eps, flags, radius)
        oi, od, touched, _ = otree.knn(query, k, eps, flags, radius, heap=heap, stats=True)
        assert_bit_exact(gi, gd, oi, od, f"dim {dim} {heap} k={k} flags={flags} r={radius}")
        assert nns.last_touched == int(touched.sum())
