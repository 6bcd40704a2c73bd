"""GPU brute force (BRUTE_FORCE_HIP) against the oracle.

FP32 path: dists2 bit-identical to the oracle's brute force (same sequential
sum), indices identical except among equal distances (the reference's STL
heap orders ties by its own heap moves; this path orders them by index).
bf16 / fp16 shortlist paths: every returned distance is the exact fp32
distance of the returned point (rescoring), results are sorted, and the
fraction of queries whose k-best differs from the exact one (the shortlist
miss rate) stays under the bound written in each test.
"""
import numpy as np
import pytest

from libnabo_amd import datasets

pytestmark = pytest.mark.gpu


def same_up_to_ties(gi, gd, oi, od):
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32)), "dists2 differ"
    for r in np.nonzero((gi != oi).any(axis=1))[0]:
        for v in np.unique(gd[r]):
            sel = gd[r] == v
            if v == gd[r][-1]:
                continue  # boundary ties: any equal-distance point is valid
            assert sorted(gi[r][sel]) == sorted(oi[r][sel]), f"query {r}"


CASES = {
    "uniform3d": lambda: datasets.uniform_cube(20_000, 2_000, seed=3),
    "dirichlet33": lambda: datasets.dirichlet_descriptors(20_000, 1_000, seed=3),
    "scan3d": lambda: (datasets.load_reference_scan("scan.3d"),
                       datasets.load_reference_scan("scan.3d")[::3]),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("k,flags,radius", [(8, 0, np.inf), (1, 1, np.inf), (10, 2, 0.5)])
def test_bf_fp32_matches_oracle(nabo, oracle, name, k, flags, radius):
    cloud, query = CASES[name]()
    print(f"case bf fp32 {name} k={k} flags={flags} r={radius}", flush=True)
    bf = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.BRUTE_FORCE_HIP)
    gi, gd = bf.knn(query, k, 0.0, flags, radius)
    oi, od = oracle.bf_knn(cloud, query, k, flags=flags | 2, max_radius=radius, threads=8)
    same_up_to_ties(gi, gd, oi, od)


def test_bf_touch_statistics(nabo):
    cloud, query = CASES["uniform3d"]()
    bf = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.BRUTE_FORCE_HIP, 3,
                                     creationOptionFlags=nabo.CreationOptionFlags.TOUCH_STATISTICS)
    bf.knn(query[:100], 4)
    assert bf.last_touched == 100 * len(cloud)


@pytest.mark.parametrize("prec,extra,max_miss", [("fp16", 32, 0.02), ("bf16", 32, 0.10),
                                                  ("fp16", 8, 0.10)])
def test_bf_shortlist(nabo, oracle, prec, extra, max_miss):
    cloud, query = CASES["dirichlet33"]()
    k = 8
    print(f"case bf {prec} m={extra}", flush=True)
    bf = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.BRUTE_FORCE_HIP, 33,
                                     {"bfPrecision": prec, "bfExtra": extra})
    gi, gd = bf.knn(query, k)
    exact = oracle.dist2_f64(cloud, query, gi)
    # every returned distance is the fp32 sequential sum of that pair
    oi, od = oracle.bf_knn(cloud, query, k, flags=2, threads=8)
    for r in range(len(query)):
        for j in range(k):
            p = cloud[gi[r, j]]
            d = np.float32(0)
            for c in range(33):
                diff = np.float32(p[c] - query[r, c])
                d = np.float32(d + np.float32(diff * diff))
            assert d == gd[r, j]
        if r > 60:
            break
    assert np.all(np.diff(gd, axis=1) >= 0)
    np.testing.assert_allclose(gd, exact, rtol=1e-5)
    miss = np.mean(np.any(gd.view(np.uint32) != od.view(np.uint32), axis=1))
    print(f"{prec} m={extra}: shortlist miss rate {miss:.4f}")
    assert miss <= max_miss


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
@pytest.mark.parametrize("k,extra,flags,radius", [(20, 24, 2, np.inf), (8, 40, 0, 0.6),
                                                  (30, 20, 3, np.inf), (5, 11, 2, 0.4)])
def test_bf_shortlist_all_points_exact(nabo, oracle, prec, k, extra, flags, radius):
    """A cloud no larger than the shortlist (n <= k + bfExtra): every point is
    a candidate, so the rescored result is the exact brute force's, ties and
    admission rules (radius, self match) included. Covers shortlists of 16,
    32 and 64 lanes per query and k above 16."""
    rng = np.random.default_rng(7)
    n = min(k + extra, 44)
    cloud = rng.random((n, 5), dtype=np.float32)
    query = np.concatenate([cloud[::2], rng.random((300, 5), dtype=np.float32)])
    print(f"case bf {prec} all points k={k} m={extra} flags={flags} r={radius}", flush=True)
    bf = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.BRUTE_FORCE_HIP, 5,
                                     {"bfPrecision": prec, "bfExtra": extra})
    gi, gd = bf.knn(query, k, 0.0, flags, radius)
    oi, od = oracle.bf_knn(cloud, query, k, flags=flags | 2, max_radius=radius)
    same_up_to_ties(gi, gd, oi, od)
    assert np.array_equal(gi == -1, oi == -1)


@pytest.mark.parametrize("prec,k,extra,max_miss", [("fp16", 40, 8, 0.10), ("bf16", 24, 16, 0.10)])
def test_bf_shortlist_large_k(nabo, oracle, prec, k, extra, max_miss):
    """k above 16 on 33-D descriptors: exact distances, sorted rows, and a
    miss rate against the exact fp32 k-best under the bound written here."""
    cloud, query = CASES["dirichlet33"]()
    print(f"case bf {prec} k={k} m={extra}", flush=True)
    bf = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.BRUTE_FORCE_HIP, 33,
                                     {"bfPrecision": prec, "bfExtra": extra})
    gi, gd = bf.knn(query, k)
    exact = oracle.dist2_f64(cloud, query, gi)
    np.testing.assert_allclose(gd, exact, rtol=1e-5)
    assert np.all(np.diff(gd, axis=1) >= 0)
    assert np.all(gi >= 0)
    assert all(len(set(r)) == k for r in gi)
    oi, od = oracle.bf_knn(cloud, query, k, flags=2, threads=8)
    miss = np.mean(np.any(gd.view(np.uint32) != od.view(np.uint32), axis=1))
    print(f"{prec} k={k} m={extra}: shortlist miss rate {miss:.4f}")
    assert miss <= max_miss
