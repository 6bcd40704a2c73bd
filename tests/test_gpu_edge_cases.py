"""Edge cases of the kd-tree search against the oracle, bit for bit.

Queries outside the cloud's box (the root starts at rd = 0 with implicit
bounds, kdtree_cpu.cpp:345-348), non-finite queries (NaN and infinite
distances are never admitted, :392-394, and never open a far child,
:415-417), a zero radius (dist <= 0 admits exact duplicates only), a cloud
of one repeated point, k equal to the cloud size, empty or one-query calls,
a tree too deep for stacks in LDS, and clouds and queries with column
strides above their row counts (host and device buffers). Every case compares indices, dists2 and the touched count with the
oracle on the same tree.
"""
import numpy as np
import pytest

from libnabo_amd import datasets

pytestmark = pytest.mark.gpu

SELF, SORT = 1, 2


def check(nabo, oracle, cloud, query, k, eps=0.0, flags=0, radius=np.inf, bucket=8, heap="linear"):
    st = nabo.SearchType.KDTREE_HIP if heap == "linear" else nabo.SearchType.KDTREE_TREE_HEAP
    nns = nabo.NearestNeighbourSearch(cloud, st, cloud.shape[1], {"bucketSize": bucket},
                                      creationOptionFlags=nabo.CreationOptionFlags.TOUCH_STATISTICS)
    gi, gd = nns.knn(query, k, eps, flags, radius)
    oi, od, touched, _ = oracle.kdtree(cloud, bucket_size=bucket).knn(
        query, k, eps, flags, radius, heap=heap, threads=8, stats=True)
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32)), "dists2 differ"
    assert np.array_equal(gi, oi), "indices differ"
    assert nns.last_touched == int(touched.sum())
    return gi, gd


def test_queries_outside_the_box(nabo, oracle):
    cloud, query = datasets.uniform_cube(20_000, 2_000, seed=41)
    rng = np.random.default_rng(41)
    far = query + rng.choice([-1.0, 1.0], query.shape).astype(np.float32) * \
        rng.uniform(0.5, 50.0, query.shape).astype(np.float32)
    print("case outside the box", flush=True)
    for heap in ("linear", "stl"):
        check(nabo, oracle, cloud, far, 8, heap=heap)
        check(nabo, oracle, cloud, far, 8, radius=2.0, heap=heap)


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf, 1e30])
def test_non_finite_queries(nabo, oracle, bad):
    cloud, query = datasets.uniform_cube(8_000, 1_000, seed=42)
    query = query.copy()
    query[::3, 0] = bad
    query[1::7, 2] = bad
    print(f"case non-finite query {bad}", flush=True)
    gi, gd = check(nabo, oracle, cloud, query, 6)
    hit = np.zeros(len(query), bool)
    hit[::3] = True
    hit[1::7] = True
    # every coordinate difference with such a query is NaN or overflows
    assert (gi[hit] == -1).all() and np.isinf(gd[hit]).all()
    assert (gi[~hit] >= 0).all()
    check(nabo, oracle, cloud, query, 6, eps=0.5, heap="stl")


def test_zero_radius_admits_exact_duplicates(nabo, oracle):
    cloud, _ = datasets.uniform_cube(5_000, 1, seed=43)
    cloud = np.concatenate([cloud, cloud[:500]])  # 500 exact duplicates
    print("case zero radius", flush=True)
    gi, gd = check(nabo, oracle, cloud, cloud, 3, flags=SELF, radius=0.0)
    assert (gd[:, 0] == 0).all()
    assert ((gi[:500, :2] >= 0).all() and (gi[500:5000, 1] == -1).all())
    check(nabo, oracle, cloud, cloud, 3, flags=0, radius=0.0)  # self match excluded: nothing


def test_one_repeated_point(nabo, oracle):
    cloud = np.tile(np.array([[1.5, -2.0, 0.25]], np.float32), (3_000, 1))
    query = np.array([[1.5, -2.0, 0.25], [0.0, 0.0, 0.0], [9.0, 9.0, 9.0]], np.float32)
    print("case one repeated point", flush=True)
    check(nabo, oracle, cloud, query, 10, flags=SELF)
    check(nabo, oracle, cloud, query, 10, flags=SELF | SORT, heap="stl")


@pytest.mark.parametrize("n", [9, 64, 128])
def test_k_equals_cloud_size(nabo, oracle, n):
    cloud, query = datasets.uniform_cube(n, 200, seed=44)
    print(f"case k = n = {n}", flush=True)
    gi, gd = check(nabo, oracle, cloud, query, n, flags=SELF)
    assert (np.sort(gi, axis=1) == np.arange(n)).all()
    check(nabo, oracle, cloud, query, n, flags=SELF | SORT, heap="stl")


def test_empty_and_single_query(nabo, oracle):
    cloud, query = datasets.uniform_cube(4_000, 5, seed=45)
    nns = nabo.NearestNeighbourSearch(cloud)
    print("case empty / single query", flush=True)
    gi, gd = nns.knn(np.empty((0, 3), np.float32), 4)
    assert gi.shape == (0, 4) and gd.shape == (0, 4)
    check(nabo, oracle, cloud, query[:1], 4)


@pytest.mark.parametrize("k", [1, 8, 40, 96])
def test_deep_tree_stacks_in_hbm(nabo, oracle, k):
    # Points at x = 2^-i split off one per level (midpoint cuts of the
    # widest side): a tree of depth 89, beyond the 32 levels whose stacks fit
    # in LDS, so the search keeps its stacks in HBM; the IndexHeapSTL form
    # then holds its heap in LDS slots of its own.
    rng = np.random.default_rng(46)
    e = np.zeros((90, 3), np.float32)
    e[:, 0] = np.float32(2.0) ** (-np.arange(90, dtype=np.float32))
    e[:, 1] = rng.random(90).astype(np.float32) * np.float32(1e-30)
    cloud = np.concatenate([e, (2 + rng.random((3_000, 3))).astype(np.float32)])
    query = np.concatenate([e[::3] * np.float32(1.5), (rng.random((300, 3)) * 3).astype(np.float32)])
    print(f"case deep tree k={k}", flush=True)
    for heap in ("linear", "stl"):
        check(nabo, oracle, cloud, query, k, bucket=2, heap=heap)
        check(nabo, oracle, cloud, query, k, flags=SORT, radius=1.0, bucket=2, heap=heap)
    nns = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_HIP, 3, {"bucketSize": 2})
    assert nns.info().depth == oracle.kdtree(cloud, bucket_size=2).depth > 32


def test_unread_hip_error_of_another_call_is_not_reported(nabo, oracle):
    # A HIP error that an earlier call on the thread left unread (here an
    # invalid device ordinal, which changes nothing) is not the library's:
    # builds and searches still succeed.
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library is linked against
    cloud, query = datasets.uniform_cube(3_000, 200, seed=47)
    print("case unread HIP error", flush=True)
    assert hip.hipSetDevice(ctypes.c_int(1 << 20)) != 0
    nns = nabo.NearestNeighbourSearch(cloud)
    assert hip.hipSetDevice(ctypes.c_int(1 << 20)) != 0
    gi, gd = nns.knn(query, 4)
    oi, od = oracle.kdtree(cloud).knn(query, 4, threads=8)
    assert np.array_equal(gi, oi) and np.array_equal(gd.view(np.uint32), od.view(np.uint32))


@pytest.mark.parametrize("stype", ["kd", "bf"])
def test_strided_cloud_and_queries(nabo, oracle, stype):
    # Eigen-style outer strides through the C ABI: a cloud whose points are 5
    # floats apart, host queries 4 floats apart, and a strided device view;
    # the extra floats hold values that must not be read.
    import ctypes

    import torch

    from tests.helpers import assert_bit_exact

    rng = np.random.default_rng(91)
    n, m, k = 20_000, 3_000, 8
    cloud = rng.random((n, 3)).astype(np.float32)
    query = rng.random((m, 3)).astype(np.float32)
    c5 = np.full((n, 5), 1e6, np.float32)
    c5[:, :3] = cloud
    q4 = np.full((m, 4), -1e6, np.float32)
    q4[:, :3] = query
    print(f"case strided {stype}", flush=True)
    lib = nabo.load_library()
    # brute force with libnabo's IndexHeapSTL (BRUTE_FORCE) is BRUTE_FORCE_HIP + the STL heap
    p = nabo.make_params(heap="linear" if stype == "kd" else "stl")
    st = nabo.SearchType.KDTREE_HIP if stype == "kd" else nabo.SearchType.BRUTE_FORCE_HIP
    h = ctypes.c_void_p()
    assert lib.nabo_hip_create(c5.ctypes.data_as(ctypes.c_void_p), n, 3, 5, 3, int(st), 0,
                               ctypes.byref(p), ctypes.byref(h)) == 0, lib.nabo_hip_last_error()
    try:
        gi = np.empty((m, k), np.int32)
        gd = np.empty((m, k), np.float32)
        assert lib.nabo_hip_knn(h, q4.ctypes.data_as(ctypes.c_void_p), 3, 4, m, k, 0.0, 2,
                                float("inf"), None, gi.ctypes.data_as(ctypes.c_void_p),
                                gd.ctypes.data_as(ctypes.c_void_p), None) == 0, \
            lib.nabo_hip_last_error()
        if stype == "kd":
            oi, od = oracle.kdtree(cloud, bucket_size=8).knn(query, k, flags=2, threads=8)
        else:
            oi, od = oracle.bf_knn(cloud, query, k, flags=2, threads=8)
        assert_bit_exact(gi, gd, oi, od, "strided host queries")
        dq = torch.from_numpy(q4).cuda()[:, :3]  # stride(0) == 4
        di = torch.empty((m, k), dtype=torch.int32, device="cuda")
        dd = torch.empty((m, k), dtype=torch.float32, device="cuda")
        assert lib.nabo_hip_knn_device(h, ctypes.c_void_p(dq.data_ptr()), 3, 4, m, k, 0.0, 2,
                                       float("inf"), None, ctypes.c_void_p(di.data_ptr()),
                                       ctypes.c_void_p(dd.data_ptr()), None, None) == 0, \
            lib.nabo_hip_last_error()
        torch.cuda.synchronize()
        assert_bit_exact(di.cpu().numpy(), dd.cpu().numpy(), oi, od, "strided device queries")
    finally:
        lib.nabo_hip_destroy(h)
