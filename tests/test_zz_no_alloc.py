"""nabo_hip_knn_device_noalloc: knn with no device allocation.

Called through the C ABI (ctypes) and compared, bit for bit, with
nabo_hip_knn_device on the same index and inputs, on two calls in a row of
which the second is smaller (the static scratch keeps the first call's
contents), and with the oracle. Calls it refuses (brute-force indices,
IndexHeapSTL with k <= 96, more temporaries than the array holds) must
return NABO_HIP_ERR_UNSUPPORTED without a search.
"""
import ctypes as C

import numpy as np
import pytest

from libnabo_amd import datasets
from tests.helpers import assert_bit_exact

pytestmark = pytest.mark.gpu

UNSUPPORTED = 4


def _run(nabo, fn_name, nns, dq, k, eps=0.0, flags=0, radius=np.inf, touched=False):
    import torch

    lib = nabo.load_library()
    m = dq.shape[0]
    di = torch.full((m, k), -7, dtype=torch.int32, device="cuda")
    dd = torch.full((m, k), -7.0, dtype=torch.float32, device="cuda")
    dt = torch.zeros(1, dtype=torch.int64, device="cuda") if touched else None
    stream = torch.cuda.current_stream().cuda_stream
    rc = getattr(lib, fn_name)(nns._h, C.c_void_p(dq.data_ptr()), 3, dq.stride(0), m, k,
                              float(eps), int(flags), float(radius), None,
                              C.c_void_p(di.data_ptr()), C.c_void_p(dd.data_ptr()),
                              None if dt is None else C.c_void_p(dt.data_ptr()),
                              C.c_void_p(stream))
    torch.cuda.synchronize()
    msg = lib.nabo_hip_last_error().decode()
    return rc, msg, di.cpu().numpy(), dd.cpu().numpy(), (None if dt is None else int(dt.item()))


def test_noalloc_matches_knn_device_two_calls(nabo, oracle):
    import torch

    cloud, query = datasets.lidar_scene(200_000, seed=41)
    nns = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_HIP, 3, {"bucketSize": 8},
                                      creationOptionFlags=nabo.CreationOptionFlags.TOUCH_STATISTICS)
    tree = oracle.kdtree(cloud, bucket_size=8)
    # first call: all 200k queries, k = 16, maxRadius 1; second: a smaller
    # call (50k queries, k = 8, epsilon 0.1) on the same scratch
    for m, k, eps, radius in ((200_000, 16, 0.0, 1.0), (50_000, 8, 0.1, np.inf)):
        print(f"case noalloc m={m} k={k} eps={eps} r={radius}", flush=True)
        dq = torch.from_numpy(np.ascontiguousarray(query[:m])).cuda()
        rc, msg, ni, nd, nt = _run(nabo, "nabo_hip_knn_device_noalloc", nns, dq, k, eps, 0,
                                   radius, touched=True)
        assert rc == 0, msg
        rc, msg, gi, gd, gt = _run(nabo, "nabo_hip_knn_device", nns, dq, k, eps, 0, radius,
                                   touched=True)
        assert rc == 0, msg
        assert_bit_exact(ni, nd, gi, gd, f"noalloc vs knn_device m={m} k={k}")
        assert nt == gt
        oi, od, ot, _ = tree.knn(query[:m], k, eps, 0, radius, stats=True)
        assert_bit_exact(ni, nd, oi, od, f"noalloc vs oracle m={m} k={k}")
        assert nt == int(ot.sum())
    nns.close()


def test_noalloc_tree_heap_large_k(nabo):
    import torch

    cloud, query = datasets.uniform_cube(20_000, 3_000, seed=42)
    nns = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_TREE_HEAP, 3)
    for m, k in ((3_000, 100), (1_000, 97)):
        print(f"case noalloc tree heap m={m} k={k}", flush=True)
        dq = torch.from_numpy(np.ascontiguousarray(query[:m])).cuda()
        rc, msg, ni, nd, _ = _run(nabo, "nabo_hip_knn_device_noalloc", nns, dq, k)
        assert rc == 0, msg
        rc, msg, gi, gd, _ = _run(nabo, "nabo_hip_knn_device", nns, dq, k)
        assert rc == 0, msg
        assert_bit_exact(ni, nd, gi, gd, f"noalloc tree heap k={k}")
    nns.close()


def test_noalloc_refusals(nabo):
    import torch

    cloud, query = datasets.uniform_cube(5_000, 1_000, seed=43)
    dq = torch.from_numpy(query).cuda()
    bf = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.BRUTE_FORCE_HIP, 3)
    rc, msg, ni, _, _ = _run(nabo, "nabo_hip_knn_device_noalloc", bf, dq, 4)
    assert rc == UNSUPPORTED and "kd-tree" in msg
    assert (ni == -7).all(), "a refused call launches no search"
    stl = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_TREE_HEAP, 3)
    rc, msg, ni, _, _ = _run(nabo, "nabo_hip_knn_device_noalloc", stl, dq, 16)
    assert rc == UNSUPPORTED and "admission logs" in msg
    assert (ni == -7).all()
    # more temporaries than the static array holds (about 24 bytes a query)
    kd = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_HIP, 3)
    big = torch.rand((4_000_000, 3), device="cuda")
    rc, msg, ni, _, _ = _run(nabo, "nabo_hip_knn_device_noalloc", kd, big, 1)
    assert rc == UNSUPPORTED and "bytes of scratch" in msg, msg
    assert (ni == -7).all()
    # the entry point still works after the refusals
    rc, msg, ni, nd, _ = _run(nabo, "nabo_hip_knn_device_noalloc", kd, dq, 4)
    assert rc == 0, msg
    gi, gd = kd.knn(query, 4)
    assert_bit_exact(ni, nd, gi, gd, "after refusals")
    for x in (bf, stl, kd):
        x.close()
