"""GPU parity of batched kd-trees (nabo_hip_batch_*) against the oracle.

Bar: the tree of each cloud inside the batch is, node for node, the oracle's
tree of that cloud alone (child ids and bucket starts shifted by the cloud's
base), and each query set's indices and dists2 are bit-identical to the
oracle's search of that cloud.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_batch(rng, sizes, scale=True):
    clouds, queries = [], []
    for s in sizes:
        sc = rng.uniform(1, 50) if scale else 1.0
        c = (rng.random((s, 3)) * sc).astype(np.float32)
        q = (c + rng.normal(scale=0.01 * sc, size=c.shape)).astype(np.float32)
        clouds.append(c)
        queries.append(q)
    return clouds, queries


def offsets_of(parts):
    return np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.int64)


def build(nabo, clouds, bucket=8, flags=0):
    import torch

    pts = torch.from_numpy(np.concatenate(clouds)).cuda()
    return nabo.BatchSearch(pts, offsets_of(clouds), 3, flags, {"bucketSize": bucket}), pts


@pytest.mark.parametrize("bucket", [8, 3])
def test_batch_trees_identical(nabo, oracle, bucket):
    rng = np.random.default_rng(4)
    sizes = [2000, 9, 8, 1, 17, 5000, 37, 12000, 300, 2]
    clouds, _ = make_batch(rng, sizes)
    print(f"case batch trees bucket={bucket}", flush=True)
    b, _ = build(nabo, clouds, bucket)
    nodes, bucket_idx = b.tree()
    roots = b.roots()
    off = offsets_of(clouds)
    ends = list(roots[1:]) + [len(nodes)]
    lo, hi = b.bounds()
    for c, cl in enumerate(clouds):
        ot = oracle.kdtree(cl, bucket_size=bucket)
        onodes, obucket, omin, omax = ot.export()
        mine = nodes[roots[c]:ends[c]].copy()
        bits = ot.dim_bits
        leaf = (mine[:, 0] & ((1 << bits) - 1)) == 3
        mine[~leaf, 0] = (mine[~leaf, 0] & 3) | (((mine[~leaf, 0] >> bits) - roots[c]) << bits)
        mine[leaf, 1] = mine[leaf, 1] - np.uint32(off[c])
        assert np.array_equal(mine[:, 0], onodes[:, 0]), f"cloud {c} node words"
        assert np.array_equal(mine[leaf, 1], onodes[leaf, 1]), f"cloud {c} bucket starts"
        assert np.array_equal(mine[~leaf, 1].view(np.float32), onodes[~leaf, 1].view(np.float32))
        assert np.array_equal(bucket_idx[off[c]:off[c + 1]], obucket), f"cloud {c} buckets"
        assert np.array_equal(lo[c], omin) and np.array_equal(hi[c], omax), f"cloud {c} bounds"


@pytest.mark.parametrize("k,eps,flags", [(8, 0.0, 0), (1, 0.0, 1), (5, 0.1, 2), (8, 0.0, 3)])
def test_batch_knn_bit_exact(nabo, oracle, k, eps, flags):
    import torch

    rng = np.random.default_rng(5)
    sizes = list(np.round(np.exp(rng.uniform(np.log(20), np.log(20000), size=60))).astype(int))
    clouds, queries = make_batch(rng, sizes)
    print(f"case batch knn k={k} eps={eps} flags={flags}", flush=True)
    b, _ = build(nabo, clouds, 8, nabo.CreationOptionFlags.TOUCH_STATISTICS)
    dq = torch.from_numpy(np.concatenate(queries)).cuda()
    qoff = torch.from_numpy(offsets_of(queries)).cuda()
    touched = torch.zeros(1, dtype=torch.int64, device="cuda")
    gi, gd = b.knn(dq, qoff, k, eps, flags, 2.0, touched=touched)
    gi, gd = gi.cpu().numpy(), gd.cpu().numpy()
    qo = offsets_of(queries)
    total = 0
    for c, (cl, q) in enumerate(zip(clouds, queries)):
        oi, od, t, _ = oracle.kdtree(cl, bucket_size=8).knn(q, k, eps, flags, 2.0, stats=True)
        total += int(t.sum())
        assert np.array_equal(gi[qo[c]:qo[c + 1]], oi), f"cloud {c} indices"
        assert np.array_equal(gd[qo[c]:qo[c + 1]].view(np.uint32), od.view(np.uint32)), c
    assert int(touched.item()) == total


def test_batch_rejects_too_large_k(nabo):
    import torch

    rng = np.random.default_rng(6)
    clouds, queries = make_batch(rng, [100, 4, 100])
    b, _ = build(nabo, clouds)
    dq = torch.from_numpy(np.concatenate(queries)).cuda()
    qoff = torch.from_numpy(offsets_of(queries)).cuda()
    with pytest.raises(nabo.NaboError, match="minus 1"):
        b.knn(dq, qoff, 4)
    b.knn(dq, qoff, 4, 0.0, 1)


@pytest.mark.parametrize("k,flags", [(6, 0), (12, 2)])
def test_batch_knn_tree_heap(nabo, oracle, k, flags):
    # heap "stl": every cloud's results in IndexHeapSTL order (KDTREE_TREE_HEAP)
    import torch

    rng = np.random.default_rng(7)
    sizes = list(np.round(np.exp(rng.uniform(np.log(20), np.log(20000), size=40))).astype(int))
    clouds, queries = make_batch(rng, sizes)
    print(f"case batch knn tree heap k={k} flags={flags}", flush=True)
    pts = torch.from_numpy(np.concatenate(clouds)).cuda()
    b = nabo.BatchSearch(pts, offsets_of(clouds), 3, nabo.CreationOptionFlags.TOUCH_STATISTICS,
                         {"bucketSize": 8, "heap": "stl"})
    dq = torch.from_numpy(np.concatenate(queries)).cuda()
    qoff = torch.from_numpy(offsets_of(queries)).cuda()
    touched = torch.zeros(1, dtype=torch.int64, device="cuda")
    gi, gd = b.knn(dq, qoff, k, 0.0, flags, 2.0, touched=touched)
    gi, gd = gi.cpu().numpy(), gd.cpu().numpy()
    qo = offsets_of(queries)
    total = 0
    for c, (cl, q) in enumerate(zip(clouds, queries)):
        oi, od, t, _ = oracle.kdtree(cl, bucket_size=8).knn(q, k, 0.0, flags, 2.0, heap="stl",
                                                            stats=True)
        total += int(t.sum())
        assert np.array_equal(gi[qo[c]:qo[c + 1]], oi), f"cloud {c} indices"
        assert np.array_equal(gd[qo[c]:qo[c + 1]].view(np.uint32), od.view(np.uint32)), c
    assert int(touched.item()) == total
