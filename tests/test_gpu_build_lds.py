"""GPU parity of the kd-tree build where segments are finished in LDS.

Segments of at most T points (T = 2048 for 2-D and 3-D clouds, and
228 * (bucketSize + 1) below bucketSize 8) leave the build's level loop and
one workgroup finishes each subtree in LDS. Bar, as for every build: the
tree is the oracle's node for node (node words, cut values, bucket order,
bounds, depth) and knn results on it are bit-identical to the oracle's.
The cases sit at the edges of that path: clouds just under, at and over T;
bucket sizes whose T is smaller; 1-D and 2-D clouds; a 4-D cloud (no LDS
finish); clustered, duplicated and identical points, whose subtrees are
deep or split in the equal-to-cut block; batches mixing those sizes.
"""
import numpy as np
import pytest

from tests.helpers import assert_bit_exact
from tests.test_gpu_batch import build as build_batch
from tests.test_gpu_batch import offsets_of
from tests.test_gpu_kdtree import assert_same_tree, make

pytestmark = pytest.mark.gpu


def _check(nabo, oracle, cloud, bucket, k=8, queries=None):
    nns = make(nabo, cloud, bucket)
    otree = oracle.kdtree(cloud, bucket_size=bucket)
    assert_same_tree(nns, otree)
    q = cloud[:2000] if queries is None else queries
    gi, gd = nns.knn(q, k, 0.0, 1)
    oi, od = otree.knn(q, k, flags=1)
    assert_bit_exact(gi, gd, oi, od, "knn on the LDS-finished tree")


@pytest.mark.parametrize("n", [2047, 2048, 2049, 4097, 30_000])
def test_sizes_around_the_lds_limit(nabo, oracle, n):
    rng = np.random.default_rng(n)
    cloud = rng.random((n, 3)).astype(np.float32)
    print(f"case lds n={n}", flush=True)
    _check(nabo, oracle, cloud, 8)


@pytest.mark.parametrize("bucket,n", [(2, 684), (2, 685), (2, 9000), (3, 5000), (17, 20_000)])
def test_bucket_sizes(nabo, oracle, bucket, n):
    rng = np.random.default_rng(bucket * 100 + n)
    cloud = rng.random((n, 3)).astype(np.float32)
    print(f"case lds bucket={bucket} n={n}", flush=True)
    _check(nabo, oracle, cloud, bucket)


@pytest.mark.parametrize("dim", [1, 2, 4])
def test_dimensions(nabo, oracle, dim):
    rng = np.random.default_rng(dim)
    cloud = rng.random((12_000, dim)).astype(np.float32)
    print(f"case lds dim={dim}", flush=True)
    _check(nabo, oracle, cloud, 8)


def test_clustered_points(nabo, oracle):
    # exponential coordinates: the sliding midpoint keeps slicing off a few
    # points, so the LDS subtrees are deep
    rng = np.random.default_rng(7)
    cloud = rng.exponential(scale=1.0, size=(40_000, 3)).astype(np.float32) ** 3
    print("case lds clustered", flush=True)
    _check(nabo, oracle, cloud, 8)


def test_duplicated_points(nabo, oracle):
    # a 4 x 4 x 4 grid: every cut is clamped onto a coordinate shared by many
    # points (the equal-to-cut block of the second Hoare pass)
    rng = np.random.default_rng(8)
    cloud = rng.integers(0, 4, size=(20_000, 3)).astype(np.float32)
    print("case lds duplicates", flush=True)
    _check(nabo, oracle, cloud, 8, k=4)


def test_identical_points(nabo, oracle):
    cloud = np.full((3000, 3), 0.25, np.float32)
    print("case lds identical", flush=True)
    _check(nabo, oracle, cloud, 8, k=3)


def test_batch_of_sizes_around_the_limit(nabo, oracle):
    rng = np.random.default_rng(9)
    sizes = [2048, 2049, 9, 2047, 700, 4100, 5, 30_000, 4]  # k = 4 needs 4 points
    clouds = [(rng.random((s, 3)) * rng.uniform(1, 20)).astype(np.float32) for s in sizes]
    print("case lds batch", flush=True)
    b, pts = build_batch(nabo, clouds, 8)
    import torch

    off = offsets_of(clouds)
    q = pts.clone()
    dqoff = torch.from_numpy(off).cuda()
    gi, gd = b.knn(q, dqoff, 4, optionFlags=1)
    gi, gd = gi.cpu().numpy(), gd.cpu().numpy()
    for c, cl in enumerate(clouds):
        oi, od = oracle.kdtree(cl, bucket_size=8).knn(cl, 4, flags=1)
        assert_bit_exact(gi[off[c]:off[c + 1]], gd[off[c]:off[c + 1]], oi, od, f"cloud {c}")
    b.close()
