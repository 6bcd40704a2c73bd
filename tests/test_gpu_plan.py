"""Batched build + search recorded as one HIP graph (nabo_hip_batch_plan_*).

Each launch rebuilds every tree from the point buffer and searches the query
buffer: results must equal the oracle's kd-tree on every cloud, bit for bit,
at every launch, including launches after the buffers' contents changed (new
data, same sizes) and launches with other GPU work between them. A level
budget too small for the data must be reported (complete = False), with no
search run.
"""
import numpy as np
import pytest

from tests.helpers import assert_bit_exact

pytestmark = pytest.mark.gpu


def _batch(seed, count=40, lo=50, hi=20_000, scale_hi=30.0, tiny=True):
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(int)
    if tiny:
        sizes[0] = 5  # a single-bucket tree (at most bucketSize points)
    clouds, queries = [], []
    for s in sizes:
        scale = rng.uniform(1.0, scale_hi)
        c = (rng.random((s, 3)) * scale).astype(np.float32)
        q = (c[rng.integers(0, s, max(1, s // 2))] +
             rng.normal(scale=0.01 * scale, size=(max(1, s // 2), 3))).astype(np.float32)
        clouds.append(c)
        queries.append(q)
    return clouds, queries


def _offsets(parts):
    return np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.int64)


def _check_against_oracle(oracle, clouds, queries, gi, gd, k, radius, what):
    qoff = _offsets(queries)
    for c, (cl, q) in enumerate(zip(clouds, queries)):
        oi, od = oracle.kdtree(cl, bucket_size=8).knn(q, k, 0.0, 0, radius)
        assert_bit_exact(gi[qoff[c]:qoff[c + 1]], gd[qoff[c]:qoff[c + 1]], oi, od,
                         f"{what}, cloud {c} ({len(cl)} points)")


def test_plan_matches_oracle_over_launches(nabo, oracle):
    import torch

    k, radius = 4, 2.0
    clouds, queries = _batch(31)
    pts = torch.from_numpy(np.concatenate(clouds)).cuda()
    qs = torch.from_numpy(np.concatenate(queries)).cuda()
    print(f"case plan {len(clouds)} clouds, {len(pts)} points, {len(qs)} queries", flush=True)
    plan = nabo.BatchPlan(pts, _offsets(clouds), qs, _offsets(queries), k, maxRadius=radius,
                          params={"bucketSize": 8})
    complete, levels = plan.status()
    assert not complete and 2 <= levels <= 33  # nothing launched yet
    for launch in range(3):
        plan.indices.fill_(-7)
        plan.dists2.fill_(-7.0)
        plan.launch()
        complete, _ = plan.status()
        assert complete
        gi, gd = plan.indices.cpu().numpy(), plan.dists2.cpu().numpy()
        _check_against_oracle(oracle, clouds, queries, gi, gd, k, radius, f"launch {launch}")
        torch.randn(4096, device="cuda").sum().item()  # other GPU work between launches
    # new contents of the same buffers: other clouds and queries of the same sizes
    rng = np.random.default_rng(32)
    clouds2 = [(rng.random(c.shape) * 7).astype(np.float32) for c in clouds]
    queries2 = [(cl[rng.integers(0, len(cl), len(q))] + rng.normal(scale=0.05, size=q.shape))
                .astype(np.float32) for cl, q in zip(clouds2, queries)]
    pts.copy_(torch.from_numpy(np.concatenate(clouds2)))
    qs.copy_(torch.from_numpy(np.concatenate(queries2)))
    plan.launch()
    complete, _ = plan.status()
    if complete:  # the new data may need more levels than the first
        gi, gd = plan.indices.cpu().numpy(), plan.dists2.cpu().numpy()
        _check_against_oracle(oracle, clouds2, queries2, gi, gd, k, radius, "new data")
    plan.close()


def test_plan_equals_batch_search(nabo):
    import torch

    k = 8
    clouds, queries = _batch(33, count=12, lo=2_000, hi=60_000, tiny=False)
    pts = torch.from_numpy(np.concatenate(clouds)).cuda()
    qs = torch.from_numpy(np.concatenate(queries)).cuda()
    off, qoff = _offsets(clouds), _offsets(queries)
    print("case plan vs BatchSearch", flush=True)
    plan = nabo.BatchPlan(pts, off, qs, qoff, k, params={"bucketSize": 16})
    plan.launch()
    assert plan.status()[0]
    b = nabo.BatchSearch(pts, off, 3, 0, {"bucketSize": 16})
    bi, bd = b.knn(qs, torch.from_numpy(qoff).cuda(), k)
    torch.cuda.synchronize()
    assert torch.equal(plan.indices, bi) and torch.equal(plan.dists2, bd)
    b.close()
    plan.close()


def test_plan_level_budget_too_small(nabo):
    import torch

    clouds, queries = _batch(34, count=6, lo=5_000, hi=20_000, tiny=False)
    pts = torch.from_numpy(np.concatenate(clouds)).cuda()
    qs = torch.from_numpy(np.concatenate(queries)).cuda()
    print("case plan with 3 levels", flush=True)
    plan = nabo.BatchPlan(pts, _offsets(clouds), qs, _offsets(queries), 4, levels=3)
    plan.indices.fill_(-7)
    plan.launch()
    complete, levels = plan.status()
    assert not complete and levels == 3
    assert bool((plan.indices == -7).all()), "no search runs on an unfinished build"
    plan.close()


def test_plan_argument_errors(nabo):
    import torch

    pts = torch.rand((100, 3), device="cuda")
    qs = torch.rand((10, 3), device="cuda")
    off = np.array([0, 100], np.int64)
    qoff = np.array([0, 10], np.int64)
    with pytest.raises(nabo.NaboError, match="levels"):
        nabo.BatchPlan(pts, off, qs, qoff, 2, levels=40)
    with pytest.raises(nabo.NaboError) as e:
        nabo.BatchPlan(pts, off, qs, qoff, 2, params={"heap": "stl"})
    assert e.value.code == 4
    with pytest.raises(nabo.NaboError, match="minus 1"):
        nabo.BatchPlan(pts, off, qs, qoff, 100)
