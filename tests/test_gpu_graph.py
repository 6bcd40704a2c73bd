"""kNN calls captured in a HIP graph (torch.cuda.CUDAGraph) and replayed.

Single-tree knn_device calls are capturable; the batched search refuses
capture (test_batched_capture_refused) since a replay of it faulted.

A knn_device call launches its whole sequence (counter zeroing, query sort
kernels, scan, search) on the caller's stream and allocates its scratch
stream-ordered, so the call can be captured and replayed. Each replay runs
with eager GPU work and synchronizations between replays, and every replay
must reproduce the eager call's indices and dists2 bit for bit (which the
other GPU tests tie to the oracle).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _capture(fn):
    import torch

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()  # warm-up outside the capture
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        fn()
    torch.cuda.synchronize()
    return graph


def _replays_match(graph, outs, want, reps=3):
    import torch

    for r in range(reps):
        for o in outs:
            o.fill_(-7)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for o, w in zip(outs, want):
            assert torch.equal(o, w), f"replay {r} differs from the eager call"
        # eager work between the replays
        torch.randn(4096, device="cuda").sum().item()


@pytest.mark.parametrize("k,sort", [(16, 1), (5, 1), (16, 0)])
def test_kdtree_knn_graph_replay(nabo, k, sort):
    import torch

    from libnabo_amd import datasets

    cloud, query = datasets.lidar_scene(60_000, seed=11)
    nns = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_HIP, 3, 0,
                                      {"bucketSize": 8, "querySort": sort})
    dq = torch.from_numpy(query).cuda()
    m = len(query)
    ei = torch.empty((m, k), dtype=torch.int32, device="cuda")
    ed = torch.empty((m, k), dtype=torch.float32, device="cuda")
    nns.knn_device(dq, k, ei, ed, 0.0, 0, 1.0)
    torch.cuda.synchronize()
    gi, gd = torch.empty_like(ei), torch.empty_like(ed)
    graph = _capture(lambda: nns.knn_device(dq, k, gi, gd, 0.0, 0, 1.0))
    _replays_match(graph, [gi, gd], [ei, ed])


@pytest.mark.skip(reason="a graph replay of the batched knn made an illegal memory access on the "
                         "GPU (cause not found); nabo_hip_batch_knn refuses stream capture")
def test_batched_knn_graph_replay(nabo):
    import torch

    rng = np.random.default_rng(12)
    sizes = [3000, 17, 9, 12000, 500, 12, 7000]  # each at least k + 1 points
    clouds = [(rng.random((s, 3)) * rng.uniform(1, 30)).astype(np.float32) for s in sizes]
    queries = [(c + rng.normal(scale=0.05, size=c.shape)).astype(np.float32) for c in clouds]
    off = np.concatenate([[0], np.cumsum([len(c) for c in clouds])]).astype(np.int64)
    qoff = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.int64)
    pts = torch.from_numpy(np.concatenate(clouds)).cuda()
    qs = torch.from_numpy(np.concatenate(queries)).cuda()
    dqoff = torch.from_numpy(qoff).cuda()
    b = nabo.BatchSearch(pts, off, 3, 0, {"bucketSize": 8})
    k = 8
    ei, ed = b.knn(qs, dqoff, k)
    torch.cuda.synchronize()
    gi, gd = torch.empty_like(ei), torch.empty_like(ed)
    graph = _capture(lambda: b.knn(qs, dqoff, k, indices=gi, dists2=gd))
    _replays_match(graph, [gi, gd], [ei, ed])
    b.close()


def test_batched_capture_refused(nabo):
    """Capture of the batched search is refused on the host, before any launch."""
    import torch

    rng = np.random.default_rng(13)
    clouds = [(rng.random((s, 3)) * 10).astype(np.float32) for s in (300, 500)]
    off = np.array([0, 300, 800], np.int64)
    pts = torch.from_numpy(np.concatenate(clouds)).cuda()
    qs = pts.clone()
    dqoff = torch.from_numpy(off).cuda()
    b = nabo.BatchSearch(pts, off, 3, 0, {"bucketSize": 8})
    gi = torch.empty((800, 4), dtype=torch.int32, device="cuda")
    gd = torch.empty((800, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    errors = []
    with torch.cuda.graph(graph):  # the capture itself ends normally
        try:
            b.knn(qs, dqoff, 4, indices=gi, dists2=gd)
        except nabo.NaboError as e:
            errors.append(str(e))
    torch.cuda.synchronize()
    assert errors and "captured" in errors[0]
    # eager calls still work afterwards
    i, _ = b.knn(qs, dqoff, 4, optionFlags=int(nabo.SearchOptionFlags.ALLOW_SELF_MATCH))
    torch.cuda.synchronize()
    assert bool((i[:, 0] == torch.from_numpy(np.r_[np.arange(300), np.arange(500)]).cuda()).all())
    b.close()
