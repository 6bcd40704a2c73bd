"""kNN calls and HIP graph (stream) capture.

Replays of captured knn calls made illegal memory accesses on the GPU,
intermittently: once for a batched search, once for a single-tree search
after other captured graphs had run in the same process (eager calls never
did). The captured calls then allocated and freed their temporaries
stream-ordered inside the capture (graph memory nodes) and, later, took
them from an event-ordered pool; which of these faulted was not shown. The
plain knn entry points therefore refuse stream capture with
NABO_HIP_ERR_UNSUPPORTED before they launch anything, and the replay test
below stays skipped. The graph form is the batch plan
(nabo_hip_batch_plan_*, tests/test_gpu_plan.py): kernel nodes only, recorded
explicitly, with every buffer owned by the plan.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPLAY_SKIP = ("replays of captured knn calls faulted the GPU intermittently (cause not found); "
               "knn refuses stream capture")
UNSUPPORTED = 4  # NABO_HIP_ERR_UNSUPPORTED


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
        torch.randn(4096, device="cuda").sum().item()


def _refused_in_capture(fn, nabo):
    """Runs fn inside a capture that then ends normally; returns its errors."""
    import torch

    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    errors = []
    with torch.cuda.graph(graph):
        try:
            fn()
        except nabo.NaboError as e:
            errors.append(e)
    torch.cuda.synchronize()
    return errors


@pytest.mark.skip(reason=REPLAY_SKIP)
@pytest.mark.parametrize("k,sort", [(16, 1), (5, 1), (16, 0)])
def test_kdtree_knn_graph_replay(nabo, k, sort):
    import torch

    from libnabo_amd import datasets

    cloud, query = datasets.lidar_scene(60_000, seed=11)
    nns = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_HIP, 3,
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


def test_kdtree_capture_refused(nabo):
    import torch

    from libnabo_amd import datasets

    cloud, query = datasets.uniform_cube(5_000, 2_000, seed=14)
    nns = nabo.NearestNeighbourSearch(cloud, nabo.SearchType.KDTREE_HIP, 3, {"bucketSize": 8})
    dq = torch.from_numpy(query).cuda()
    gi = torch.empty((len(query), 4), dtype=torch.int32, device="cuda")
    gd = torch.empty((len(query), 4), dtype=torch.float32, device="cuda")
    errors = _refused_in_capture(lambda: nns.knn_device(dq, 4, gi, gd), nabo)
    assert errors and errors[0].code == UNSUPPORTED and "captured" in str(errors[0])
    # eager calls still work and match the host path
    nns.knn_device(dq, 4, gi, gd)
    torch.cuda.synchronize()
    hi, hd = nns.knn(query, 4)
    assert np.array_equal(gi.cpu().numpy(), hi)
    assert np.array_equal(gd.cpu().numpy().view(np.uint32), hd.view(np.uint32))


def test_batched_capture_refused(nabo):
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
    errors = _refused_in_capture(lambda: b.knn(qs, dqoff, 4, indices=gi, dists2=gd), nabo)
    assert errors and errors[0].code == UNSUPPORTED and "captured" in str(errors[0])
    i, _ = b.knn(qs, dqoff, 4, optionFlags=int(nabo.SearchOptionFlags.ALLOW_SELF_MATCH))
    torch.cuda.synchronize()
    assert bool((i[:, 0] == torch.from_numpy(np.r_[np.arange(300), np.arange(500)]).cuda()).all())
    b.close()
