"""Seeded synthetic workloads (numpy PCG64), float32, one point per row.

A cloud of n points in d dimensions is an (n, d) C-order array, i.e. the
column-major d x n matrix libnabo takes (nabo/nabo.h:260-264).

    uniform_cube       10k x 10k uniform [0,1)^3 kd-tree case (bucket 8, k=10)
    lidar_scene        1M points on 20 planes + 10 cylinders of a 100 m scene,
                       queries = the cloud under a small rigid motion + noise
    dirichlet_descriptors  200k x 33 histogram-like descriptors + jittered queries
    uneven_clouds      512 clouds of log-uniform size 2k..200k for batching
    served_pool        64 clouds of log-uniform size 10k..2M and a Poisson
                       request stream over them (query batches 1k..500k)
"""
from __future__ import annotations

import math

import numpy as np


def uniform_cube(n: int = 10_000, m: int = 10_000, dim: int = 3, seed: int = 1):
    rng = np.random.default_rng(seed)
    cloud = rng.random((n, dim), dtype=np.float32)
    query = rng.random((m, dim), dtype=np.float32)
    return cloud, query


def _unit(rng, count):
    v = rng.normal(size=(count, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _basis(normal):
    a = np.array([1.0, 0.0, 0.0]) if abs(normal[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(normal, a)
    u /= np.linalg.norm(u)
    v = np.cross(normal, u)
    return u, v


def rotation(axis, angle):
    """Rodrigues rotation matrix."""
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(angle) * k + (1 - math.cos(angle)) * (k @ k)


def lidar_scene(n: int = 1_000_000, seed: int = 2, sigma: float = 0.01, planes: int = 20,
                cylinders: int = 10, size: float = 100.0, max_rot_deg: float = 2.0,
                max_trans: float = 0.2):
    """Returns (cloud, query): points ∝ surface area on planes and cylinders
    in a size^3 cube with Gaussian noise; query = cloud under a random rigid
    motion (rotation <= max_rot_deg, translation <= max_trans) + fresh noise."""
    rng = np.random.default_rng(seed)
    objs = []
    for _ in range(planes):
        normal = _unit(rng, 1)[0]
        u, v = _basis(normal)
        lu, lv = rng.uniform(10, 100, size=2)
        center = rng.uniform(0, size, size=3)
        objs.append(("plane", lu * lv, (center, u, v, lu, lv)))
    for _ in range(cylinders):
        axis = _unit(rng, 1)[0]
        u, v = _basis(axis)
        radius = rng.uniform(0.2, 2.0)
        length = rng.uniform(5, 50)
        center = rng.uniform(0, size, size=3)
        objs.append(("cyl", 2 * math.pi * radius * length, (center, axis, u, v, radius, length)))
    areas = np.array([o[1] for o in objs])
    counts = rng.multinomial(n, areas / areas.sum())
    parts = []
    for (kind, _, prm), c in zip(objs, counts):
        if c == 0:
            continue
        if kind == "plane":
            center, u, v, lu, lv = prm
            a = rng.uniform(-lu / 2, lu / 2, size=(c, 1))
            b = rng.uniform(-lv / 2, lv / 2, size=(c, 1))
            parts.append(center + a * u + b * v)
        else:
            center, axis, u, v, radius, length = prm
            t = rng.uniform(0, 2 * math.pi, size=(c, 1))
            h = rng.uniform(-length / 2, length / 2, size=(c, 1))
            parts.append(center + h * axis + radius * (np.cos(t) * u + np.sin(t) * v))
    pts = np.concatenate(parts, axis=0)
    cloud = (pts + rng.normal(scale=sigma, size=pts.shape)).astype(np.float32)
    rot = rotation(_unit(rng, 1)[0], math.radians(rng.uniform(0, max_rot_deg)))
    trans = _unit(rng, 1)[0] * rng.uniform(0, max_trans)
    moved = pts @ rot.T + trans
    query = (moved + rng.normal(scale=sigma, size=pts.shape)).astype(np.float32)
    return cloud, query


def dirichlet_descriptors(n: int = 200_000, m: int = 50_000, dim: int = 33, seed: int = 3,
                          alpha: float = 0.5, total: float = 100.0, jitter: float = 0.05):
    rng = np.random.default_rng(seed)
    cloud = (rng.dirichlet(np.full(dim, alpha), size=n) * total).astype(np.float32)
    pick = rng.integers(0, n, size=m)
    base = cloud[pick].astype(np.float64)
    query = base * (1 + jitter * rng.normal(size=base.shape))
    query = np.maximum(query, 0).astype(np.float32)
    return cloud, query


def uneven_clouds(count: int = 512, lo: int = 2_000, hi: int = 200_000, seed: int = 4):
    """List of (cloud, query) pairs: U[0,1)^3 x scale U[1,50] m, queries =
    copy + N(0, 0.01 * scale) noise."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(math.log(lo), math.log(hi), size=count))).astype(int)
    out = []
    for s in sizes:
        scale = rng.uniform(1, 50)
        cloud = (rng.random((s, 3)) * scale).astype(np.float32)
        query = (cloud + rng.normal(scale=0.01 * scale, size=cloud.shape)).astype(np.float32)
        out.append((cloud, query))
    return out


def load_reference_scan(name: str) -> np.ndarray:
    """The reference's tests/data scans, committed as float32 fixtures."""
    import os

    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    return np.load(os.path.join(here, "tests", "golden", f"{name}.npy"))


def served_pool(count: int = 64, lo: int = 10_000, hi: int = 2_000_000, seed: int = 5):
    """Clouds U[0,1)^3 x scale U[1,50] m, sizes log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(math.log(lo), math.log(hi), size=count))).astype(int)
    return [(rng.random((s, 3)) * rng.uniform(1, 50)).astype(np.float32) for s in sizes]


def served_requests(pool, count: int = 150, lo: int = 1_000, hi: int = 500_000,
                    ks=(1, 8, 32), radius: float = 0.5, seed: int = 6):
    """Requests (cloud id, queries, k, radius): each query batch is a
    log-uniform number of cloud points perturbed by N(0, 0.01 m)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        c = int(rng.integers(0, len(pool)))
        m = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        base = pool[c][rng.integers(0, len(pool[c]), m)]
        q = (base + rng.normal(scale=0.01, size=base.shape)).astype(np.float32)
        out.append((c, q, int(rng.choice(ks)), radius))
    return out
