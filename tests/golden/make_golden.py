"""Writes the committed fixtures under tests/golden/ (run in the build container).

1. scan.2d / scan.3d / scan.3d.large: the point clouds of libnabo's own
   validation tests (tests/data/*.txt, read as libnabo's loader does:
   tests/helpers.h:46-80 parses each token with atof, then the matrix is
   float), stored as float32 .npy. These are data files of the reference.
2. ckdtree_*.npz: exact kNN answers from scipy.spatial.cKDTree (scipy
   1.15.3, an independent exact kNN in float64 on the float32 inputs) for
   seeded inputs: indices and float64 squared distances.

Usage: python tests/golden/make_golden.py [/path/to/libnabo]
"""
from __future__ import annotations

import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def load_scan(path: str) -> np.ndarray:
    rows = []
    dim = None
    with open(path, "r", encoding="utf-8") as f:
        for line in f:
            toks = [t for t in re.split(r"[ \t,;]+", line.strip()) if t]
            if not toks:
                continue
            vals = [float(t) for t in toks]
            if dim is None:
                dim = len(vals)
            rows.append(vals)
    flat = np.array([v for r in rows for v in r], dtype=np.float64)
    return flat.astype(np.float32).reshape(-1, dim)


def ckdtree_case(name, cloud, query, k, allow_self=False, max_radius=np.inf):
    from scipy.spatial import cKDTree

    c64 = cloud.astype(np.float64)
    q64 = query.astype(np.float64)
    tree = cKDTree(c64)
    kk = k if allow_self else k + 4  # room to drop self / zero-distance hits
    d, i = tree.query(q64, k=kk, distance_upper_bound=max_radius)
    d2 = np.where(np.isfinite(d), d * d, np.inf)
    i = np.where(np.isfinite(d), i, -1).astype(np.int32)
    if not allow_self:
        eps = np.finfo(np.float32).eps
        out_i = np.full((len(query), k), -1, np.int32)
        out_d = np.full((len(query), k), np.inf)
        for r in range(len(query)):
            keep = d2[r] > eps
            out_i[r] = np.concatenate([i[r][keep], np.full(k, -1, np.int32)])[:k]
            out_d[r] = np.concatenate([d2[r][keep], np.full(k, np.inf)])[:k]
        i, d2 = out_i, out_d
    np.savez_compressed(os.path.join(HERE, f"ckdtree_{name}.npz"), cloud=cloud, query=query,
                        k=k, allow_self=allow_self, max_radius=max_radius, indices=i, dists2=d2)


def main(ref_root: str = "/root/reference"):
    data = os.path.join(ref_root, "tests", "data")
    for name in ("scan.2d", "scan.3d", "scan.3d.large"):
        arr = load_scan(os.path.join(data, f"{name}.txt"))
        np.save(os.path.join(HERE, f"{name}.npy"), arr)
        print(name, arr.shape)

    from libnabo_amd import datasets

    cloud, query = datasets.uniform_cube(10_000, 2_000, seed=1)
    ckdtree_case("uniform_k10", cloud, query, 10)
    scan3 = np.load(os.path.join(HERE, "scan.3d.npy"))
    ckdtree_case("scan3d_self_k10", scan3, scan3, 10, allow_self=True)
    rng = np.random.default_rng(7)
    q = scan3[rng.integers(0, len(scan3), 500)] + rng.normal(scale=0.05, size=(500, 3)).astype(
        np.float32)
    ckdtree_case("scan3d_radius_k8", scan3, q.astype(np.float32), 8, max_radius=0.5)
    print("done")


if __name__ == "__main__":
    main(*sys.argv[1:])
