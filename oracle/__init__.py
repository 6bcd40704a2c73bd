"""ctypes front-end of the CPU oracle (oracle/nabo_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, as the checker. The product (libnabo_amd)
never imports it.

Arrays follow libnabo's layout: a cloud is ``n x rows`` in numpy C order,
i.e. the column-major ``rows x n`` Eigen matrix of nabo/nabo.h:260-264 with
one point per column; results are ``m x k`` (column i of the reference's
``k x m`` index/dists2 matrices is row i here).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnabo_oracle.so")

_i64, _i32, _u32, _f32 = C.c_int64, C.c_int, C.c_uint, C.c_float
_vp = C.c_void_p


def build(mutant: int = 0) -> str:
    """Compiles the oracle (or a planted-mistake build) with make; returns its path."""
    if mutant:
        subprocess.run(["make", "-s", "-C", HERE, f"mutant{mutant}"], check=True)
        return os.path.join(HERE, "_mutants", f"libnabo_oracle_m{mutant}.so")
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_vp)


class Oracle:
    """One loaded oracle library (the real one or a mutant)."""

    def __init__(self, path: str = LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library missing: {path} (run make -C oracle)")
        lib = C.CDLL(path)
        lib.oracle_last_error.restype = C.c_char_p
        lib.oracle_kd_build.restype = _vp
        lib.oracle_kd_build.argtypes = [_vp, _i32, _i64, _i32, _u32]
        lib.oracle_kd_free.argtypes = [_vp]
        lib.oracle_kd_node_count.restype = _i64
        lib.oracle_kd_node_count.argtypes = [_vp]
        lib.oracle_kd_depth.argtypes = [_vp]
        lib.oracle_kd_dim_bits.argtypes = [_vp]
        lib.oracle_kd_export.argtypes = [_vp, _vp, _vp, _vp, _vp]
        lib.oracle_kd_knn.argtypes = [_vp, _vp, _i32, _i64, _i32, _f32, _u32, _f32, _vp, _i32,
                                      _vp, _vp, _vp, _vp, _i32]
        lib.oracle_bf_knn.argtypes = [_vp, _i32, _i64, _i32, _vp, _i32, _i64, _i32, _u32, _f32,
                                      _vp, _vp, _vp, _i32]
        lib.oracle_dist2_f64.argtypes = [_vp, _i32, _i32, _vp, _i32, _i64, _i32, _vp, _vp]
        lib.oracle_bf_kth_f64.argtypes = [_vp, _i32, _i64, _i32, _vp, _i32, _i64, _i32, _u32, _vp]
        lib.oracle_now.restype = C.c_double
        self.lib = lib
        self.mutant = lib.oracle_mutant()

    def error(self) -> str:
        return self.lib.oracle_last_error().decode()

    # -- kd-tree ---------------------------------------------------------
    def kdtree(self, cloud: np.ndarray, dim: int | None = None, bucket_size: int = 8) -> "OracleTree":
        return OracleTree(self, cloud, dim, bucket_size)

    # -- brute force -----------------------------------------------------
    def bf_knn(self, cloud, query, k, flags=0, max_radius=np.inf, radii=None, dim=None,
               threads=0):
        cloud = np.ascontiguousarray(cloud, dtype=np.float32)
        query = np.ascontiguousarray(query, dtype=np.float32)
        n, rows = cloud.shape
        m, qrows = query.shape
        dim = rows if dim is None else dim
        oi = np.empty((m, k), np.int32)
        od = np.empty((m, k), np.float32)
        r = None if radii is None else np.ascontiguousarray(radii, np.float32)
        rc = self.lib.oracle_bf_knn(_ptr(cloud), rows, n, dim, _ptr(query), qrows, m, k, flags,
                                    float(max_radius), _ptr(r), _ptr(oi), _ptr(od), threads)
        if rc:
            raise RuntimeError(self.error())
        return oi, od

    def dist2_f64(self, cloud, query, idx, dim=None):
        cloud = np.ascontiguousarray(cloud, dtype=np.float32)
        query = np.ascontiguousarray(query, dtype=np.float32)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        n, rows = cloud.shape
        m, qrows = query.shape
        k = idx.shape[1]
        out = np.empty((m, k), np.float64)
        self.lib.oracle_dist2_f64(_ptr(cloud), rows, rows if dim is None else dim, _ptr(query),
                                  qrows, m, k, _ptr(idx), _ptr(out))
        return out

    def bf_kth_f64(self, cloud, query, k, flags=0, dim=None):
        cloud = np.ascontiguousarray(cloud, dtype=np.float32)
        query = np.ascontiguousarray(query, dtype=np.float32)
        n, rows = cloud.shape
        m, qrows = query.shape
        out = np.empty((m, k), np.float64)
        self.lib.oracle_bf_kth_f64(_ptr(cloud), rows, n, rows if dim is None else dim,
                                   _ptr(query), qrows, m, k, flags, _ptr(out))
        return out

    def now(self) -> float:
        return self.lib.oracle_now()


class OracleTree:
    """Oracle kd-tree (KDTREE_LINEAR_HEAP / KDTREE_TREE_HEAP semantics)."""

    def __init__(self, oracle: Oracle, cloud, dim, bucket_size):
        self.o = oracle
        self.cloud = np.ascontiguousarray(cloud, dtype=np.float32)
        n, rows = self.cloud.shape
        self.n, self.rows = n, rows
        dim = rows if dim is None else dim
        h = oracle.lib.oracle_kd_build(_ptr(self.cloud), rows, n, dim, bucket_size)
        if not h:
            raise RuntimeError(oracle.error())
        self.h = h
        self.dim = min(dim, rows)

    def __del__(self):
        if getattr(self, "h", None):
            self.o.lib.oracle_kd_free(self.h)
            self.h = None

    @property
    def depth(self) -> int:
        return self.o.lib.oracle_kd_depth(self.h)

    @property
    def dim_bits(self) -> int:
        return self.o.lib.oracle_kd_dim_bits(self.h)

    def export(self):
        """Returns (nodes[count,2] uint32, bucket_index[n] int32, min_bound, max_bound)."""
        cnt = self.o.lib.oracle_kd_node_count(self.h)
        nodes = np.empty((cnt, 2), np.uint32)
        bucket = np.empty(self.n, np.int32)
        mn = np.empty(self.dim, np.float32)
        mx = np.empty(self.dim, np.float32)
        self.o.lib.oracle_kd_export(self.h, _ptr(nodes), _ptr(bucket), _ptr(mn), _ptr(mx))
        return nodes, bucket, mn, mx

    def knn(self, query, k, eps=0.0, flags=0, max_radius=np.inf, radii=None, heap="linear",
            threads=0, stats=False):
        query = np.ascontiguousarray(query, dtype=np.float32)
        m, qrows = query.shape
        oi = np.empty((m, k), np.int32)
        od = np.empty((m, k), np.float32)
        touched = np.empty(m, np.uint64) if stats else None
        visits = np.empty(m, np.uint32) if stats else None
        r = None if radii is None else np.ascontiguousarray(radii, np.float32)
        rc = self.o.lib.oracle_kd_knn(self.h, _ptr(query), qrows, m, k, float(eps), flags,
                                      float(max_radius), _ptr(r), 0 if heap == "linear" else 1,
                                      _ptr(oi), _ptr(od), _ptr(touched), _ptr(visits), threads)
        if rc:
            raise RuntimeError(self.o.error())
        if stats:
            return oi, od, touched, visits
        return oi, od


_DEFAULT: Oracle | None = None


def default() -> Oracle:
    """The real (non-mutant) oracle, built on first use when missing."""
    global _DEFAULT
    if _DEFAULT is None:
        if not os.path.exists(LIB_PATH):
            build()
        _DEFAULT = Oracle(LIB_PATH)
    return _DEFAULT
