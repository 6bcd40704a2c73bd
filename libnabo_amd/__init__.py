"""libnabo_amd: libnabo's kNN search on MI355X, behind libnabo's own interface.

Python front-end of the C ABI in include/nabo_hip.h (libnabo_hip.so, built
in-tree by libnabo_amd.build). It mirrors pynabo (python/nabo.cpp:95-217 of
libnabo): ``NearestNeighbourSearch(cloud, searchType, dim, params)`` and
``knn(query, k, epsilon, optionFlags, maxRadius) -> (indices, dists2)``, with
the same layout rule: the contiguous axis of a 2-D array holds the
coordinates of one point, and results keep the query's order with the k
results of a query contiguous. Unlike pynabo (double only, python/nabo.cpp:10)
the arithmetic is float32, as in ``Nabo::NNSearchF``.

There is no CPU fallback: without libnabo_hip.so, or without a GPU, calls
raise.
"""
from __future__ import annotations

import ctypes as C
import enum
import os

import numpy as np

__all__ = [
    "SearchType", "SearchOptionFlags", "CreationOptionFlags", "NearestNeighbourSearch",
    "NaboError", "load_library", "LIB_PATH", "Params", "BatchSearch", "BatchPlan", "Server",
]

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libnabo_hip.so")


class SearchType(enum.IntEnum):
    """nabo/nabo.h:290-299, with the HIP types appended."""
    BRUTE_FORCE = 0
    KDTREE_LINEAR_HEAP = 1
    KDTREE_TREE_HEAP = 2
    KDTREE_CL_PT_IN_NODES = 3
    KDTREE_CL_PT_IN_LEAVES = 4
    BRUTE_FORCE_CL = 5
    KDTREE_HIP = 6
    BRUTE_FORCE_HIP = 7
    SEARCH_TYPE_COUNT = 8


class CreationOptionFlags(enum.IntFlag):
    """nabo/nabo.h:302-305."""
    TOUCH_STATISTICS = 1


class SearchOptionFlags(enum.IntFlag):
    """nabo/nabo.h:308-312."""
    ALLOW_SELF_MATCH = 1
    SORT_RESULTS = 2


STATUS = {0: "ok", 1: "argument", 2: "hip", 3: "memory", 4: "unsupported"}
BF_PRECISION = {"fp32": 0, "bf16": 1, "fp16": 2}
HEAP = {"linear": 0, "stl": 1}  # NABO_HIP_HEAP_*


class NaboError(RuntimeError):
    """Raised where the reference throws Nabo::runtime_error (nabo/nabo.cpp:50-66)."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code
        self.kind = STATUS.get(code, "unknown")


class Params(C.Structure):
    _fields_ = [
        ("bucket_size", C.c_uint32),
        ("query_sort", C.c_int32),
        ("bf_precision", C.c_int32),
        ("bf_extra", C.c_int32),
        ("device", C.c_int32),
        ("heap", C.c_int32),
        ("wave_queries", C.c_int32),
        ("stack_place", C.c_int32),
        ("stl_log", C.c_int32),
        ("reserved", C.c_int32 * 7),
    ]


class Info(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("rows", C.c_int32),
        ("dim", C.c_int32),
        ("search_type", C.c_int32),
        ("bucket_size", C.c_uint32),
        ("dim_bits", C.c_uint32),
        ("depth", C.c_int32),
        ("node_count", C.c_int64),
        ("device_bytes", C.c_int64),
        ("creation_flags", C.c_uint32),
        ("device", C.c_int32),
        ("clouds", C.c_int32),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_int64),
        ("slots", C.c_int32),
        ("max_merge_queries", C.c_int64),
        ("device", C.c_int32),
        ("host_cache_bytes", C.c_int64),
        ("reserved", C.c_int32 * 6),
    ]


class ServerStats(C.Structure):
    _fields_ = [(name, C.c_int64) for name in (
        "submitted", "completed", "launches", "merged_requests", "builds", "evictions",
        "resident_trees", "resident_bytes", "peak_bytes", "readmits", "host_cached_bytes")]


EXPORTED = [
    "nabo_hip_last_error", "nabo_hip_params_default", "nabo_hip_create",
    "nabo_hip_create_device", "nabo_hip_destroy", "nabo_hip_knn", "nabo_hip_knn_device",
    "nabo_hip_get_info", "nabo_hip_get_bounds", "nabo_hip_get_tree", "nabo_hip_device_count",
    "nabo_hip_kernel_timing", "nabo_hip_kernel_timing_collect", "nabo_hip_batch_create",
    "nabo_hip_batch_knn", "nabo_hip_batch_roots", "nabo_hip_server_config_default",
    "nabo_hip_server_create", "nabo_hip_server_destroy", "nabo_hip_server_add_cloud",
    "nabo_hip_server_submit", "nabo_hip_server_wait", "nabo_hip_server_get_stats",
    "nabo_hip_serialize", "nabo_hip_deserialize", "nabo_hip_knn_device_noalloc",
    "nabo_hip_batch_plan_create", "nabo_hip_batch_plan_launch", "nabo_hip_batch_plan_status",
    "nabo_hip_batch_plan_destroy",
]

_LIBS: dict = {}


def load_library(path: str | None = None):
    """Loads libnabo_hip.so (raises when it has not been built). The default
    path is the module's LIB_PATH as it is at the time of the call."""
    if path is None:
        path = LIB_PATH
    if path in _LIBS:
        return _LIBS[path]
    if not os.path.exists(path):
        raise NaboError(2, f"{path} is missing: build it with `python -m libnabo_amd.build`")
    # torch bundles its own HIP runtime under the same soname
    # (libamdhip64.so.7); whichever is loaded first serves the whole process.
    # Load torch's first when torch is present, so that torch tensors and
    # streams and this library share one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    vp, i32, i64, u32, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_float
    lib.nabo_hip_last_error.restype = C.c_char_p
    lib.nabo_hip_params_default.argtypes = [C.POINTER(Params)]
    lib.nabo_hip_create.argtypes = [vp, i64, i32, i64, i32, i32, u32, C.POINTER(Params),
                                    C.POINTER(vp)]
    lib.nabo_hip_create_device.argtypes = [vp, i64, i32, i64, i32, i32, u32, C.POINTER(Params),
                                           vp, C.POINTER(vp)]
    lib.nabo_hip_destroy.argtypes = [vp]
    lib.nabo_hip_knn.argtypes = [vp, vp, i32, i64, i64, i32, f32, u32, f32, vp, vp, vp,
                                 C.POINTER(C.c_uint64)]
    lib.nabo_hip_knn_device.argtypes = [vp, vp, i32, i64, i64, i32, f32, u32, f32, vp, vp, vp,
                                        vp, vp]
    lib.nabo_hip_knn_device_noalloc.argtypes = [vp, vp, i32, i64, i64, i32, f32, u32, f32, vp,
                                                vp, vp, vp, vp]
    lib.nabo_hip_batch_plan_create.argtypes = [vp, vp, i32, i32, i64, i32, C.POINTER(Params), vp,
                                               i64, vp, i32, f32, u32, f32, vp, vp, i32,
                                               C.POINTER(vp)]
    lib.nabo_hip_batch_plan_launch.argtypes = [vp, vp]
    lib.nabo_hip_batch_plan_status.argtypes = [vp, C.POINTER(i32), C.POINTER(i32)]
    lib.nabo_hip_batch_plan_destroy.argtypes = [vp]
    lib.nabo_hip_get_info.argtypes = [vp, C.POINTER(Info)]
    lib.nabo_hip_get_bounds.argtypes = [vp, vp, vp]
    lib.nabo_hip_get_tree.argtypes = [vp, vp, vp]
    lib.nabo_hip_device_count.argtypes = [C.POINTER(C.c_int32)]
    lib.nabo_hip_batch_create.argtypes = [vp, vp, i32, i32, i64, i32, u32, C.POINTER(Params), vp,
                                          C.POINTER(vp)]
    lib.nabo_hip_batch_knn.argtypes = [vp, vp, i32, i64, i64, vp, i32, f32, u32, f32, vp, vp, vp,
                                       vp]
    lib.nabo_hip_batch_roots.argtypes = [vp, vp]
    lib.nabo_hip_server_config_default.argtypes = [C.POINTER(ServerConfig)]
    lib.nabo_hip_server_config_default.restype = None
    lib.nabo_hip_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(vp)]
    lib.nabo_hip_server_destroy.argtypes = [vp]
    lib.nabo_hip_server_add_cloud.argtypes = [vp, vp, i64, i32, i64, i32, C.POINTER(Params),
                                              C.POINTER(i32)]
    lib.nabo_hip_server_submit.argtypes = [vp, i32, vp, i32, i64, i64, i32, f32, u32, f32, vp,
                                           vp, C.POINTER(i64)]
    lib.nabo_hip_server_wait.argtypes = [vp, i64, C.POINTER(C.c_double)]
    lib.nabo_hip_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
    lib.nabo_hip_serialize.argtypes = [vp, vp, C.c_uint64, C.POINTER(C.c_uint64)]
    lib.nabo_hip_deserialize.argtypes = [vp, C.c_uint64, i32, C.POINTER(vp)]
    lib.nabo_hip_kernel_timing.argtypes = [i32]
    lib.nabo_hip_kernel_timing_collect.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    for name in EXPORTED:
        if name != "nabo_hip_server_config_default":
            getattr(lib, name).restype = getattr(lib, name).restype or C.c_int
    lib.nabo_hip_last_error.restype = C.c_char_p
    _LIBS[path] = lib
    return lib


def _check(lib, rc: int):
    if rc != 0:
        raise NaboError(rc, lib.nabo_hip_last_error().decode())


def device_count() -> int:
    lib = load_library()
    c = C.c_int32(0)
    _check(lib, lib.nabo_hip_device_count(C.byref(c)))
    return c.value


def kernel_timing(enable: bool) -> None:
    """Brackets every search-kernel launch with HIP events on its stream."""
    lib = load_library()
    _check(lib, lib.nabo_hip_kernel_timing(1 if enable else 0))


def kernel_timing_collect():
    """(summed kernel milliseconds, launches) since the last collect."""
    lib = load_library()
    ms = C.c_double(0)
    n = C.c_int64(0)
    _check(lib, lib.nabo_hip_kernel_timing_collect(C.byref(ms), C.byref(n)))
    return ms.value, n.value


def _as_points(a, name: str) -> np.ndarray:
    """(count, rows) float32 C-order view: pynabo's layout rule (python/nabo.cpp:35-49)."""
    a = np.asarray(a)
    if a.ndim != 2:
        raise NaboError(1, f'Argument "{name}" must be a two-dimensional array')
    if a.flags.f_contiguous and not a.flags.c_contiguous:
        a = a.T  # contiguous axis = coordinates
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


STACK_PLACE = {"auto": 0, "hbm": 1}  # NABO_HIP_STACK_*


def make_params(bucket_size: int = 8, query_sort: bool = True, bf_precision: str = "fp32",
                bf_extra: int = 32, device: int = -1, heap="linear", wave_queries: int = 0,
                stack_place="auto", stl_log: int = 0, **unknown) -> Params:
    """Parameters map (nabo/nabo.h:231-254): only "bucketSize" is a reference key.

    ``heap``: "linear" (IndexHeapBruteForceVector results) or "stl"
    (IndexHeapSTL results: heap order unless SORT_RESULTS). The kd-tree
    search's schedule (results do not depend on it): ``wave_queries``
    (queries a wavefront claims at a time: 64, 128, 256 or 512; 0 = 64),
    ``stack_place`` ("auto": LDS stacks for trees up to 32 levels; "hbm"),
    ``stl_log`` (IndexHeapSTL, k <= 96: admission-log entries per lane).
    """
    if unknown:
        raise NaboError(1, f"unknown parameters {sorted(unknown)}")
    p = Params()
    load_library().nabo_hip_params_default(C.byref(p))
    p.bucket_size = int(bucket_size)
    p.query_sort = 1 if query_sort else 0
    p.bf_precision = BF_PRECISION[bf_precision]
    p.bf_extra = int(bf_extra)
    p.device = int(device)
    p.heap = HEAP[heap] if isinstance(heap, str) else int(heap)
    p.wave_queries = int(wave_queries)
    p.stack_place = STACK_PLACE[stack_place] if isinstance(stack_place, str) else int(stack_place)
    p.stl_log = int(stl_log)
    return p


_PARAM_KEYS = {"bucketSize": "bucket_size", "querySort": "query_sort",
               "bfPrecision": "bf_precision", "bfExtra": "bf_extra", "device": "device",
               "heap": "heap", "waveQueries": "wave_queries", "stackPlace": "stack_place",
               "stlLog": "stl_log"}


class _Serialized:
    def __init__(self, data, device):
        self.data = data
        self.device = device


class NearestNeighbourSearch:
    """pynabo.NearestNeighbourSearch on the GPU (python/nabo.cpp:95-163).

    The positional arguments are pynabo's, in its order (python/nabo.cpp:98):
    ``cloud`` (array of points, contiguous axis = coordinates),
    ``searchType`` (default KDTREE_LINEAR_HEAP, as in pynabo; every type runs
    on the GPU, see SearchType), ``dim`` and ``params`` ({"bucketSize": 8,
    ...}). ``creationOptionFlags`` (TOUCH_STATISTICS), which pynabo does not
    expose, is keyword-only.
    """

    def __init__(self, cloud, searchType=SearchType.KDTREE_LINEAR_HEAP, dim: int = 2**31 - 1,
                 params: dict | None = None, *, creationOptionFlags: int = 0):
        if params is not None and not isinstance(params, dict):
            raise TypeError("params must be a dict (pynabo's fourth argument); pass "
                            "creationOptionFlags by keyword")
        self._lib = load_library()
        self._h = None
        if isinstance(cloud, _Serialized):
            h = C.c_void_p()
            buf = cloud.data
            _check(self._lib, self._lib.nabo_hip_deserialize(_ptr(buf), buf.nbytes, int(cloud.device),
                                                             C.byref(h)))
            self._h = h
            inf = self.info()
            self.cloud = None
            self.searchType = SearchType(inf.search_type)
            self.creationOptionFlags = int(inf.creation_flags)
            self.last_touched = 0
            return
        kw = {}
        for key, val in (params or {}).items():
            if key not in _PARAM_KEYS:
                raise NaboError(1, f"unknown parameter {key!r}")
            kw[_PARAM_KEYS[key]] = val
        # The reference's types keep their k-best containers: KDTREE_TREE_HEAP
        # and BRUTE_FORCE use IndexHeapSTL (nabo.cpp:143, brute_force_cpu.cpp:79).
        if int(searchType) in (SearchType.KDTREE_TREE_HEAP, SearchType.BRUTE_FORCE):
            kw.setdefault("heap", "stl")
        p = make_params(**kw)
        if int(searchType) in (SearchType.KDTREE_LINEAR_HEAP, SearchType.KDTREE_TREE_HEAP):
            searchType = SearchType.KDTREE_HIP
        elif int(searchType) == SearchType.BRUTE_FORCE:
            searchType = SearchType.BRUTE_FORCE_HIP
        pts = _as_points(cloud, "cloud")
        n, rows = pts.shape
        h = C.c_void_p()
        _check(self._lib, self._lib.nabo_hip_create(_ptr(pts), n, rows, rows, int(dim),
                                                    int(searchType), int(creationOptionFlags),
                                                    C.byref(p), C.byref(h)))
        self._h = h
        self.cloud = pts
        self.searchType = SearchType(int(searchType))
        self.creationOptionFlags = int(creationOptionFlags)
        self.last_touched = 0

    def close(self):
        if self._h:
            self._lib.nabo_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- properties -------------------------------------------------------
    def info(self) -> Info:
        inf = Info()
        _check(self._lib, self._lib.nabo_hip_get_info(self._h, C.byref(inf)))
        return inf

    @property
    def dim(self) -> int:
        return self.info().dim

    @property
    def bounds(self):
        d = self.dim
        lo = np.empty(d, np.float32)
        hi = np.empty(d, np.float32)
        _check(self._lib, self._lib.nabo_hip_get_bounds(self._h, _ptr(lo), _ptr(hi)))
        return lo, hi

    def serialize(self) -> np.ndarray:
        """The built index as a uint8 array in host memory (nabo_hip_serialize)."""
        size = C.c_uint64(0)
        _check(self._lib, self._lib.nabo_hip_serialize(self._h, None, 0, C.byref(size)))
        buf = np.empty(size.value, np.uint8)
        _check(self._lib, self._lib.nabo_hip_serialize(self._h, _ptr(buf), buf.nbytes,
                                                       C.byref(size)))
        return buf

    @classmethod
    def deserialize(cls, data, device: int = -1) -> "NearestNeighbourSearch":
        """Recreates an index from serialize() output without a build."""
        return cls(_Serialized(np.ascontiguousarray(data, dtype=np.uint8), device))

    def tree(self):
        """(nodes[count, 2] uint32, bucket point indices[n]) as built on the GPU."""
        inf = self.info()
        nodes = np.empty((inf.node_count, 2), np.uint32)
        bucket = np.empty(inf.n, np.int32)
        _check(self._lib, self._lib.nabo_hip_get_tree(self._h, _ptr(nodes), _ptr(bucket)))
        return nodes, bucket

    # -- search -----------------------------------------------------------
    def knn(self, query, k: int = 1, epsilon: float = 0.0, optionFlags: int = 0,
            maxRadius: float = np.inf, maxRadii=None):
        """Returns (indices, dists2), each (query count, k); see nabo/nabo.h:338,351."""
        q = _as_points(query, "query")
        m, qrows = q.shape
        k = int(k)
        idx = np.empty((m, max(k, 0)), np.int32)
        d2 = np.empty((m, max(k, 0)), np.float32)
        radii = None
        if maxRadii is not None:
            radii = np.ascontiguousarray(maxRadii, dtype=np.float32).ravel()
            if radii.size != m:
                raise NaboError(1, f"Maximum radii vector has not the same length ({radii.size}) "
                                   f"than query has columns ({m})")
        touched = C.c_uint64(0)
        _check(self._lib, self._lib.nabo_hip_knn(self._h, _ptr(q), qrows, qrows, m, k,
                                                 float(epsilon), int(optionFlags),
                                                 float(maxRadius), _ptr(radii), _ptr(idx),
                                                 _ptr(d2), C.byref(touched)))
        self.last_touched = touched.value
        return idx, d2

    def knn_device(self, query, k: int, indices, dists2, epsilon: float = 0.0,
                   optionFlags: int = 0, maxRadius: float = np.inf, maxRadii=None,
                   touched=None, stream=None):
        """Enqueues a search on torch device tensors (query: (m, rows) float32
        contiguous; indices/dists2: (m, k) int32/float32) on ``stream``
        (a torch.cuda.Stream, default: the current one)."""
        import torch

        dev = _index_device(self)
        m, qrows = _check_query_tensor(query, dev)
        k = int(k)
        _check_out_tensor(indices, "indices", torch.int32, (m, k), dev)
        _check_out_tensor(dists2, "dists2", torch.float32, (m, k), dev)
        if maxRadii is not None:
            _check_out_tensor(maxRadii, "maxRadii", torch.float32, (m,), dev)
        if touched is not None:
            _check_out_tensor(touched, "touched", (torch.int64, torch.uint64), (1,), dev,
                              any_shape=True)
        if stream is None:
            stream = torch.cuda.current_stream(query.device)
        _check(self._lib, self._lib.nabo_hip_knn_device(
            self._h, C.c_void_p(query.data_ptr()), qrows, query.stride(0), m, int(k),
            float(epsilon), int(optionFlags), float(maxRadius),
            None if maxRadii is None else C.c_void_p(maxRadii.data_ptr()),
            C.c_void_p(indices.data_ptr()), C.c_void_p(dists2.data_ptr()),
            None if touched is None else C.c_void_p(touched.data_ptr()),
            C.c_void_p(stream.cuda_stream)))


def _index_device(index) -> int:
    if getattr(index, "_device", None) is None:
        index._device = int(index.info().device)
    return index._device


def _tensor_fail(msg):
    raise NaboError(1, msg)


def _check_query_tensor(query, dev: int):
    """A (m, rows) float32 CUDA tensor on device ``dev`` whose rows are
    contiguous (the C ABI reads point i at query + i * stride(0))."""
    import torch

    if not isinstance(query, torch.Tensor) or not query.is_cuda:
        _tensor_fail("query must be a CUDA tensor")
    if query.device.index != dev:
        _tensor_fail(f"query is on cuda:{query.device.index}, the index on cuda:{dev}")
    if query.dtype != torch.float32:
        _tensor_fail(f"query must be float32, not {query.dtype}")
    if query.dim() != 2:
        _tensor_fail("query must be a two-dimensional (count, rows) tensor")
    m, qrows = query.shape
    if m > 0 and (query.stride(1) != 1 or query.stride(0) < qrows):
        _tensor_fail("query rows must be contiguous (stride(1) == 1, stride(0) >= rows)")
    return m, qrows


def _check_out_tensor(t, name: str, dtype, shape, dev: int, any_shape: bool = False):
    """A contiguous CUDA tensor of the given dtype and shape on device ``dev``:
    the kernels write through its raw pointer."""
    import torch

    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        _tensor_fail(f"{name} must be a CUDA tensor")
    if t.device.index != dev:
        _tensor_fail(f"{name} is on cuda:{t.device.index}, the index on cuda:{dev}")
    dtypes = dtype if isinstance(dtype, tuple) else (dtype,)
    if t.dtype not in dtypes:
        _tensor_fail(f"{name} must be {' or '.join(str(d) for d in dtypes)}, not {t.dtype}")
    if any_shape:
        if t.numel() < 1:
            _tensor_fail(f"{name} must hold at least one element")
    elif tuple(t.shape) != tuple(shape):
        _tensor_fail(f"{name} has shape {tuple(t.shape)}, {tuple(shape)} is needed")
    if not t.is_contiguous():
        _tensor_fail(f"{name} must be contiguous")


def _stream_handle(stream, device):
    import torch

    if stream is None:
        stream = torch.cuda.current_stream(device)
    return C.c_void_p(stream.cuda_stream)


class BatchSearch:
    """Many independent clouds: one GPU build pass, one search launch.

    ``points``: torch float32 CUDA tensor (total, rows), contiguous, the
    clouds one after the other; ``offsets``: clouds + 1 point offsets. Tree c
    is the KDTREE_HIP tree of cloud c alone; results of a query set are those
    of that cloud's own search, with indices local to the cloud.
    """

    def __init__(self, points, offsets, dim: int = 2**31 - 1, creationOptionFlags: int = 0,
                 params: dict | None = None, stream=None):
        import torch

        self._lib = load_library()
        self._h = None
        kw = {}
        for key, val in (params or {}).items():
            if key not in _PARAM_KEYS:
                raise NaboError(1, f"unknown parameter {key!r}")
            kw[_PARAM_KEYS[key]] = val
        p = make_params(**kw)
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        if not isinstance(points, torch.Tensor) or not points.is_cuda:
            _tensor_fail("points must be a CUDA tensor")
        if points.dtype != torch.float32 or points.dim() != 2:
            _tensor_fail("points must be a two-dimensional float32 tensor")
        total, rows = points.shape
        if total > 0 and (points.stride(1) != 1 or points.stride(0) < rows):
            _tensor_fail("points rows must be contiguous (stride(1) == 1, stride(0) >= rows)")
        if off.ndim != 1 or off.size < 2:
            raise NaboError(1, "a batch needs at least one cloud (clouds + 1 offsets)")
        if int(off[-1]) != total:
            raise NaboError(1, "the last offset must equal the number of points")
        h = C.c_void_p()
        _check(self._lib, self._lib.nabo_hip_batch_create(
            C.c_void_p(points.data_ptr()), _ptr(off), len(off) - 1, rows, points.stride(0),
            int(dim), int(creationOptionFlags), C.byref(p),
            _stream_handle(stream, points.device), C.byref(h)))
        self._h = h
        self.offsets = off
        self.device = points.device

    def close(self):
        if self._h:
            self._lib.nabo_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def clouds(self) -> int:
        return len(self.offsets) - 1

    def info(self) -> Info:
        inf = Info()
        _check(self._lib, self._lib.nabo_hip_get_info(self._h, C.byref(inf)))
        return inf

    def roots(self) -> np.ndarray:
        r = np.empty(self.clouds, np.int32)
        _check(self._lib, self._lib.nabo_hip_batch_roots(self._h, _ptr(r)))
        return r

    def tree(self):
        """(nodes[count, 2] uint32 with global child ids, cloud-local bucket indices[n])."""
        inf = self.info()
        nodes = np.empty((inf.node_count, 2), np.uint32)
        bucket = np.empty(inf.n, np.int32)
        _check(self._lib, self._lib.nabo_hip_get_tree(self._h, _ptr(nodes), _ptr(bucket)))
        return nodes, bucket

    def bounds(self):
        inf = self.info()
        lo = np.empty((self.clouds, inf.dim), np.float32)
        hi = np.empty((self.clouds, inf.dim), np.float32)
        _check(self._lib, self._lib.nabo_hip_get_bounds(self._h, _ptr(lo), _ptr(hi)))
        return lo, hi

    def knn(self, queries, query_offsets, k: int, epsilon: float = 0.0, optionFlags: int = 0,
            maxRadius: float = np.inf, indices=None, dists2=None, touched=None, stream=None):
        """queries: (m, rows) float32 CUDA tensor; query_offsets: clouds + 1
        int64 CUDA tensor. Returns (indices, dists2) CUDA tensors (m, k)."""
        import torch

        dev = _index_device(self)
        m, qrows = _check_query_tensor(queries, dev)
        k = int(k)
        _check_out_tensor(query_offsets, "query_offsets", torch.int64, (self.clouds + 1,), dev)
        if indices is None:
            indices = torch.empty((m, k), dtype=torch.int32, device=queries.device)
        if dists2 is None:
            dists2 = torch.empty((m, k), dtype=torch.float32, device=queries.device)
        _check_out_tensor(indices, "indices", torch.int32, (m, k), dev)
        _check_out_tensor(dists2, "dists2", torch.float32, (m, k), dev)
        if touched is not None:
            _check_out_tensor(touched, "touched", (torch.int64, torch.uint64), (1,), dev,
                              any_shape=True)
        _check(self._lib, self._lib.nabo_hip_batch_knn(
            self._h, C.c_void_p(queries.data_ptr()), qrows, queries.stride(0), m,
            C.c_void_p(query_offsets.data_ptr()), int(k), float(epsilon), int(optionFlags),
            float(maxRadius), C.c_void_p(indices.data_ptr()), C.c_void_p(dists2.data_ptr()),
            None if touched is None else C.c_void_p(touched.data_ptr()),
            _stream_handle(stream, queries.device)))
        return indices, dists2


class BatchPlan:
    """A batched build + search recorded as one HIP graph (nabo_hip_batch_plan_*).

    ``points`` (total, rows) and ``queries`` (m, rows): float32 CUDA tensors,
    row-contiguous, read at every launch; ``offsets`` / ``query_offsets``:
    clouds + 1 host offsets. ``indices`` / ``dists2`` (m, k) are allocated
    here (or given) and rewritten at every launch. ``levels``: build levels
    recorded (0: the data's depth + 1 + 4, from one build at creation).
    """

    def __init__(self, points, offsets, queries, query_offsets, k: int, epsilon: float = 0.0,
                 optionFlags: int = 0, maxRadius: float = np.inf, params: dict | None = None,
                 levels: int = 0, indices=None, dists2=None):
        import torch

        self._lib = load_library()
        self._h = None
        kw = {}
        for key, val in (params or {}).items():
            if key not in _PARAM_KEYS:
                raise NaboError(1, f"unknown parameter {key!r}")
            kw[_PARAM_KEYS[key]] = val
        p = make_params(**kw)
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        qoff = np.ascontiguousarray(query_offsets, dtype=np.int64)
        if off.ndim != 1 or off.size < 2 or qoff.shape != off.shape:
            raise NaboError(1, "offsets and query_offsets need clouds + 1 entries each")
        dev = points.device.index
        for name, t in (("points", points), ("queries", queries)):
            if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32 \
                    or t.dim() != 2 or t.device.index != dev:
                _tensor_fail(f"{name} must be a 2-D float32 CUDA tensor on cuda:{dev}")
            if t.shape[0] > 0 and (t.stride(1) != 1 or t.stride(0) < t.shape[1]):
                _tensor_fail(f"{name} rows must be contiguous")
        total, rows = points.shape
        m = queries.shape[0]
        if int(off[-1]) != total or int(qoff[-1]) != m:
            raise NaboError(1, "the last offsets must equal the numbers of points and queries")
        if queries.shape[1] != rows:
            raise NaboError(1, "queries must have the points' rows")
        k = int(k)
        if indices is None:
            indices = torch.empty((m, k), dtype=torch.int32, device=points.device)
        if dists2 is None:
            dists2 = torch.empty((m, k), dtype=torch.float32, device=points.device)
        _check_out_tensor(indices, "indices", torch.int32, (m, k), dev)
        _check_out_tensor(dists2, "dists2", torch.float32, (m, k), dev)
        self.points, self.queries, self.indices, self.dists2 = points, queries, indices, dists2
        h = C.c_void_p()
        _check(self._lib, self._lib.nabo_hip_batch_plan_create(
            C.c_void_p(points.data_ptr()), _ptr(off), len(off) - 1, rows, points.stride(0), rows,
            C.byref(p), C.c_void_p(queries.data_ptr()), queries.stride(0), _ptr(qoff), k,
            float(epsilon), int(optionFlags), float(maxRadius), C.c_void_p(indices.data_ptr()),
            C.c_void_p(dists2.data_ptr()), int(levels), C.byref(h)))
        self._h = h
        self.offsets, self.query_offsets = off, qoff

    def launch(self, stream=None):
        """Rebuilds every tree from ``points`` and searches ``queries``: one
        graph launch on ``stream`` (default: the current one)."""
        _check(self._lib, self._lib.nabo_hip_batch_plan_launch(
            self._h, _stream_handle(stream, self.points.device)))
        return self.indices, self.dists2

    def status(self):
        """(complete, levels): waits for the last launch; complete is False when
        its build needed more levels than recorded (results not written)."""
        done, lv = C.c_int32(0), C.c_int32(0)
        _check(self._lib, self._lib.nabo_hip_batch_plan_status(self._h, C.byref(done),
                                                               C.byref(lv)))
        return bool(done.value), lv.value

    def close(self):
        if self._h:
            self._lib.nabo_hip_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServedRequest:
    """A submitted request: holds its query and output arrays until waited."""

    def __init__(self, server, rid, query, indices, dists2):
        self.server, self.id = server, rid
        self.query, self.indices, self.dists2 = query, indices, dists2
        self.latency_ms = None

    def wait(self):
        """Blocks until done; returns (indices, dists2) of shape (m, k)."""
        if self.latency_ms is None:
            lat = C.c_double(0)
            _check(self.server._lib, self.server._lib.nabo_hip_server_wait(
                self.server._h, self.id, C.byref(lat)))
            self.latency_ms = lat.value
        return self.indices, self.dists2


class Server:
    """Served kNN over a pool of resident kd-trees (nabo_hip_server_*).

    ``budget_bytes``: HBM target for resident trees and request buffers (0:
    unlimited); ``slots``: concurrent streams. Register clouds with
    add_cloud, then submit requests from any thread and wait on them.
    """

    def __init__(self, budget_bytes: int = 0, slots: int = 4, max_merge_queries: int = 1 << 22,
                 device: int = -1, host_cache_bytes: int = 0):
        self._lib = load_library()
        self._h = None
        cfg = ServerConfig()
        self._lib.nabo_hip_server_config_default(C.byref(cfg))
        cfg.hbm_budget_bytes = int(budget_bytes)
        cfg.slots = int(slots)
        cfg.max_merge_queries = int(max_merge_queries)
        cfg.device = int(device)
        cfg.host_cache_bytes = int(host_cache_bytes)
        h = C.c_void_p()
        _check(self._lib, self._lib.nabo_hip_server_create(C.byref(cfg), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            _check(self._lib, self._lib.nabo_hip_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_cloud(self, cloud, dim: int = 2**31 - 1, params: dict | None = None) -> int:
        kw = {_PARAM_KEYS[key]: val for key, val in (params or {}).items()}
        p = make_params(**kw)
        pts = _as_points(cloud, "cloud")
        n, rows = pts.shape
        cid = C.c_int32(-1)
        _check(self._lib, self._lib.nabo_hip_server_add_cloud(self._h, _ptr(pts), n, rows, rows,
                                                               int(dim), C.byref(p),
                                                               C.byref(cid)))
        return cid.value

    def submit(self, cloud_id: int, query, k: int = 1, epsilon: float = 0.0,
               optionFlags: int = 0, maxRadius: float = np.inf) -> ServedRequest:
        q = _as_points(query, "query")
        m, qrows = q.shape
        idx = np.empty((m, k), np.int32)
        d2 = np.empty((m, k), np.float32)
        rid = C.c_int64(0)
        _check(self._lib, self._lib.nabo_hip_server_submit(
            self._h, int(cloud_id), _ptr(q), qrows, qrows, m, int(k), float(epsilon),
            int(optionFlags), float(maxRadius), _ptr(idx), _ptr(d2), C.byref(rid)))
        return ServedRequest(self, rid.value, q, idx, d2)

    def stats(self) -> dict:
        st = ServerStats()
        _check(self._lib, self._lib.nabo_hip_server_get_stats(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in ServerStats._fields_}
