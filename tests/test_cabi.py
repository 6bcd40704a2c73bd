"""CPU checks of the boundary: libnabo_hip.so loads and exports every
function include/nabo_hip.h declares, and the Python mirror's constants keep
the reference's values (nabo/nabo.h:290-312). No compute call here."""
import ctypes
import os
import re

import libnabo_amd as nabo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "nabo_hip.h"), encoding="utf-8").read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nabo_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(nabo.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(nabo.EXPORTED)


def test_enums_keep_reference_values():
    assert [int(t) for t in nabo.SearchType][:6] == [0, 1, 2, 3, 4, 5]
    assert nabo.SearchType.KDTREE_HIP == 6 and nabo.SearchType.BRUTE_FORCE_HIP == 7
    assert nabo.SearchOptionFlags.ALLOW_SELF_MATCH == 1
    assert nabo.SearchOptionFlags.SORT_RESULTS == 2
    assert nabo.CreationOptionFlags.TOUCH_STATISTICS == 1


def test_params_defaults():
    p = nabo.make_params()
    assert p.bucket_size == 8 and p.query_sort == 1 and p.device == -1
    assert ctypes.sizeof(nabo.Params) == 64


def test_last_error_is_empty_at_start():
    lib = nabo.load_library()
    assert lib.nabo_hip_last_error() in (b"", None)


def test_struct_sizes_match_the_header():
    # nabo_hip_server_config / _stats as laid out by the C compiler
    assert ctypes.sizeof(nabo.ServerConfig) == 64
    assert ctypes.sizeof(nabo.ServerStats) == 11 * 8


def test_deserialize_rejects_bad_buffers_before_any_device_call():
    lib = nabo.load_library()
    h = ctypes.c_void_p()
    short = (ctypes.c_uint8 * 16)()
    rc = lib.nabo_hip_deserialize(ctypes.cast(short, ctypes.c_void_p), 16, -1, ctypes.byref(h))
    assert rc == 1 and b"truncated" in lib.nabo_hip_last_error()
    junk = (ctypes.c_uint8 * 4096)(*([7] * 4096))
    rc = lib.nabo_hip_deserialize(ctypes.cast(junk, ctypes.c_void_p), 4096, -1, ctypes.byref(h))
    assert rc == 1 and b"not a serialized index" in lib.nabo_hip_last_error()
    assert not h.value
    size = ctypes.c_uint64(0)
    assert lib.nabo_hip_serialize(None, None, 0, ctypes.byref(size)) == 1


def test_heap_param_and_unknown_heap_rejected_before_any_device_call():
    import numpy as np

    assert nabo.make_params().heap == 0 and nabo.make_params(heap="stl").heap == 1
    lib = nabo.load_library()
    p = nabo.make_params()
    p.heap = 5
    cloud = np.zeros((10, 3), np.float32)
    h = ctypes.c_void_p()
    rc = lib.nabo_hip_create(cloud.ctypes.data_as(ctypes.c_void_p), 10, 3, 3, 3,
                             int(nabo.SearchType.KDTREE_HIP), 0, ctypes.byref(p), ctypes.byref(h))
    assert rc == 1 and b"Unknown heap type" in lib.nabo_hip_last_error()
    assert not h.value


def test_schedule_params_checked_before_any_device_call():
    import numpy as np

    lib = nabo.load_library()
    cloud = np.random.default_rng(0).random((50, 3)).astype(np.float32)
    for field, bad in (("wave_queries", 96), ("stack_place", 5), ("stl_log", 8)):
        p = nabo.make_params()
        setattr(p, field, bad)
        h = ctypes.c_void_p()
        rc = lib.nabo_hip_create(cloud.ctypes.data_as(ctypes.c_void_p), 50, 3, 3, 3, 6, 0,
                                 ctypes.byref(p), ctypes.byref(h))
        assert rc == 1, field
        assert field.split("_")[0] in lib.nabo_hip_last_error().decode()
    try:
        nabo.make_params(bogus=1)
    except nabo.NaboError as e:
        assert "bogus" in str(e)
    else:
        raise AssertionError("unknown parameter accepted")


def test_cloud_size_limits_checked_before_any_device_call():
    # The point count is checked before the cloud is read: a 10-point buffer
    # stands for clouds of 2^29 and 2^30 points.
    import numpy as np

    lib = nabo.load_library()
    p = nabo.make_params()
    cloud = np.zeros((10, 3), np.float32)
    for n, stype, text in ((1 << 29, int(nabo.SearchType.KDTREE_HIP), b"kd-tree clouds of 2^29"),
                           (1 << 30, int(nabo.SearchType.BRUTE_FORCE_HIP), b"clouds of 2^30")):
        h = ctypes.c_void_p()
        rc = lib.nabo_hip_create(cloud.ctypes.data_as(ctypes.c_void_p), n, 3, 3, 3, stype, 0,
                                 ctypes.byref(p), ctypes.byref(h))
        assert rc == 4 and text in lib.nabo_hip_last_error(), (n, lib.nabo_hip_last_error())
        assert not h.value
