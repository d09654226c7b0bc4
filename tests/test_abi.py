"""The C-ABI library loads and exports every entry point include/yafgpu.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "yafgpu.h")).read()
    return sorted(set(re.findall(r"^int\s+(yg_\w+)\s*\(", src, re.M)))


def test_header_declares_core_entry_points():
    names = declared()
    for n in ("yg_scene_create", "yg_mesh_add", "yg_build", "yg_trace_closest", "yg_trace_any",
              "yg_render", "yg_photon_build", "yg_photon_gather", "yg_last_error",
              "yg_scene_destroy"):
        assert n in names


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(os.path.join(ROOT, "yafray_amd", "libyafgpu.so"))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_error_path_without_gpu():
    """Argument errors are reported through the status code and yg_last_error."""
    from yafray_amd import lib as Y
    L = Y.lib()
    assert L.yg_abi_version() == 4
    assert L.yg_scene_create(None) == -1
    assert "null" in Y.last_error()
    h = ctypes.c_void_p()
    assert L.yg_scene_create(ctypes.byref(h)) == 0
    # index out of range is rejected before anything reaches the device
    v = (ctypes.c_float * 9)(0, 0, 0, 1, 0, 0, 0, 1, 0)
    f = (ctypes.c_uint32 * 3)(0, 1, 5)
    assert L.yg_mesh_add(h, v, 3, f, 1, None, 7, None) == -1
    assert "out of range" in Y.last_error()
    # queries before a build report the state error
    assert L.yg_trace_closest(h, None, 1, None, None) == -3
    assert L.yg_scene_destroy(h) == 0


def test_oracle_loads():
    from oracle import pyoracle as po
    assert po.lib().yo_next_prime(7) == 11
