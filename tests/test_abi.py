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
    assert L.yg_abi_version() == 14
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


STRUCTS = ["yg_ray", "yg_hit", "yg_camera", "yg_texture", "yg_modulator", "yg_material", "yg_light",
           "yg_render_opts", "yg_render_stats", "yg_build_opts", "yg_tune_opts", "yg_tune_result",
           "yg_photon_stats"]


def test_ctypes_layouts_match_header(tmp_path):
    """Every ABI struct has the same size and field offsets in the ctypes
    mirror as a C compiler gives it from include/yafgpu.h."""
    import subprocess
    from yafray_amd import descriptors as D
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "yafgpu.h"', 'int main(void) {']
    for s in STRUCTS:
        lines.append('printf("%s sizeof() %%zu\\n", sizeof(%s));' % (s, s))
        for f, _ in getattr(D, s)._fields_:
            # (a trailing underscore marks a C field named by a Python keyword)
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (s, f, s, f.rstrip("_")))
    lines.append('return 0; }')
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for s in STRUCTS:
        cls = getattr(D, s)
        assert got[(s, "sizeof()")] == ctypes.sizeof(cls), s
        for f, _ in cls._fields_:
            assert got[(s, f)] == getattr(cls, f).offset, (s, f)


def test_autotune_argument_errors_without_gpu():
    from yafray_amd import descriptors as D
    from yafray_amd import lib as Y
    L = Y.lib()
    r = D.yg_tune_result()
    assert L.yg_autotune(None, None, None, None, ctypes.byref(r), None) == -1
    h = ctypes.c_void_p()
    assert L.yg_scene_create(ctypes.byref(h)) == 0
    o = D.yg_render_opts()
    dummy = ctypes.c_void_p(16)  # never dereferenced: the masks are checked first
    for widths, leaves, bins, rays, compacts, msg in ((1 << 3, 0, 0, 0, 0, "widths"), (0, 1 << 9, 0, 0, 0, "leaves"),
                                                      (0, 0, 1 << 7, 0, 0, "bins"), (0, 0, 0, 1 << 3, 0, "rays"),
                                                      (0, 0, 0, 0, 1 << 5, "compacts")):
        t = D.yg_tune_opts(widths=widths, leaves=leaves, bins=bins, rays=rays, compacts=compacts)
        assert L.yg_autotune(h, ctypes.byref(o), dummy, ctypes.byref(t), ctypes.byref(r), None) == -1
        assert msg in Y.last_error()
    assert L.yg_scene_destroy(h) == 0


def test_build_launch_settings_are_checked_without_gpu():
    """yg_build refuses a width, rays_per_lane, compact_below or builder
    outside their ranges before anything reaches the device."""
    from yafray_amd import descriptors as D
    from yafray_amd import lib as Y
    L = Y.lib()
    h = ctypes.c_void_p()
    assert L.yg_scene_create(ctypes.byref(h)) == 0
    for kw, msg in ((dict(width=3), "width"), (dict(rays_per_lane=3), "rays_per_lane"),
                    (dict(compact_below=1.5), "compact_below"), (dict(compact_below=-0.5), "compact_below"),
                    (dict(builder=3), "builder")):
        o = D.yg_build_opts(**kw)
        assert L.yg_build(h, ctypes.byref(o)) == -1
        assert msg in Y.last_error()
    assert L.yg_scene_destroy(h) == 0


def test_render_flags_match_header():
    import re
    from yafray_amd import descriptors as D
    hdr = open(os.path.join(ROOT, "include", "yafgpu.h")).read()
    flags = dict(re.findall(r"#define (YG_RENDER_\w+) (\d+)", hdr))
    assert flags and all(getattr(D, k) == int(v) for k, v in flags.items()), flags
