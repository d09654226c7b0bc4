"""YafRay XML scene files and image files on the host side (no GPU): the
loader builds the same descriptors as the scene built through the C ABI,
malformed files are refused with a line number, and the targa / RGBE
writers produce the reference's byte layout and 8-bit / shared-exponent
values (targaIO.cc:83-131, color.cc:35-40, HDR_io.cc:76-160)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import descriptors as D
from yafray_amd import scenes as S

from xml_scene import read_hdr, read_tga, rgbe, scene_xml, tga_bytes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
D_ERR_ARG, D_ERR_UNSUPPORTED = -1, -4  # YG_ERR_ARG, YG_ERR_UNSUPPORTED


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("scene") / "scene_file_check")
    lib = os.path.join(ROOT, "yafray_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(lib, "host"),
                    os.path.join(ROOT, "tests", "scene_file_check.cpp"), "-L", lib, "-lyafgpu",
                    "-Wl,-rpath," + lib, "-Wl,-rpath,/opt/rocm/lib", "-o", exe], check=True, capture_output=True)
    return exe


def run(exe, *args):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return [l.split() for l in r.stdout.strip().splitlines()]


def test_loader_builds_the_same_scene(checker, tmp_path):
    d = S.terrain_gi(po.heightfield, nx=32, ny=16, resx=64, resy=40, samples=2, depth=3, nmat=8)
    d.opts = D.render_opts(bias=2e-3, raydepth=3, background=(0.4, 0.5, 0.7), aa_passes=1, aa_minsamples=3,
                           exposure=0.5, gamma=2.2)
    path = tmp_path / "t.xml"
    path.write_text(scene_xml(d, "t.tga"))
    out = run(checker, "describe", str(path))
    assert out[0] == ["load", "0"]
    lights = [l for l in out if l[0] == "light"]
    assert len(lights) == len(d.lights)
    for l, L in zip(lights, d.lights):
        assert int(l[1]) == L.type and int(l[2]) == L.use_in_render and int(l[3]) == L.use_in_indirect
        assert np.float32(float(l[4])) == np.float32(L.power)
        if L.type == D.YG_LIGHT_SUN:
            assert [np.float32(float(x)) for x in l[8:11]] == list(np.float32(L.dir))
        if L.type == D.YG_LIGHT_PATH:
            assert (int(l[12]), int(l[13]), int(l[14])) == (L.samples, L.depth, L.use_qmc)
    mats = [[np.float32(float(x)) for x in l[1:4]] for l in out if l[0] == "mat"]
    assert mats == [list(np.float32(m.color)) for m in d.mats]
    r = [l for l in out if l[0] == "render"][0]
    assert r[1:6] == ["cam", "64", "40", "t.tga", "tga"]
    assert int(r[6]) == 3 and np.float32(float(r[7])) == np.float32(2e-3)
    assert (int(r[8]), int(r[9]), int(r[10])) == (1, 3, 1)  # AA passes, samples, tone mapping for targa
    assert np.float32(float(r[11])) == np.float32(0.5) and np.float32(float(r[12])) == np.float32(2.2)
    path.write_text(scene_xml(d, "t.hdr", output_type="hdr"))
    r = [l for l in run(checker, "describe", str(path)) if l[0] == "render"][0]
    assert r[5] == "hdr" and int(r[10]) == 0  # no tone mapping for hdr (render.cc:1233-1237)


def test_malformed_files_are_refused(checker):
    out = run(checker, "errors")
    codes = [int(l[1]) for l in out]
    assert codes[:6] == [D_ERR_ARG] * 5 + [D_ERR_UNSUPPORTED]
    assert codes[6] == 0
    assert "line 3" in " ".join(out[1]) and "not quoted" in " ".join(out[1])
    assert "closes" in " ".join(out[0])



def test_image_writers(checker, tmp_path):
    tga, hdr = str(tmp_path / "i.tga"), str(tmp_path / "i.hdr")
    out = run(checker, "images", tga, hdr)
    assert out[0] == ["tga", "0"] and out[1] == ["hdr", "0"]
    rgb = np.array([float(l[0]) for l in out[2:]], np.float32).reshape(5, 37, 3)
    head, px = read_tga(tga)
    assert head == bytes([0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 37, 0, 5, 0, 0x18, 0x20])
    assert np.array_equal(px, tga_bytes(rgb))
    head, img = read_hdr(hdr)
    assert head.startswith(b"#?RADIANCE\n") and b"FORMAT=32-bit_rle_rgbe\n" in head
    assert np.array_equal(img, rgbe(rgb))


def test_unsupported_object_settings_are_refused(checker, tmp_path):
    """Object settings that change the reference's image and that the GPU
    path does not implement make the load fail with YG_ERR_UNSUPPORTED
    instead of being dropped: caustic settings (msin_object.cc:448-476 turn
    on obj->caustics).  Face uv, vertex colours and has_orco load (their
    texture coordinates are read by modulators), and so does a sphere object,
    which the reference's loader makes nothing of (render.cc:768-776).
    autosmooth below 1 degree is a no-op in the reference (mesh.cc:204) and
    loads."""
    d = S.terrain_gi(po.heightfield, nx=8, ny=6, resx=16, resy=12, samples=1, depth=1, nmat=2)
    base = scene_xml(d, "t.tga")
    assert '<mesh autosmooth="0">' in base and "<f a=" in base
    cases = {
        "caus_ior": base.replace('<object name="O0000"', '<object name="O0000" caus_IOR="1.5"', 1),
        "caus_rcolor": base.replace("<attributes></attributes><mesh",
                                    '<attributes><caus_rcolor r="1" g="1" b="1"/></attributes><mesh', 1),
        "caus_tcolor": base.replace("<attributes></attributes><mesh",
                                    '<attributes><caus_tcolor r="1" g="1" b="1"/></attributes><mesh', 1),
    }
    for name, text in cases.items():
        assert text != base, name
        path = tmp_path / ("%s.xml" % name)
        path.write_text(text)
        out = run(checker, "describe", str(path))
        assert out[0] == ["load", str(D_ERR_UNSUPPORTED)], name
    loads = {
        "uv": base.replace("<f a=", '<f u_a="0" v_a="0" u_b="1" v_b="0" u_c="0" v_c="1" a=', 1),
        "vcol": base.replace("<f a=", '<f vcol_a_r="1" vcol_a_g="1" vcol_a_b="1" a=', 1),
        "orco": base.replace('<mesh autosmooth="0">', '<mesh autosmooth="0" has_orco="on">', 1),
        "sphere": base.replace('<object name="O0000"',
                               '<object name="A_sphere" shader_name="M000"><attributes></attributes>'
                               '<sphere radius="1"><center x="0" y="0" z="0"/></sphere></object>'
                               '<object name="O0000"', 1),
    }
    for name, text in loads.items():
        assert text != base, name
        path = tmp_path / ("%s.xml" % name)
        path.write_text(text)
        assert run(checker, "describe", str(path))[0] == ["load", "0"], name
    for angle in ("0", "0.5"):
        path = tmp_path / "smooth.xml"
        path.write_text(base.replace('<mesh autosmooth="0">', '<mesh autosmooth="%s">' % angle))
        assert run(checker, "describe", str(path))[0] == ["load", "0"]
