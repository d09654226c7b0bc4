"""The C++ host adapter maps the reference plugin parameters to the same
descriptors as yafray_amd.descriptors (no GPU needed)."""
import os
import subprocess

import pytest

from yafray_amd import descriptors as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def adapter_output(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("host") / "host_adapter_check")
    lib = os.path.join(ROOT, "yafray_amd")
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "yafray_amd", "host"),
           os.path.join(ROOT, "tests", "host_adapter_check.cpp"), "-L", lib, "-lyafgpu",
           "-Wl,-rpath," + lib, "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return [l.split() for l in r.stdout.strip().splitlines()]


def _f(x):
    return float(x)


def test_lights_sorted_and_defaults(adapter_output):
    lights = [l for l in adapter_output if l[0] == "light"]
    # alphabetical: area, path, photons, sun
    assert [int(l[1]) for l in lights] == [D.YG_LIGHT_AREA, D.YG_LIGHT_PATH, D.YG_LIGHT_PHOTON, D.YG_LIGHT_SUN]
    area, path, ph, sun = lights
    ref = D.pathlight(samples=4, use_qmc=True, threshold=1000.0)
    assert int(path[12]) == ref.samples and int(path[13]) == ref.depth and _f(path[15]) == pytest.approx(1000.0)
    assert int(path[3]) == 0  # pathlight is not used in indirect light
    r = D.photonlight(photons=3998792, depth=1, mindepth=0, fixedradius=5.0, use_qmc=True)
    assert int(ph[19]) == r.photons and int(ph[20]) == r.search and int(ph[21]) == r.mindepth
    assert _f(ph[18]) == pytest.approx(r.angle) and _f(ph[22]) == pytest.approx(r.bias)
    assert _f(ph[24]) == pytest.approx(r.cluster)  # cluster defaults to fixedradius / 4
    s = D.sunlight((1, 0.5, 1.5), power=3.0)
    assert _f(sun[7]) == pytest.approx(s.power) and int(sun[11]) == 1
    assert int(area[12]) == 16 and _f(area[7]) == pytest.approx(10.0)


def test_shader_modulator_texture(adapter_output):
    mats = [l for l in adapter_output if l[0] == "mat"]
    mods = [l for l in adapter_output if l[0] == "mod"]
    texs = [l for l in adapter_output if l[0] == "tex"]
    assert mats == [["mat", "0.5", "0.4", "0.3", "0", "1"]]
    assert int(mods[0][2]) == D.YG_MOD_MUL and _f(mods[0][3]) == pytest.approx(0.6)
    assert [_f(x) for x in mods[0][4:7]] == [3.0, 3.0, 3.0]
    t = D.musgrave(octaves=4.0, intensity=0.5)
    assert int(texs[0][1]) == D.YG_TEX_MUSGRAVE
    assert [_f(x) for x in texs[0][2:6]] == [t.H, t.lacunarity, t.octaves, t.iscale]


def test_unsupported_is_refused(adapter_output):
    assert [l for l in adapter_output if l[0] == "refused"] == [["refused", str(D_ERR_UNSUPPORTED)]]


def test_cache_parameters(adapter_output):
    """pathLight_t::factory with cache on reads cache_size, shadow_threshold
    and search, raising search to 3 (pathlight.cc:785-797)."""
    row = [l for l in adapter_output if l[0] == "pcache"][0]
    assert row[1:] == ["0", "1", "0.02", "0.25", "3"]


D_ERR_UNSUPPORTED = -4


def test_texco_names(adapter_output):
    """The modulator's texco name maps as modulator_t::string2texcotype
    (texture.h:135-149): unknown names are uv; without the parameter the
    factory's default "orco" (basicshaders.cc:270)."""
    got = {l[1]: (int(l[2]), int(l[3])) for l in adapter_output if l[0] == "texco"}
    assert got == {"global": (0, D.YG_TEXCO_GLOBAL), "orco": (0, D.YG_TEXCO_ORCO), "uv": (0, D.YG_TEXCO_UV),
                   "window": (0, D.YG_TEXCO_WIN), "normal": (0, D.YG_TEXCO_NOR), "reflect": (0, D.YG_TEXCO_REFL),
                   "sticky": (0, D.YG_TEXCO_UV), "-": (0, D.YG_TEXCO_ORCO)}


def test_render_options_match_python_spec(adapter_output):
    rows = [l for l in adapter_output if l[0] == "ropts"]
    specs = [D.render_opts(),
             D.render_opts(bias=0.002, aa_passes=2, aa_minsamples=0, aa_pixelwidth=2.5, aa_threshold=0.03,
                           aa_jitterfirst=True),
             D.render_opts(aa_passes=2, aa_threshold=0.1, aa_pixelwidth=0.5)]
    assert len(rows) == len(specs)
    for r, o in zip(rows, specs):
        assert _f(r[1]) == pytest.approx(o.self_bias) and int(r[2]) == o.raydepth
        assert (int(r[3]), int(r[4]), int(r[7])) == (o.aa_passes, o.aa_minsamples, o.aa_jitterfirst)
        assert _f(r[5]) == pytest.approx(o.aa_pixelwidth) and _f(r[6]) == pytest.approx(o.aa_threshold)


def test_noise_type_names(adapter_output):
    """blender / stdperlin select those generators; voronoi and cellnoise are
    refused; any other name (or none) is newPerlin_t, as in newNoise."""
    got = {l[1]: (int(l[2]), int(l[3])) for l in adapter_output if l[0] == "noise"}
    assert got["blender"] == (0, 2) and got["stdperlin"] == (0, 1)
    assert got["newperlin"] == got["-"] == got["anything"] == (0, 0)
    assert got["voronoi_f1"][0] == D_ERR_UNSUPPORTED and got["cellnoise"][0] == D_ERR_UNSUPPORTED


def test_adaptive_repetitions_are_reported(adapter_output):
    """A render whose pathlight asked for adaptive extra repetitions
    (pathlight.cc:349-374), which the GPU path counts but does not run, is
    reported by the command-line renderer; a render without them is not."""
    w0 = [l for l in adapter_output if l[0] == "warn0"][0]
    w1 = [l for l in adapter_output if l[0] == "warn1"][0]
    assert w0 == ["warn0", "0"]
    assert w1 == ["warn1", "1", "1"]


def test_musgrave_types_offset_gain(adapter_output):
    """The Musgrave "type" names select the generators as textureMusgrave_t's
    ctor does (basictex.cc:657-672); any other name, or none, is fBm;
    offset and gain are read."""
    got = {l[1]: (int(l[2]), int(l[3]), _f(l[4]), _f(l[5])) for l in adapter_output if l[0] == "mtype"}
    assert got["fBm"][:2] == got["-"][:2] == got["other"][:2] == (0, D.YG_MUSGRAVE_FBM)
    assert got["multifractal"][1] == D.YG_MUSGRAVE_MULTIFRACTAL
    assert got["heteroterrain"][1] == D.YG_MUSGRAVE_HETEROTERRAIN
    assert got["hybridmf"][1] == D.YG_MUSGRAVE_HYBRIDMF and got["ridgedmf"][1] == D.YG_MUSGRAVE_RIDGEDMF
    assert all(v[2] == pytest.approx(0.7) and v[3] == pytest.approx(2.0) for v in got.values())


def test_modulator_projection_and_matrix(adapter_output):
    """proj_x/y/z follow "nxyz".find (texture.h:166-173): z -> 3, n -> none
    (-1), x -> 1; m03 / m11 land in the 3x4 texture matrix; without them
    the projection is the identity and there is no matrix."""
    mm = [l for l in adapter_output if l[0] == "modmap"][0]
    assert [int(x) for x in mm[1:6]] == [0, 3, -1, 1, 1]
    assert [_f(x) for x in mm[6:9]] == [0.25, 2.0, 1.0]
    md = [l for l in adapter_output if l[0] == "moddef"][0]
    assert [int(x) for x in md[1:6]] == [0, 1, 2, 3, 0]
