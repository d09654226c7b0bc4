"""Modulator texture coordinates on the GPU (modulator_t::doMapping,
texture.cc:125-165; triangle_t::getSurface's uv and orco interpolation,
triangle.cc:138-167; meshObject_t::toObjectOrco, mesh.cc:510-514): the
closed-form wall of every texco, and GI / antialiasing / photon renders with
uv, orco points, smoothed normals and every texco against the oracle."""
import numpy as np
import pytest

import closed_form_scenes as CF
from oracle import pyoracle as po
from yafray_amd import descriptors as D
from yafray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _Y():
    from yafray_amd import lib as Y
    return Y


def gpu_render(desc, torch, build_kw=None):
    Y = _Y()
    sc = Y.Scene.from_desc(desc, **(build_kw or {}))
    out = torch.zeros(desc.cam.resx * desc.cam.resy * 3, dtype=torch.float32, device="cuda")
    st = sc.render(out)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(desc.cam.resy, desc.cam.resx, 3), st, sc


def rel_rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)) / max(np.sqrt(np.mean(b.astype(np.float64) ** 2)), 1e-30))


@pytest.mark.parametrize("kind", list(CF.TEXCO_KINDS))
def test_texco_wall_closed_form(torch_cuda, kind):
    d = CF.texco_wall(kind)
    img, _, _ = gpu_render(d, torch_cuda)
    exp, mask = CF.texco_wall_expected(kind)
    assert np.abs(img[mask] - exp[mask]).max() < 2e-5
    ref, _ = po.OracleScene.from_desc(d).render(d)
    assert np.mean(np.all(img == ref, axis=2)) > 0.999
    assert np.abs(img - ref).max() < 1e-5


def textured_room(resx=96, resy=72, samples=4, depth=3, aa=False):
    """The direct room with uv on the room box (planar in x, z), orco points
    on box 0 (its vertices scaled), box 1 without data (orco from its
    matrix), the sphere autoSmooth'ed; each of the four materials has a
    modulator of another texco, the sphere's two (reflect, normal), and a
    pathlight and a sun light the paths' vertices."""
    r = S.direct_room(resx, resy, subdiv=3, samples=samples)
    tex = [D.marble(color1=(0.15, 0.3, 0.2), color2=(0.9, 0.7, 0.4), turbulence=0.0),
           D.wood(color1=(0.3, 0.2, 0.1), color2=(0.8, 0.6, 0.5)),
           D.clouds(color1=(0.2, 0.2, 0.6), color2=(0.9, 0.9, 0.8))]
    mods = [D.modulator(0, D.YG_MOD_MIX, color=0.8, size=2.0, texco="orco"),      # box 0: orco points
            D.modulator(1, D.YG_MOD_MUL, color=0.7, texco="orco"),                # box 1: orco matrix
            D.modulator(2, D.YG_MOD_MIX, color=0.9, size=3.0, texco="uv"),        # room: uv
            D.modulator(0, D.YG_MOD_MIX, color=0.6, texco="window"),              # room: window
            D.modulator(1, D.YG_MOD_MIX, color=0.7, size=0.7, texco="reflect"),   # sphere
            D.modulator(2, D.YG_MOD_ADD, color=0.2, texco="normal")]              # sphere
    mats = [D.material((0.6, 0.5, 0.4), 0, 1), D.material((0.5, 0.6, 0.5), 1, 1),
            D.material((0.7, 0.7, 0.6), 2, 2), D.material((0.5, 0.5, 0.7), 4, 2)]
    objs = r.tri_obj
    texd = {}
    # room (object 2): u = (x + 3) / 6, v = (z) / 6 + y / 12 at each corner
    sel = np.nonzero(objs == 2)[0]
    corners = r.verts[r.idx[sel]]
    uv = np.stack([(corners[..., 0] + 3) / 6, corners[..., 2] / 6 + corners[..., 1] / 12], -1).reshape(-1, 6)
    flags = np.ones(len(sel), np.uint8)
    flags[::4] = 2  # every fourth face: vertex colours, no uv
    texd[2] = {"uv": uv.astype(np.float32), "flags": flags}
    orco = np.zeros_like(r.verts)
    sel0 = np.unique(r.idx[objs == 0])
    orco[sel0] = r.verts[sel0] * np.array([0.5, -0.25, 1.5]) + np.array([0.1, 0.2, 0.3])
    texd[0] = {"orco": True}
    lights = list(r.lights) + [D.pathlight(samples=samples, depth=depth, use_qmc=True, threshold=1e3),
                               D.sunlight((0.2, -0.3, 1.0), power=0.6)]
    opts = r.opts
    if aa:
        opts = D.render_opts(bias=1e-3, raydepth=3, aa_passes=2, aa_minsamples=2, aa_threshold=0.02,
                             aa_jitterfirst=True)
    return D.SceneDesc(r.verts, r.idx, r.tri_obj, r.tri_mat, r.obj_flags, r.cam, mats, mods, tex, lights, opts,
                       smooth={3: 180.0}, tex=texd, orco=orco)


@pytest.mark.parametrize("width", [2, 4, 8])
def test_textured_room_gi_matches_oracle(torch_cuda, width):
    """Camera hits, path vertices and the area light's and sun's fromLight
    read the modulators through every texco; GPU = oracle with the same ray
    and draw counts."""
    d = textured_room()
    img, st, _ = gpu_render(d, torch_cuda, build_kw={"width": width})
    ref, rs = po.OracleScene.from_desc(d).render(d)
    assert np.all(np.isfinite(img))
    assert rel_rmse(img, ref) <= 1e-5
    assert np.mean(np.all(img == ref, axis=2)) > 0.99
    assert (st.rays_camera, st.rays_extension, st.rays_shadow) == tuple(int(x) for x in rs[:3])
    assert st.lcg_draws == rs[4]
    # the coordinates matter: the same room with every texco global differs
    g = textured_room()
    for m in g.mods:
        m.texco = D.YG_TEXCO_GLOBAL
    flat, _, _ = gpu_render(g, torch_cuda)
    assert rel_rmse(img, flat) > 1e-2


def test_textured_room_antialiasing_matches_oracle(torch_cuda):
    """Antialiasing passes: the window coordinates follow each jittered
    sample's screen position (scene.cc:798-799)."""
    d = textured_room(resx=64, resy=48, samples=2, depth=2, aa=True)
    d.lights[0].use_in_indirect = 0  # antialiasing with area lights inside GI paths is refused
    img, st, _ = gpu_render(d, torch_cuda)
    ref, rs = po.OracleScene.from_desc(d).render(d)
    assert rel_rmse(img, ref) <= 1e-5
    assert np.mean(np.all(img == ref, axis=2)) > 0.99
    assert st.rays_camera == rs[0]


def test_textured_photon_field_matches_oracle(torch_cuda):
    """Photon bounces filter by the diffuse colour at the previous hit with
    eye = the new direction (photonlight.cc:237): normal and reflect
    coordinates on the photon field's spheres."""
    d = S.photon_field(n_side=4, subdiv=2, extent=10.0, resx=80, resy=60, photons=2 * 150 * 150, search=20,
                       nmat=4)
    d.texs = D.carray(D.yg_texture, [D.marble(color1=(0.2, 0.3, 0.2), color2=(0.9, 0.8, 0.6), turbulence=0.0)])
    d.mods = D.carray(D.yg_modulator, [D.modulator(0, D.YG_MOD_MIX, color=0.8, texco="normal"),
                                       D.modulator(0, D.YG_MOD_MUL, color=0.6, size=0.5, texco="reflect")])
    mats = list(d.mats)
    mats[1] = D.material(tuple(mats[1].color), 0, 1)
    mats[2] = D.material(tuple(mats[2].color), 1, 1)
    d.mats = D.carray(D.yg_material, mats)
    Y = _Y()
    sc = Y.Scene.from_desc(d)
    ps = sc.photon_build(0)
    oc = po.OracleScene.from_desc(d)
    os_ = oc.photon_build(d, 0)
    assert (ps.emitted, ps.stored, ps.clustered) == tuple(int(x) for x in os_[:3])
    buf = torch_cuda.zeros(ps.clustered * 9, dtype=torch_cuda.float32, device="cuda")
    assert sc.photon_read(buf) == ps.clustered
    g, c = buf.cpu().numpy().reshape(-1, 9), oc.photon_read()
    assert np.mean(np.all(g == c, axis=1)) > 0.99
    assert np.allclose(g, c, rtol=1e-5, atol=1e-6)


def test_cache_with_window_coordinates_is_refused(torch_cuda):
    d = CF.texco_wall("window")
    d.lights = D.carray(D.yg_light, list(d.lights) + [D.pathlight(samples=4, cache=True, threshold=1e3, use_qmc=True)])
    Y = _Y()
    sc = Y.Scene.from_desc(d)
    out = torch_cuda.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch_cuda.float32, device="cuda")
    with pytest.raises(RuntimeError, match="window"):
        sc.render(out)
