"""The pathlight's irradiance cache on the GPU (yg_render with cache on)
against the oracle's restatement of the reference (pathlight.cc:405-723,
lightcache.cc, cacheproxy.cc, the fake pass of scene.cc:559-583).

The fake pass takes the same samples (same visits, positions, normals,
first-hit distances and estimates: the takeSample paths are traced by the
GI levels with the per-area Halton indices), and the render pass gathers
them with the same radius chain per area and the same tie order, so images
agree to rounding.  Differences allowed by construction: the cache's tree
order is built without the reference's move-to-front of box lists, which
only changes the float rounding of the tree's split means."""
import numpy as np
import pytest

import closed_form_scenes as CF
from oracle import pyoracle as po
from test_oracle_cache import cache_room, cache_two_bounce
from yafray_amd import descriptors as D
from yafray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _Y():
    from yafray_amd import lib as Y
    return Y


def rel_rmse(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-30))


def gpu_render(d, torch, width=4):
    sc = _Y().Scene.from_desc(d, width=width)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    st = sc.render(out)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(d.cam.resy, d.cam.resx, 3), st, sc.cache_read()


def compare(d, torch, width=4, img_tol=1e-5, rank_min=0.999, exact_min=0.9999):
    img, st, gs = gpu_render(d, torch, width)
    oc = po.OracleScene.from_desc(d)
    ref, rs = oc.render(d, threads=8)
    os_, nnew = oc.cache_read()
    assert nnew == 0
    # the same samples in the same order: visits, positions, normals, distances, estimates
    assert len(gs) == len(os_) > 0
    assert np.array_equal(gs[:, 16], os_[:, 16])
    for a, b in ((0, 9), (9, 12), (12, 15)):
        np.testing.assert_allclose(gs[:, a:b], os_[:, a:b], rtol=1e-6, atol=1e-7)
    same_rank = float(np.mean(gs[:, 15] == os_[:, 15]))
    rel = rel_rmse(img, ref)
    exact = float(np.mean(np.all(img == ref, axis=2)))
    print("samples %d, rank equal %.5f, image rel rmse %.3g, exact pixels %.5f" % (len(gs), same_rank, rel, exact))
    assert same_rank >= rank_min
    assert rel <= img_tol and exact >= exact_min
    assert (st.rays_extension, st.rays_shadow) == (int(rs[1]), int(rs[2]))
    return img, ref, gs


@pytest.mark.parametrize("width", [2, 4, 8])
def test_two_bounce_cache_matches_oracle(torch_cuda, width):
    d, _ = cache_two_bounce(samples=6, resx=64, resy=48, cache_size=0.04)
    compare(d, torch_cuda, width)


def test_room_cache_matches_oracle(torch_cuda):
    d = cache_room(resx=160, resy=120, cache_size=0.02, samples=16, depth=2)
    compare(d, torch_cuda)


def test_terrain_cache_matches_oracle(torch_cuda):
    Y = _Y()
    d0 = S.terrain_gi(Y.device_heights, nx=256, ny=128, resx=320, resy=180, samples=8, depth=3)
    cached = D.pathlight(samples=8, depth=3, use_qmc=True, threshold=1e3, cache=True, cache_size=0.01, search=50)
    lights = [cached if L.type == D.YG_LIGHT_PATH else L for L in d0.lights]
    d = D.SceneDesc(d0.verts, d0.idx, d0.tri_obj, d0.tri_mat, d0.obj_flags, d0.cam, list(d0.mats), list(d0.mods),
                    list(d0.texs), lights, d0.opts)
    compare(d, torch_cuda, img_tol=1e-4, rank_min=0.99, exact_min=0.999)


def test_cache_refusals(torch_cuda):
    Y = _Y()
    d = cache_room(resx=32, resy=24)
    out = torch_cuda.zeros(32 * 24 * 3, dtype=torch_cuda.float32, device="cuda")
    L = D.yg_light.from_buffer_copy(d.lights[0])
    L.threshold = 0.5
    d1 = D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, d.cam, list(d.mats), [], [],
                     [L, d.lights[1]], d.opts)
    with pytest.raises(RuntimeError, match="refinement"):
        Y.Scene.from_desc(d1).render(out)
    o = D.yg_render_opts.from_buffer_copy(d.opts)
    o.aa_passes, o.aa_minsamples = 1, 1
    with pytest.raises(RuntimeError, match="antialiasing"):
        Y.Scene.from_desc(d).render(out, o)
    lights = list(d.lights) + [S.direct_room(8, 8, samples=4, subdiv=1).lights[0]]
    d3 = D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, d.cam, list(d.mats), [], [], lights, d.opts)
    with pytest.raises(RuntimeError, match="area lights"):
        Y.Scene.from_desc(d3).render(out)
    s1, s2 = Y.Scene.from_desc(d), Y.Scene.from_desc(d)
    with pytest.raises(RuntimeError, match="batched"):
        Y.render_batch([s1, s2], [out, out.clone()])
