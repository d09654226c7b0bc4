"""Area lights with psamples > 0 (guessLight, arealight.cc:176-200) and the
one-sample area light on the GPU, against the oracle and the closed forms
of tests/test_oracle_guess.py."""
import numpy as np
import pytest

import closed_form_scenes as CF
from oracle import pyoracle as po
from test_oracle_guess import _floor_points
from yafray_amd import descriptors as D
from yafray_amd import scenes as S

pytestmark = pytest.mark.gpu


def gpu_render(d, torch, width=4):
    from yafray_amd import lib as Y
    sc = Y.Scene.from_desc(d, width=width)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    st = sc.render(out)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(d.cam.resy, d.cam.resx, 3), st


def test_fully_lit_and_shadowed_floor(torch_cuda):
    d = CF.area_over_floor(samples=16, psamples=4)
    img, st = gpu_render(d, torch_cuda)
    L = d.lights[0]
    exp = CF.lambert_from_points(_floor_points(d), CF.area_grid_points(L), L, (0.7, 0.6, 0.5))
    np.testing.assert_allclose(img.reshape(-1, 3), exp, rtol=2e-4, atol=1e-7)
    assert st.lcg_draws == 0
    d2 = CF.area_over_floor(samples=16, psamples=3, slab=True)
    img2, st2 = gpu_render(d2, torch_cuda)
    assert np.all(img2 == 0) and st2.lcg_draws == 0


def test_single_sample_light_matches_oracle(torch_cuda):
    d = CF.area_over_floor(samples=1, psamples=0, resx=40, resy=30)
    img, st = gpu_render(d, torch_cuda)
    ref, rs = po.OracleScene.from_desc(d).render(d)
    assert np.array_equal(img, ref) and st.lcg_draws == int(rs[4])


@pytest.mark.parametrize("width", [2, 4, 8])
@pytest.mark.parametrize("psamples", [1, 4])
def test_room_psamples_matches_oracle(torch_cuda, width, psamples):
    """The room's icosphere and boxes shadow part of the floor: LIGHT,
    SHADOW and PENUMBRA verdicts all occur, and the penumbra points take
    their ourRandom() draws after every earlier point's, whatever verdict."""
    r = S.direct_room(96, 72, samples=16, subdiv=3)
    L = D.yg_light.from_buffer_copy(r.lights[0])
    L.psamples = psamples
    d = D.SceneDesc(r.verts, r.idx, r.tri_obj, r.tri_mat, r.obj_flags, r.cam, list(r.mats), [], [], [L], r.opts)
    img, st = gpu_render(d, torch_cuda, width)
    ref, rs = po.OracleScene.from_desc(d).render(d)
    rel = float(np.sqrt(np.mean((img.astype(np.float64) - ref) ** 2)) / np.sqrt(np.mean(ref.astype(np.float64) ** 2)))
    exact = float(np.mean(np.all(img == ref, axis=2)))
    print("psamples %d width %d: rel %.3g, exact %.4f, draws %d" % (psamples, width, rel, exact, st.lcg_draws))
    assert rel <= 1e-5 and exact >= 0.999
    assert st.lcg_draws == int(rs[4])
    assert (st.rays_camera, st.rays_shadow) == (int(rs[0]), int(rs[2]))
    # fewer draws than the penumbra loop everywhere
    L0 = D.yg_light.from_buffer_copy(L)
    L0.psamples = 0
    d0 = D.SceneDesc(r.verts, r.idx, r.tri_obj, r.tri_mat, r.obj_flags, r.cam, list(r.mats), [], [], [L0], r.opts)
    _, st0 = gpu_render(d0, torch_cuda, width)
    assert st.lcg_draws < st0.lcg_draws


def test_psamples_refusals(torch_cuda):
    from yafray_amd import lib as Y
    r = S.direct_room(32, 24, samples=16, subdiv=1)
    L = D.yg_light.from_buffer_copy(r.lights[0])
    L.psamples = 2
    o = D.yg_render_opts.from_buffer_copy(r.opts)
    o.aa_passes, o.aa_minsamples = 1, 1
    d = D.SceneDesc(r.verts, r.idx, r.tri_obj, r.tri_mat, r.obj_flags, r.cam, list(r.mats), [], [], [L], o)
    out = torch_cuda.zeros(32 * 24 * 3, dtype=torch_cuda.float32, device="cuda")
    with pytest.raises(RuntimeError, match="antialiasing"):
        Y.Scene.from_desc(d).render(out)
    lights = [L, D.pathlight(samples=2, depth=1, use_qmc=True, threshold=1e3)]
    d2 = D.SceneDesc(r.verts, r.idx, r.tri_obj, r.tri_mat, r.obj_flags, r.cam, list(r.mats), [], [], lights, r.opts)
    with pytest.raises(RuntimeError, match="inside GI paths"):
        Y.Scene.from_desc(d2).render(out)
