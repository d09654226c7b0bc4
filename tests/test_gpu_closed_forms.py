"""The closed forms of tests/test_oracle_closed_forms.py on the GPU (through
the C ABI): the pathlight past level 0, the photon bounce and cell merge,
the four modulator blends and the pinhole camera.  Each GPU result is also
compared with the oracle on the same inputs."""
import numpy as np
import pytest

import closed_form_scenes as CF
from oracle import pyoracle as po
from yafray_amd import descriptors as D
from test_oracle_closed_forms import _bounce_check, exact_halton

pytestmark = pytest.mark.gpu


def _Y():
    from yafray_amd import lib as Y
    return Y


def _render(desc, torch, **kw):
    sc = _Y().Scene.from_desc(desc, **kw)
    out = torch.zeros(desc.cam.resx * desc.cam.resy * 3, dtype=torch.float32, device="cuda")
    st = sc.render(out)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(desc.cam.resy, desc.cam.resx, 3), st


def _oracle(d):
    return po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)


@pytest.mark.parametrize("width", [2, 4, 8])
def test_two_bounce_pathlight_closed_form_gpu(torch_cuda, width):
    S_, rx, ry = 4, 80, 72
    d, sun = CF.two_bounce(samples=S_, depth=3, resx=rx, resy=ry)
    img, st = _render(d, torch_cuda, width=width)
    exp = CF.two_bounce_expected(S_, rx, ry, sun)
    assert np.max(np.abs(img - exp) / exp) < 3e-5
    assert st.rays_extension == 2 * S_ * st.rays_camera


@pytest.mark.parametrize("width", [2, 4, 8])
def test_level1_hits_on_a_sunlit_wall_gpu(torch_cuda, width):
    S_ = 4
    d, sun = CF.wall_bounce(samples=S_, resx=48, resy=40)
    img, st = _render(d, torch_cuda, width=width)
    exp, mask = CF.wall_bounce_expected(d, sun, samples=S_, halton=exact_halton)
    assert np.max((np.abs(img - exp) / exp)[mask]) < 2e-4
    ref, rs = _oracle(d).render(d)
    assert np.max(np.abs(img - ref) / ref) < 1e-5
    assert (st.rays_camera, st.rays_extension, st.rays_shadow) == tuple(int(x) for x in rs[:3])


def _gpu_photons(d, torch):
    sc = _Y().Scene.from_desc(d)
    ps = sc.photon_build(0)
    buf = torch.zeros(max(ps.clustered, 1) * 9, dtype=torch.float32, device="cuda")
    assert sc.photon_read(buf) == ps.clustered
    torch.cuda.synchronize()
    return ps, buf.cpu().numpy()[: ps.clustered * 9].reshape(-1, 9)


def test_photon_bounce_closed_form_gpu(torch_cuda):
    d = CF.photon_bounce()
    ps, ph = _gpu_photons(d, torch_cuda)
    assert ps.emitted == 60 * 60 and ps.stored == ps.emitted and ps.clustered == ps.emitted
    _bounce_check(ph, int(ps.emitted))
    oc = _oracle(d)
    oc.photon_build(d, 0)
    assert np.allclose(ph, oc.photon_read(), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("cell", [100.0, 0.05])
def test_cluster_merge_gpu(torch_cuda, cell):
    at = (5.013, 5.027)
    _, marks = _gpu_photons(CF.photon_spot_plane(cluster=1e-6, at=at), torch_cuda)
    ps, got = _gpu_photons(CF.photon_spot_plane(cluster=cell, at=at), torch_cuda)
    exp = CF.merge_expected(marks, cell)
    assert len(got) == len(exp) == ps.clustered
    assert np.allclose(got[:, 0:3], exp[:, 0:3], atol=2e-5)
    assert np.allclose(got[:, 3:6], exp[:, 3:6], rtol=1e-5, atol=1e-6)
    assert np.allclose(got[:, 6:9], exp[:, 6:9], rtol=2e-4)


@pytest.mark.parametrize("mode", [D.YG_MOD_MIX, D.YG_MOD_MUL, D.YG_MOD_ADD, D.YG_MOD_SUB])
def test_modulated_plane_render_gpu(torch_cuda, mode):
    """Each modulator mode on the device equals the oracle's render bit for
    bit up to transcendental rounding, alone and chained after another mode."""
    for seq in ([(mode, 0.6)], [(D.YG_MOD_ADD, 0.3), (mode, 0.8)], [(mode, 1.4)]):
        d = CF.textured_plane(seq, resx=64, resy=48)
        img, _ = _render(d, torch_cuda)
        ref, _ = _oracle(d).render(d)
        assert np.allclose(img, ref, rtol=1e-5, atol=1e-6), seq
        assert np.mean(img == ref) > 0.99


def test_camera_pinhole_projection_gpu(torch_cuda):
    for aspect, focal, res in ((1.3, 1.7, (96, 60)), (1.0, 1.6, (50, 70))):
        kw = dict(resx=res[0], resy=res[1], aspect=aspect, focal=focal)
        d = CF.camera_grid(**kw)
        img, _ = _render(d, torch_cuda)
        exp, mask = CF.camera_grid_expected(**kw)
        assert np.allclose(img[mask], exp[mask], atol=1e-6)


def test_modulator_projection_matrix_and_multifractals_gpu(torch_cuda):
    """Projection axes (swapped, "n"), a texture matrix and each Musgrave
    multifractal on the device equal the oracle's render."""
    M = np.array([[0.8, -0.6, 0.0, 0.3], [0.6, 0.8, 0.0, -1.2], [0.0, 0.0, 1.0, 0.7]], np.float32)
    for mt in ("multifractal", "heteroterrain", "hybridmf", "ridgedmf"):
        t = D.musgrave(color1=(0.1, 0.2, 0.3), color2=(0.9, 0.8, 0.7), octaves=4.5, size=0.8, intensity=0.6,
                       mtype=mt, offset=0.8, gain=1.5)
        for proj, mtx in (((0, 0, 0), None), ((3, 1, 2), M), ((0, -1, 1), M)):
            m = D.modulator(0, D.YG_MOD_MUL, color=0.9, sizexyz=(0.7, 1.3, 0.9), ofs=(0.25, -0.5, 0.125),
                            proj=proj, mtx=mtx)
            d = CF.textured_plane([(D.YG_MOD_MUL, 0.9)], resx=64, resy=48)
            d.mods = D.carray(D.yg_modulator, [m])
            d.texs = D.carray(D.yg_texture, [t])
            img, _ = _render(d, torch_cuda)
            ref, _ = _oracle(d).render(d)
            assert np.allclose(img, ref, rtol=1e-5, atol=1e-6), (mt, proj)
            assert np.mean(img == ref) > 0.99
