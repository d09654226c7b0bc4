"""photonSampler_t (pathtools.cc:80-266) on the GPU against the oracle: the
pathlight's sampler when a global photon map is published and samples > 96
(getSampler, pathlight.cc:725-738).

The map itself is statistical (each photon draws from its own segment of the
Park-Miller sequence, see test_gpu_gphoton.py), so the oracle loads the GPU's
irradiance boxes and raw photons and renders the same frame.  Everything
after that is exact arithmetic shared by both: the areas' Halton starts
(ourRandomI() draws in area order), each shading point's 96-photon gather
(stateless: from the map's radius) and sector plan, the sector walk and the
weighted rays, and the multiplier."""
import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import descriptors as D
from yafray_amd import scenes as S

import test_oracle_psamp as TP

pytestmark = pytest.mark.gpu


def _Y():
    from yafray_amd import lib as Y
    return Y


def rel_rmse(a, b):
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - b) ** 2)) / np.sqrt(np.mean(np.asarray(b, np.float64) ** 2)))


def frame_and_map(d, torch, width=4, gl=1):
    sc = _Y().Scene.from_desc(d, width=width)
    sc.gphoton_build(gl)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    st = sc.render(out)
    torch.cuda.synchronize()
    n = sc.gphoton_read()
    boxes = torch.zeros(max(n, 1) * 15, dtype=torch.float32, device="cuda")
    sc.gphoton_read(boxes)
    nr = sc.gphoton_read_photons()
    recs = torch.zeros(max(nr, 1) * 5, dtype=torch.int32, device="cuda")
    sc.gphoton_read_photons(recs)
    torch.cuda.synchronize()
    img = out.cpu().numpy().reshape(d.cam.resy, d.cam.resx, 3)
    return (img, st, boxes.cpu().numpy()[: n * 15].reshape(n, 15),
            recs.cpu().numpy().view(np.uint32)[: nr * 5].reshape(nr, 5))


def oracle_frame(d, boxes, recs, gl=1):
    oc = po.OracleScene.from_desc(d)
    oc.gphoton_load(d, gl, boxes, stateless=True)
    oc.gphoton_load_photons(recs)
    return oc.render(d)


@pytest.mark.parametrize("width", [4, 8])
def test_photon_sampler_render_matches_oracle(torch_cuda, width):
    d = S.gphoton_room(48, 40, photons=60000, depth=2, search=50, radius=1.0, samples=128, path_depth=2,
                       sun=(0.3, -0.2, 1.0))
    img, st, boxes, recs = frame_and_map(d, torch_cuda, width)
    assert len(recs) > 20000 and np.all(np.isfinite(img)) and img.mean() > 0
    ref, rs = oracle_frame(d, boxes, recs)
    assert rel_rmse(img, ref) <= 1e-4
    assert np.mean(img == ref) > 0.999
    assert (st.rays_camera, st.rays_extension, st.rays_shadow) == tuple(int(x) for x in rs[:3])


def test_photon_sampler_open_sky_unbiased(torch_cuda):
    d = TP.floor_scene(samples=160)
    img, st, boxes, recs = frame_and_map(d, torch_cuda)
    exp = 0.5 * np.array([0.8, 0.6, 0.4]) * (2.0 / 3.0)
    assert np.allclose(img.reshape(-1, 3).mean(axis=0), exp, rtol=5e-3)
    ref, _ = oracle_frame(d, boxes, recs)
    assert rel_rmse(img, ref) <= 1e-5


def test_photon_sampler_refusals(torch_cuda):
    d = TP.floor_scene(samples=128)
    d.opts = D.render_opts(bias=1e-3, background=(0.8, 0.6, 0.4), aa_passes=1, aa_minsamples=2)
    sc = _Y().Scene.from_desc(d)
    sc.gphoton_build(1)
    out = torch_cuda.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch_cuda.float32, device="cuda")
    with pytest.raises(Exception, match="photonSampler_t"):
        sc.render(out)
