"""Pins of the oracle's photonSampler_t (pathtools.cc:80-266), the pathlight's
hemisphere sampler when a global photon map is published and samples > 96
(getSampler, pathlight.cc:725-738), from hand-worked cases and the
mathematics:

* grid 36 gives 4 x 8 sectors and a 96-photon gather;
* no photon facing N: every sector's energy difference is 0, so each starts
  with one sample and the remaining samples - 32 go round the sectors in
  order (the `noway` loop): 100 samples give 4, 4, 4, 4 and then 3s;
* photons arriving along N all fall in sector (0, 2) (sine 0, azimuth
  acos 0 = pi / 2 in the third pi / 4 slice); its energy E differs by E from
  its 5 neighbours (0..1 x 1..3), the others by 0, so each of those 6 gets
  (int)(68 / 6 + 0.5) = 11 extra samples, and the 2 left over go to the first
  two sectors holding more than one: (0, 1) and (0, 2) end at 13;
* weight = max hits / hits and multi = 1 / (max hits x sectors);
* the estimator is unbiased: under an open sky of colour B every level-0 ray
  escapes, weight x 2 z1 x cos theta averages the integral of 2 z1 sqrt(1 -
  z1^2) over z1 in [0, 1], i.e. 2/3, so the frame's mean is albedo x B x 2/3
  x power (to the stratified quadrature's error), whatever the plan."""
import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import descriptors as D


def floor_scene(samples=128, photons=20000, bg=(0.8, 0.6, 0.4), res=24):
    E = 5000.0
    v = np.array([[-E, -E, 0], [E, -E, 0], [E, E, 0], [-E, E, 0]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    L = D.arealight((-0.5, -0.5, 2), (-0.5, 0.5, 2), (0.5, 0.5, 2), (0.5, -0.5, 2), samples=1, power=10.0,
                    dummy=True)
    G = D.globalphotonlight(radius=0.5, depth=0, photons=photons, search=50)
    P = D.pathlight(samples=samples, depth=1, use_qmc=True, threshold=1e6)
    cam = D.camera((0, 0, 5), (0, 0, 0), (0, 1, 5), res, res, 1.0, 1.0)
    return D.SceneDesc(v, f, [0, 0], [0, 0], [7], cam, [D.material((0.5, 0.5, 0.5))], [], [], [L, G, P],
                       D.render_opts(bias=1e-3, background=bg))


def _loaded(recs):
    d = floor_scene()
    oc = po.OracleScene.from_desc(d)
    oc.gphoton_build(d, 1)
    oc.gphoton_load_photons(np.asarray(recs, np.uint32).reshape(-1, 5))
    return oc


def _rec(pos, th, ph, rgbe=(128, 128, 128, 129)):
    r = np.zeros(5, np.uint32)
    r[:3] = np.array(pos, np.float32).view(np.uint32)
    r[3] = th | (ph << 8)
    r[4] = rgbe[0] | (rgbe[1] << 8) | (rgbe[2] << 16) | (rgbe[3] << 24)
    return r


P0, N0, RU, RV = (0, 0, 0), (0, 0, 1), (1, 0, 0), (0, 1, 0)


def test_plan_without_photons_round_robin():
    # photons pointing down (theta byte 254: z < 0) never face N
    oc = _loaded([_rec((0.1 * k, 0, 0), 254, 0) for k in range(5)])
    hits, w, multi, nf = oc.psamp_plan(P0, N0, RU, RV, 100)
    assert nf == 0 and len(hits) == 32
    assert list(hits) == [4] * 4 + [3] * 28
    assert np.allclose(w, 4.0 / hits) and np.isclose(multi, 1 / (4 * 32))
    hits, w, multi, _ = oc.psamp_plan(P0, N0, RU, RV, 97)
    assert list(hits) == [4] + [3] * 31 and np.isclose(multi, 1 / (4 * 32))  # 65 = 2 x 32 + 1


def test_plan_with_photons_along_n():
    oc = _loaded([_rec((0.01 * k, 0.02, 0), 0, 0) for k in range(10)])
    hits, w, multi, nf = oc.psamp_plan(P0, N0, RU, RV, 100)
    assert nf == 10
    exp = np.ones((4, 8), int)
    for i in (0, 1):
        for j in (1, 2, 3):
            exp[i, j] = 12
    exp[0, 1] = exp[0, 2] = 13
    assert np.array_equal(hits.reshape(4, 8), exp) and hits.sum() == 100
    assert np.allclose(w, 13.0 / hits) and np.isclose(multi, 1 / (13 * 32))
    # beyond the radius (0.5) nothing is found
    hits2, _, _, nf2 = oc.psamp_plan((3, 3, 0), N0, RU, RV, 100)
    assert nf2 == 0 and list(hits2) == [4] * 4 + [3] * 28


def test_gather_keeps_the_96_nearest():
    rng = np.random.default_rng(5)
    pts = rng.uniform(-0.3, 0.3, (400, 2))
    oc = _loaded([_rec((x, y, 0), 0, 0) for x, y in pts])
    _, _, _, nf = oc.psamp_plan(P0, N0, RU, RV, 128)
    assert nf == 96


@pytest.mark.parametrize("samples", [128, 200])
def test_open_sky_estimate_is_unbiased(samples):
    d = floor_scene(samples=samples)
    oc = po.OracleScene.from_desc(d)
    oc.gphoton_build(d, 1)
    assert len(oc.gphoton_read_photons()) > 10000
    img, st = oc.render(d)
    exp = 0.5 * np.array([0.8, 0.6, 0.4]) * (2.0 / 3.0)
    mean = img.reshape(-1, 3).mean(axis=0)
    assert np.allclose(mean, exp, rtol=5e-3)
    # every pixel's own estimate is close too (stratified over 32 sectors)
    assert np.max(np.abs(img.reshape(-1, 3) / exp - 1)) < 0.05
    # the Halton sampler (96 samples: no photon sampler) has the same mean
    d96 = floor_scene(samples=96)
    oc96 = po.OracleScene.from_desc(d96)
    oc96.gphoton_build(d96, 1)
    img96, _ = oc96.render(d96)
    assert np.allclose(img96.reshape(-1, 3).mean(axis=0), exp, rtol=5e-3)
    assert not np.allclose(img96, img)
