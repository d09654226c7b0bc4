"""Closed-form pins of the oracle functions that the open-sky plane, brute
force and the single-hit photon estimate leave open: the pathlight loop past
level 0, the photon bounce, preGathering's cell merge, the modulator blends
and the pinhole camera.  Expected values come from the reference's
definitions and from geometry (tests/closed_form_scenes.py), not from the
oracle's code."""
import math

import numpy as np
import pytest

import closed_form_scenes as CF
from oracle import pyoracle as po
from yafray_amd import descriptors as D


def _oracle(d):
    return po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)


# ------------------------------------------------ two-bounce pathlight --

@pytest.mark.parametrize("depth", [2, 3])
def test_two_bounce_pathlight_closed_form(depth):
    """pixel = rho_c P (1/S) sum_i sqrt(z1_i) rho_f ((n.s) E + B): the level-0
    cosine weights the floor's sun light and, through raycolour = rho_f, the
    sky seen by the level-1 ray (pathlight.cc:305-347).  Depth 3 changes
    nothing (every path escapes at level 1).  Both the FP32 and the FP64
    oracle."""
    S_, rx, ry = 4, 80, 72
    d, sun = CF.two_bounce(samples=S_, depth=depth, resx=rx, resy=ry)
    oc = _oracle(d)
    img, st = oc.render(d)
    exp = CF.two_bounce_expected(S_, rx, ry, sun)
    assert np.max(np.abs(img - exp) / exp) < 3e-5
    # 2 extension rays per sample; shadow rays: the ceiling's sun + one per floor vertex
    assert st[1] == 2 * S_ * st[0]
    assert st[2] == st[0] + S_ * st[0]
    f64 = oc.render_fp64(d)
    assert np.max(np.abs(f64 - exp) / exp) < 3e-5


def test_two_bounce_fails_without_the_sky_term():
    """The closed form is sensitive to each factor: dropping the level-1 sky
    term, the albedo product or the cosine moves pixels by far more than
    the tolerance (so the pin above is not vacuous)."""
    S_, rx, ry = 4, 40, 36
    d, sun = CF.two_bounce(samples=S_, resx=rx, resy=ry)
    img, _ = _oracle(d).render(d)
    exp = CF.two_bounce_expected(S_, rx, ry, sun)
    rf, rc, B = np.array(CF.RHO_F), np.array(CF.RHO_C), np.array(CF.SKY)
    k = CF.area_pixel_index(rx, ry)
    no_sky = np.zeros_like(exp)
    no_cos = np.zeros_like(exp)
    for i in range(ry):
        for j in range(rx):
            z = np.array([CF.radinv(int(k[i, j]) * S_ + s + 1, 2) for s in range(S_)])
            no_sky[i, j] = rc * 1.5 * np.mean(np.sqrt(z)) * rf * sun[2] * 2.0
            no_cos[i, j] = rc * 1.5 * rf * (sun[2] * 2.0 + B)
    for wrong in (no_sky, no_cos, exp / np.array(CF.RHO_F)):
        assert np.max(np.abs(img - wrong) / wrong) > 0.05


def exact_halton(base, n):
    return np.array([CF.radinv(i, base) for i in range(1, n + 1)])


def test_level1_hits_on_a_sunlit_wall():
    """Paths whose level-0 or level-1 ray reaches a sunlit wall: the light
    found at a level-1 hit is weighted by the level-0 cosine startray.N and
    by raycolour = the level-0 albedo (pathlight.cc:323-343); which samples
    reach the wall depends on the level-1 Halton dimensions 2 and 3 and on
    the face-forwarded normal HN.  The estimator is evaluated on the
    analytic geometry (CF.wall_bounce_expected), exact radical inverses
    standing in for the incremental Halton values (within 2e-5 of them)."""
    S_ = 4
    d, sun = CF.wall_bounce(samples=S_, resx=48, resy=40)
    oc = _oracle(d)
    img, st = oc.render(d)
    exp, mask = CF.wall_bounce_expected(d, sun, samples=S_, halton=exact_halton)
    assert mask.mean() > 0.9
    rel = np.abs(img - exp) / exp
    assert np.max(rel[mask]) < 2e-4, np.max(rel[mask])
    # some samples reach the wall at level 0 and more at level 1
    assert st[1] == 2 * S_ * st[0]
    f64 = oc.render_fp64(d)
    assert np.max((np.abs(f64 - exp) / exp)[mask]) < 2e-4


# -------------------------------------------------------- photon bounce --

def _bounce_check(ph, n_emitted, angle=30.0, power=3.0):
    """Stored marks of CF.photon_bounce: colour, count and the bounce
    directions against the Halton points of dimensions 2 and 3."""
    assert len(ph) == n_emitted
    col = CF.photon_power(angle, power) * np.array(CF.PH_COL) * np.array(CF.PH_RHO)
    assert np.allclose(ph[:, 6:9], col[None, :], rtol=2e-6)
    assert np.allclose(ph[:, 5], 2.0, atol=1e-5)          # all on the ceiling
    cos_t = -ph[:, 2].astype(np.float64)
    az = np.arctan2(ph[:, 1], ph[:, 0]).astype(np.float64)
    k = np.arange(1, n_emitted + 1)
    r1 = np.array([CF.radinv(int(i), 5) for i in k])
    r2 = np.array([CF.radinv(int(i), 7) for i in k])
    exp_cos = 1 - 0.95 * r2
    exp_az = 2 * math.pi * r1 + math.pi
    o, oe = np.argsort(cos_t), np.argsort(exp_cos)
    assert np.allclose(cos_t[o], exp_cos[oe], atol=3e-5)
    daz = np.angle(np.exp(1j * (az[o] - exp_az[oe])))
    # the azimuth is undefined for near-vertical bounces; elsewhere it must match
    steep = exp_cos[oe] > 0.9999
    assert np.max(np.abs(daz[~steep])) < 2e-3
    # cosines uniform on [0.05, 1]: mean 0.525 within the sampling error
    assert abs(np.mean(cos_t) - 0.525) < 3 * 0.95 / math.sqrt(12 * n_emitted)


def test_photon_bounce_closed_form():
    """shoot_photon_diffuse (photonlight.cc:203-244) with a bounce: every
    emitted photon is stored once, on the ceiling, with colour power x colour
    x 2 pi (1 - cos angle) x rho_floor, along a cone direction whose cosine is
    1 - 0.95 r2 and azimuth 2 pi r1 (+ pi, pointing back), r1 and r2 the
    base-5 and base-7 radical inverses of its bounce index."""
    d = CF.photon_bounce()
    oc = _oracle(d)
    st = oc.photon_build(d, 0)
    assert st[0] == 60 * 60 and st[1] == st[0] and st[2] == st[0]
    _bounce_check(oc.photon_read(), int(st[0]))


# -------------------------------------------------------- cluster merge --

def test_cluster_merge_one_cell():
    """All photons in one hash cell give one mark (preGathering,
    photonlight.cc:397-411): the mean of their positions, the normalised sum
    of their directions and the sum of their colours."""
    d0 = CF.photon_spot_plane(cluster=1e-6)
    oc = _oracle(d0)
    st0 = oc.photon_build(d0, 0)
    assert st0[1] == st0[2] == st0[0]
    marks = oc.photon_read()
    d1 = CF.photon_spot_plane(cluster=100.0)
    st1 = oc.photon_build(d1, 0)
    assert st1[1] == st0[1] and st1[2] == 1
    got = oc.photon_read()
    exp = CF.merge_expected(marks, 100.0)
    assert exp.shape == (1, 9)
    assert np.allclose(got[0, 0:3], exp[0, 0:3], atol=1e-5)
    assert np.allclose(got[0, 3:6], exp[0, 3:6], rtol=1e-5)
    assert np.allclose(got[0, 6:9], exp[0, 6:9], rtol=2e-4)


def test_cluster_merge_several_cells():
    """Cells of size 0.05 over the spot's footprint: marks are the per-cell
    merges in (x, y, z) cell order (the std::map order of hash3d_t)."""
    at = (5.013, 5.027)
    d0 = CF.photon_spot_plane(cluster=1e-6, at=at)
    oc = _oracle(d0)
    oc.photon_build(d0, 0)
    marks = oc.photon_read()
    cell = 0.05
    # no unmerged mark lies on a cell boundary (the assignment is unambiguous)
    q = marks[:, 3:5].astype(np.float64) / cell
    assert np.min(np.abs(q - np.round(q))) > 1e-4
    d1 = CF.photon_spot_plane(cluster=cell, at=at)
    st = oc.photon_build(d1, 0)
    got = oc.photon_read()
    exp = CF.merge_expected(marks, cell)
    assert len(got) == len(exp) == st[2] and 10 < len(exp) < len(marks)
    assert np.allclose(got[:, 0:3], exp[:, 0:3], atol=2e-5)
    assert np.allclose(got[:, 3:6], exp[:, 3:6], rtol=1e-5, atol=1e-6)
    assert np.allclose(got[:, 6:9], exp[:, 6:9], rtol=1e-4)


# ------------------------------------------------------------ modulators --

MODES = [D.YG_MOD_MIX, D.YG_MOD_MUL, D.YG_MOD_ADD, D.YG_MOD_SUB]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("factor", [0.0, 0.3, 1.0, 1.7])
def test_modulator_blend_closed_form(mode, factor):
    """getDiffuse through one modulator at marble points whose texture value
    f is 0, 1/2 and 1 (texture colour c1, (c1 + c2)/2, c2): MIX, MUL, ADD and
    SUB by the formulas of texture.cc:499-534, factor <= 0 leaving the
    colour alone and a factor above 1 taking the texture colour in MIX."""
    fvals = np.array([0.0, 0.5, 1.0, 0.25])
    pts = CF.marble_points(fvals)
    mats, mods, tex = CF.modulated_scene([(mode, factor)])
    d = D.SceneDesc(np.zeros((3, 3), np.float32), np.zeros((1, 3), np.uint32), [0], [0], [7],
                    D.camera((0, 0, 0), (0, 1, 0), (0, 0, 1), 4, 4), mats, mods, tex, [], D.render_opts())
    got = po.diffuse_at(d, 0, pts)
    c1, c2 = np.array([0.1, 0.2, 0.3]), np.array([0.9, 0.6, 0.5])
    for q, f in enumerate(fvals):
        tc = c1 + f * (c2 - c1)
        exp = CF.modulate_expected((0.4, 0.5, 0.6), tc, mode, factor)
        assert np.allclose(got[q], exp, atol=3e-6), (mode, factor, f, got[q], exp)


def test_modulators_apply_in_order():
    """Several modulators on one shader are applied in list order
    (basicshaders.cc:191-204): MUL then ADD differs from ADD then MUL."""
    pts = CF.marble_points([0.25, 0.75])
    c1, c2 = np.array([0.1, 0.2, 0.3]), np.array([0.9, 0.6, 0.5])
    for seq in ([(D.YG_MOD_MUL, 0.6), (D.YG_MOD_ADD, 0.4)], [(D.YG_MOD_ADD, 0.4), (D.YG_MOD_MUL, 0.6)],
                [(D.YG_MOD_SUB, 0.2), (D.YG_MOD_MIX, 0.5), (D.YG_MOD_MUL, 0.9)]):
        mats, mods, tex = CF.modulated_scene(seq)
        d = D.SceneDesc(np.zeros((3, 3), np.float32), np.zeros((1, 3), np.uint32), [0], [0], [7],
                        D.camera((0, 0, 0), (0, 1, 0), (0, 0, 1), 4, 4), mats, mods, tex, [], D.render_opts())
        got = po.diffuse_at(d, 0, pts)
        for q, f in enumerate([0.25, 0.75]):
            C = np.array([0.4, 0.5, 0.6])
            for mode, c in seq:
                C = CF.modulate_expected(C, c1 + f * (c2 - c1), mode, c)
            assert np.allclose(got[q], C, atol=3e-6)


@pytest.mark.parametrize("mode", MODES)
def test_modulated_plane_render(mode):
    """A plane under a sun straight above it renders its modulated albedo
    (sunlight.cc:35-57 with N.s = 1): the render's pixels equal getDiffuse at
    the camera rays' hit points."""
    d = CF.textured_plane([(mode, 0.6)], resx=32, resy=24)
    oc = _oracle(d)
    img, _ = oc.render(d)
    # hit points of the pixel-centre camera rays, from the FP64 render's camera
    # (pinned by test_camera_pinhole_projection below) intersected with z = 0
    cam = d.cam
    pos = np.array(cam.pos[:], np.float64)
    fwd = np.array(cam.look[:], np.float64) - pos
    up = np.array(cam.up[:], np.float64) - pos
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    upv /= np.linalg.norm(upv)
    fwd /= np.linalg.norm(fwd)
    j = (np.arange(cam.resx) + 0.5) / cam.resx - 0.5
    i = (np.arange(cam.resy) + 0.5) / cam.resy - 0.5
    asp = cam.aspect * cam.resy / cam.resx
    dirs = (right[None, None, :] * j[None, :, None] - upv[None, None, :] * asp * i[:, None, None]
            + fwd[None, None, :] * cam.focal)
    t = -pos[2] / dirs[:, :, 2]
    P = pos[None, None, :] + dirs * t[:, :, None]
    exp = po.diffuse_at(d, 0, P.reshape(-1, 3).astype(np.float32)).reshape(img.shape)
    assert np.allclose(img, exp, atol=2e-4)


# ----------------------------------------------------------------- camera --

def test_camera_pinhole_projection():
    """Each pixel's camera ray lands where a pinhole of focal length `focal`
    and an image plane one unit wide, aspect x resy / resx high, puts it
    (camera.cc:27-87 builds that plane; scene.cc:736-756 samples pixel
    centres, row 0 at the top): on a wall of colour-coded cells every pixel
    away from a cell edge shows the cell the projection predicts."""
    for aspect, focal, res in ((1.3, 1.7, (96, 60)), (1.0, 1.6, (50, 70))):
        kw = dict(resx=res[0], resy=res[1], aspect=aspect, focal=focal)
        d = CF.camera_grid(**kw)
        img, _ = _oracle(d).render(d)
        exp, mask = CF.camera_grid_expected(**kw)
        assert mask.mean() > 0.8
        assert np.allclose(img[mask], exp[mask], atol=1e-6)
        f64 = _oracle(d).render_fp64(d)
        assert np.allclose(f64[mask], exp[mask], atol=1e-6)
