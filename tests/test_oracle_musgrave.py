"""Pins of the oracle's Musgrave multifractals (noise.cc:668-791) and of the
modulator's projection axes and texture matrix (texture.cc:122-180,
311-317), from the mathematics:

* at lattice points every generator's noise is 1/2 (the signed noise 0),
  and with lacunarity 2 the scaled points stay on the lattice, so each
  multifractal reduces to a closed form in offset, gain and
  w = lacunarity^-H: multifractal 1; heteroterrain
  offset * prod_{i=1}^{k-1} (1 + offset w^i); hybridmf with gain = 1/offset
  offset (1 + sum_{i=1}^{k-1} w^{i(i+1)/2}); ridgedmf with offset 1
  1 + sum_{i=1}^{k-1} min(gain, 1)^i w^i;
* off the lattice, with one or two octaves (and a half octave), each is a
  factorised expression of the signed noise s0 = s(p), s1 = s(2p) of the
  pinned generators;
* the modulator's texture point: axis swaps, "n" axes and an affine matrix
  move the texture exactly as the same texture evaluated at the moved
  point."""
import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import descriptors as D

EVAL_NOISE, EVAL_TEX_FLOAT = 0, 3


def tex_float(t, pts):
    return po.noise_eval(EVAL_TEX_FLOAT, t, np.asarray(pts, np.float32)).astype(np.float64)


def signed(kind, pts):
    t = D.musgrave()
    t.noise = kind
    return 2.0 * po.noise_eval(EVAL_NOISE, t, np.asarray(pts, np.float32)).astype(np.float64) - 1.0


LATTICE = np.array([[0, 0, 0], [3, -2, 7], [-11, 5, 1], [100, 200, -300]], np.float32)


@pytest.mark.parametrize("H,octaves", [(1.0, 4.0), (0.5, 3.0), (1.3, 5.0), (0.8, 2.5)])
def test_lattice_closed_forms(H, octaves):
    w = 2.0 ** -H
    k = int(np.floor(octaves))
    rmd = octaves - k
    # multifractal: every factor is 1 + pwr * 0
    t = D.musgrave(H=H, lacunarity=2.0, octaves=octaves, mtype="multifractal")
    assert np.all(tex_float(t, LATTICE) == 1.0)
    # heteroterrain
    off = 0.7
    exp = off
    for i in range(1, k):
        exp *= 1.0 + off * w ** i
    exp += rmd * off * w ** k * exp
    t = D.musgrave(H=H, lacunarity=2.0, octaves=octaves, mtype="heteroterrain", offset=off)
    assert np.allclose(tex_float(t, LATTICE), exp, rtol=2e-6)
    # hybridmf, gain = 1/offset: weights w^{i(i-1)/2}, terms offset w^{i(i+1)/2}
    off = 0.8
    exp = off * (1.0 + sum(w ** (i * (i + 1) / 2) for i in range(1, k)))
    if rmd:
        exp += rmd * off * w ** k
    t = D.musgrave(H=H, lacunarity=2.0, octaves=octaves, mtype="hybridmf", offset=off, gain=1.0 / off)
    assert np.allclose(tex_float(t, LATTICE), exp, rtol=2e-6)
    # ridgedmf, offset 1: signal_i = min(gain, 1)^i, terms w^i (no half octave)
    for gain in (0.5, 2.0):
        g = min(gain, 1.0)
        exp = 1.0 + sum(g ** i * w ** i for i in range(1, k))
        t = D.musgrave(H=H, lacunarity=2.0, octaves=octaves, mtype="ridgedmf", offset=1.0, gain=gain)
        assert np.allclose(tex_float(t, LATTICE), exp, rtol=2e-6), gain


@pytest.mark.parametrize("kind", [0, 1, 2])  # newPerlin_t, stdPerlin_t, blenderNoise_t
def test_one_and_two_octaves_factorise(kind):
    rng = np.random.default_rng(5)
    p = rng.uniform(-20, 20, (4000, 3)).astype(np.float32)
    s0 = signed(kind, p)
    s1 = signed(kind, (p * np.float32(2.0)).astype(np.float32))
    H, off, gain = 0.9, 0.6, 1.7
    w = 2.0 ** -H

    def tf(mtype, octaves, offset=off):
        t = D.musgrave(H=H, lacunarity=2.0, octaves=octaves, mtype=mtype, offset=offset, gain=gain, noise=kind)
        return tex_float(t, p)

    tol = dict(rtol=1e-5, atol=1e-6)
    assert np.allclose(tf("multifractal", 1.0), 1.0 + s0, **tol)
    assert np.allclose(tf("multifractal", 2.0), (1.0 + s0) * (1.0 + w * s1), **tol)
    assert np.allclose(tf("multifractal", 1.5), (1.0 + s0) * (1.0 + 0.5 * s1 * w), **tol)
    assert np.allclose(tf("heteroterrain", 1.0), off + s0, **tol)
    assert np.allclose(tf("heteroterrain", 2.0), (off + s0) * (1.0 + (s1 + off) * w), **tol)
    assert np.allclose(tf("heteroterrain", 1.5), (off + s0) * (1.0 + 0.5 * (s1 + off) * w), **tol)
    r0 = s0 + off
    wt = gain * r0
    hyb2 = np.where(wt > 0.001, r0 + np.minimum(wt, 1.0) * (s1 + off) * w, r0)
    assert np.allclose(tf("hybridmf", 1.0), r0, **tol)
    assert np.allclose(tf("hybridmf", 2.0), hyb2, **tol)
    assert np.allclose(tf("hybridmf", 1.5), r0 + 0.5 * (s1 + off) * w, **tol)
    ro = 1.0  # ridged offset: (1 - |s0|)^2 gain crosses 1 on part of the points
    sig0 = (ro - np.abs(s0)) ** 2
    sig1 = (ro - np.abs(s1)) ** 2 * np.clip(sig0 * gain, 0.0, 1.0)
    assert np.allclose(tf("ridgedmf", 1.0, ro), sig0, **tol)
    assert np.allclose(tf("ridgedmf", 2.0, ro), sig0 + sig1 * w, **tol)
    # the ridged clamp engages on part of the points, the hybrid one too
    assert 0.05 < np.mean(sig0 * gain > 1.0) < 0.95 and 0.05 < np.mean(wt > 1.0) < 0.95


def test_size_and_intensity():
    """getFloat = intensity * generator(size * p) (basictex.cc:677-680)."""
    rng = np.random.default_rng(2)
    p = rng.uniform(-5, 5, (500, 3)).astype(np.float32)
    for mt in ("fBm", "multifractal", "heteroterrain", "hybridmf", "ridgedmf"):
        a = tex_float(D.musgrave(octaves=3.0, size=2.5, intensity=0.4, mtype=mt, offset=0.9, gain=1.3), p)
        b = tex_float(D.musgrave(octaves=3.0, size=1.0, intensity=1.0, mtype=mt, offset=0.9, gain=1.3),
                      (p * np.float32(2.5)).astype(np.float32))
        assert np.allclose(a, 0.4 * b, rtol=1e-6, atol=1e-7), mt


def test_modulator_projection_and_matrix_move_the_texture():
    """A plane lit from above shows the modulated albedo at each hit point P;
    the same scene with projection axes and a texture matrix equals the
    texture evaluated at size * (M P)[proj] + ofs computed here."""
    import closed_form_scenes as CF
    t = D.musgrave(color1=(0.1, 0.2, 0.3), color2=(0.9, 0.8, 0.7), octaves=3.0, size=1.0, intensity=0.5,
                   mtype="ridgedmf", offset=0.8, gain=1.5)
    M = np.array([[0.8, -0.6, 0.0, 0.3], [0.6, 0.8, 0.0, -1.2], [0.0, 0.0, 1.0, 0.7]], np.float32)
    for proj, mtx in (((0, 0, 0), None), ((2, 1, -1), None), ((3, 1, 2), M), ((0, -1, 1), M)):
        m = D.modulator(0, D.YG_MOD_MIX, color=1.0, sizexyz=(0.7, 1.3, 0.9), ofs=(0.25, -0.5, 0.125), proj=proj,
                        mtx=mtx)
        d = CF.textured_plane([(D.YG_MOD_MIX, 1.0)], resx=48, resy=36)
        d.mods = D.carray(D.yg_modulator, [m])
        d.texs = D.carray(D.yg_texture, [t])
        img, _ = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat).render(d)
        # the hit points of the pixel centres (plane z = 0 under the camera)
        from bench import _camera_rays
        o, dirs = _camera_rays(d.cam, np.arange(48 * 36))
        P = (o + dirs * (-o[:, 2:3] / dirs[:, 2:3])).astype(np.float32)
        T = P.astype(np.float64) if mtx is None else P.astype(np.float64) @ M[:, :3].T.astype(np.float64) + M[:, 3]
        q = np.zeros_like(T)
        for a in range(3):
            src = np.zeros(len(T)) if proj[a] < 0 else T[:, a if proj[a] == 0 else proj[a] - 1]
            q[:, a] = (m.size[a] * (src + m.ofs[a])) if proj[a] >= 0 else m.size[a] * m.ofs[a]
        f = tex_float(t, q.astype(np.float32))
        exp = np.array(t.color1)[None, :] + f[:, None] * (np.array(t.color2) - np.array(t.color1))[None, :]
        # pixels whose computed hit point rounds like the renderer's (all but a few)
        close = np.all(np.abs(img.reshape(-1, 3) - exp) < 2e-3, axis=1)
        assert np.mean(close) > 0.99, (proj, np.mean(close))
