"""Pins of the oracle's autoSmooth (mesh.cc:199-282) and of getSurface's
interpolated shading normal (triangle.cc:142-160), from geometry: a tent
roof whose two planes meet at a ridge, and a smoothed sphere."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import descriptors as D
from yafray_amd import scenes as S


def tent(n=4, h=0.5):
    """Grid over x in [-1, 1] (2n cells), y in [0, 1] (n cells), z = h (1 - |x|):
    two planes with unit normals (-h, 0, 1)/|.| (x < 0) and (h, 0, 1)/|.| (x > 0)."""
    xs = np.linspace(-1, 1, 2 * n + 1)
    ys = np.linspace(0, 1, n + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    v = np.stack([X.ravel(), Y.ravel(), h * (1 - np.abs(X.ravel()))], 1).astype(np.float32)
    nx = 2 * n + 1
    f = []
    for j in range(n):
        for i in range(2 * n):
            a, b, c, d = j * nx + i, j * nx + i + 1, (j + 1) * nx + i + 1, (j + 1) * nx + i
            f += [[a, b, c], [a, c, d]]
    nl = np.array([-h, 0, 1.0]) / math.hypot(h, 1)
    nr = np.array([h, 0, 1.0]) / math.hypot(h, 1)
    return v, np.array(f, np.uint32), nl, nr


def _expected_vertex_normals(v, f, nl, nr, smooth_across):
    """Per triangle corner: the plane normal, or at ridge vertices (x = 0) when
    both planes are smoothed together the normalised kl nl + kr nr, kl / kr
    the numbers of the vertex's triangles on each side."""
    side = np.where(v[f].mean(axis=1)[:, 0] < 0, 0, 1)   # plane of each triangle
    out = np.zeros((len(f), 3, 3))
    for t, tri in enumerate(f):
        for k, vi in enumerate(tri):
            own = nl if side[t] == 0 else nr
            if abs(v[vi, 0]) < 1e-6 and smooth_across:
                adj = np.nonzero(np.any(f == vi, axis=1))[0]
                kl, kr = np.sum(side[adj] == 0), np.sum(side[adj] == 1)
                s_ = kl * nl + kr * nr
                out[t, k] = s_ / np.linalg.norm(s_)
            else:
                out[t, k] = own
    return out


@pytest.mark.parametrize("angle,across", [(60.0, True), (200.0, True), (40.0, False)])
def test_tent_ridge_vertex_normals_and_interpolation(angle, across):
    """The planes' normals differ by 2 atan(h) = 53.1 degrees: a smoothing
    angle above that averages them at the ridge (weighted by how many
    triangles of each plane share the vertex), below it every corner keeps its
    face's normal.  The shading normal at barycentric weights (fa, fb, fc) is
    normalize(fa na + fb nb + fc nc)."""
    v, f, nl, nr = tent()
    oc = po.OracleScene(v, f, np.zeros(len(f), np.int32), np.zeros(len(f), np.int32))
    oc.autosmooth(0, angle)
    exp_vn = _expected_vertex_normals(v, f, nl, nr, across)
    rng = np.random.default_rng(3)
    w = rng.dirichlet([1, 1, 1], size=len(f))
    P = np.einsum("tk,tkc->tc", w, v[f].astype(np.float64)).astype(np.float32)
    got = oc.shading_normals(np.arange(len(f), dtype=np.int32), P)
    exp = np.einsum("tk,tkc->tc", w, exp_vn)
    exp /= np.linalg.norm(exp, axis=1, keepdims=True)
    assert np.allclose(got, exp, atol=2e-6)
    # triangles away from the ridge, and all of them without smoothing across:
    # the face normal exactly
    far = np.all(np.abs(v[f][:, :, 0]) > 1e-6, axis=1) if across else np.ones(len(f), bool)
    fn = np.where((v[f].mean(axis=1)[:, 0] < 0)[:, None], nl, nr)
    assert np.allclose(got[far], fn[far], atol=1e-6)


def test_below_one_degree_is_flat():
    """autoSmooth does nothing below 1 degree (mesh.cc:204): N = Ng exactly."""
    v, f, nl, nr = tent()
    oc = po.OracleScene(v, f, np.zeros(len(f), np.int32), np.zeros(len(f), np.int32))
    oc.autosmooth(0, 0.5)
    P = v[f].mean(axis=1).astype(np.float32)
    got = oc.shading_normals(np.arange(len(f), dtype=np.int32), P)
    n = np.cross(v[f][:, 1] - v[f][:, 0], v[f][:, 2] - v[f][:, 0])
    assert np.allclose(got, n / np.linalg.norm(n, axis=1, keepdims=True), atol=1e-7)


def sun_sphere(smooth, subdiv=4, res=96):
    """An icosphere of radius 1 at the origin seen from +y, a sun along
    direction s (power 1, albedo 1, no sky): each pixel is max(0, N.s) at
    its hit point."""
    sv, sf = S.icosphere(subdiv)
    s = np.array([0.3, -0.5, 0.8]); s /= np.linalg.norm(s)
    cam = D.camera((0, -4, 0), (0, 0, 0), (0, -4, 1), res, res, 1.0, 2.2)
    d = D.SceneDesc(sv.astype(np.float32), sf, np.zeros(len(sf), np.int32), np.zeros(len(sf), np.int32), [7], cam,
                    [D.material((1.0, 1.0, 1.0))], [], [], [D.sunlight(s, power=1.0)], D.render_opts(bias=1e-3),
                    smooth={0: 180.0} if smooth else None)
    return d, s


def test_smoothed_sphere_approaches_the_radial_normal():
    """The vertex normals of a finely tessellated sphere smoothed at 180
    degrees point nearly radially, so a sun-lit render approaches the true
    sphere's Lambert term r.s (within 0.005 here; mean 0.0007) where face
    normals deviate by the facet angle (mean 0.01)."""
    devs = {}
    for smooth in (False, True):
        d, s = sun_sphere(smooth)
        oc = po.OracleScene.from_desc(d)
        img, _ = oc.render(d)
        # camera rays through pixel centres and their hits on the unit sphere
        rx = d.cam.resx
        q = np.arange(rx * rx)
        from bench import _camera_rays
        o, dirs = _camera_rays(d.cam, q)
        b = np.sum(o * dirs, 1)
        c = np.sum(o * o, 1) - 1
        disc = b * b - c
        hit = disc > 0.05
        t = -b - np.sqrt(np.maximum(disc, 0))
        r = o + dirs * t[:, None]
        lam = np.maximum(r @ s, 0)
        # away from the terminator, where facets shadow each other differently
        m = hit & (lam > 0.05)
        dev = np.abs(img.reshape(-1, 3)[:, 0] - lam)[m]
        devs[smooth] = (dev.mean(), dev.max())
    assert devs[True][1] < 0.005 and devs[True][0] < 0.001, devs
    assert devs[False][0] > 0.005, devs
