"""Pins of the oracle's global photon map (globalphotonlight.cc, photon.cc,
color.cc rgbe_t), from the mathematics and from invariants:

* RGBE (color.cc:140-155): the decoded colour is the input truncated to 8
  bits below the largest channel's power of two, and encoding is idempotent;
* direction bytes (photon.h:14-37): theta = (int)(acos z 255 / pi) up to
  254, phi = (int)(atan2(y, x) 256 / 2 pi), plus 256 when negative; decoded
  vectors are unit and within about one byte cell of the input;
* the quad emitter (arealight.cc:200-232): a dummy light over a floor that
  stores first hits only (depth 0) puts colour x power x area x 0.5 x
  E[cos] = ... x (1 + 0.0001) / 2 on the floor, each photon's cosine
  uniform in [0.0001, 1];
* pre-gathering (setIrradiance): a dense uniform field of box photons of
  colour c arriving at cos a gives (4 pi / 3) rho c cos a (rho: boxes per
  unit area; the cone weights 1 - d / r average 1/3 over the disc);
* the path vertex term (getLight, pathlight.cc:140-185): a ceiling whose
  boxes all hold irradiance c0 lights a floor seen from below it: pixel =
  rho_f P (1/S) sum_i sqrt(z1_i) rho_c c0, z1 the exact base-2 radical
  inverse of the pixel's per-area index (the two-bounce closed form with the
  sun replaced by the map)."""
import math

import numpy as np
import pytest

import closed_form_scenes as CF
from oracle import pyoracle as po
from yafray_amd import descriptors as D


def test_rgbe_truncation_and_idempotence():
    rng = np.random.default_rng(1)
    c = (rng.uniform(0, 1, (5000, 3)) * 10.0 ** rng.uniform(-6, 6, (5000, 1))).astype(np.float32)
    d = po.rgbe_roundtrip(c)
    m = c.max(axis=1)
    e = np.floor(np.log2(m)) + 1  # frexp exponent of the largest channel
    step = np.exp2(e - 8)[:, None]
    assert np.all(d <= c * (1 + 1e-6)) and np.all(c - d < step * (1 + 1e-6))
    assert np.array_equal(po.rgbe_roundtrip(d), d)
    assert np.all(po.rgbe_roundtrip(np.zeros((3, 3))) == 0)


def test_direction_bytes():
    rng = np.random.default_rng(2)
    v = rng.normal(size=(20000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v = v.astype(np.float32)
    dec, code = po.dir_roundtrip(v)
    th = np.minimum(np.floor(np.arccos(v[:, 2].astype(np.float64)) * 255 / np.pi), 254)
    ph = np.trunc(np.arctan2(v[:, 1], v[:, 0]).astype(np.float64) * 256 / (2 * np.pi))  # (int): toward zero
    ph = np.where(ph < 0, ph + 256, ph)
    assert np.array_equal(code[:, 0], th) and np.array_equal(code[:, 1], ph)
    assert np.allclose(np.linalg.norm(dec, axis=1), 1.0, atol=1e-6)
    # the decoded vector is the cell's lower corner: within one cell of the input
    ang = np.arccos(np.clip(np.sum(dec * v, axis=1), -1, 1))
    assert np.max(ang) < math.hypot(np.pi / 255, 2 * np.pi / 256) * 1.01
    dz, _ = po.dir_roundtrip(np.array([[0, 0, 1], [0, 0, -1], [0, 0, 0]], np.float32))
    assert np.array_equal(dz[0], [0, 0, 1]) and dz[1][2] < -0.9999 and np.all(dz[2] == 0)


def floor_under_dummy(photons=40000, depth=0, search=50, radius=0.5, power=10.0, half=0.5):
    """A 10^4 x 10^4 floor z = 0 (albedo 0.5) under a dummy area light at
    z = 2 facing down, nothing else (a photon misses it only when its cosine
    is below 2 / 5000)."""
    E = 5000.0
    v = np.array([[-E, -E, 0], [E, -E, 0], [E, E, 0], [-E, E, 0]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    h = half
    L = D.arealight((-h, -h, 2), (-h, h, 2), (h, h, 2), (h, -h, 2), samples=1, power=power, dummy=True)
    G = D.globalphotonlight(radius=radius, depth=depth, photons=photons, search=search)
    cam = D.camera((0, 0, 5), (0, 0, 0), (0, 1, 5), 8, 8, 1.0, 1.0)
    return D.SceneDesc(v, f, [0, 0], [0, 0], [7], cam, [D.material((0.5, 0.5, 0.5))], [], [], [L, G],
                       D.render_opts(bias=1e-3))


def test_emitter_power_on_the_floor():
    d = floor_under_dummy()
    oc = po.OracleScene.from_desc(d)
    st = oc.gphoton_build(d, 1)
    assert st[0] == 40000 and 0.999 * 40000 <= st[1] <= 40000  # first hits stored (depth 0)
    # 4 draws to emit; 3 at every hit on a surface used for radiosity (the
    # transmission test and the bounce direction are drawn before shoot
    # returns at depth > maxdepth, globalphotonlight.cc:19-20, 64-77)
    assert st[4] == 4 * 40000 + 3 * st[1]
    boxes = oc.gphoton_read()
    area = 1.0 * 1.0
    expect = 10.0 * area * 0.5 * (1 + 0.0001) / 2
    got = boxes[:, 12].sum()
    # RGBE truncation removes up to 1/256 of the box colour per merge: the
    # sum of the box colours is the emitted total less a small bias
    assert 0.9 * expect < got <= expect * 1.01, (got, expect)


def test_pregathering_closed_form():
    """Boxes on a grid of spacing h on the plane z = 0, each the photon of
    colour c arriving at cos a (direction quantised), N = +z; the recomputed
    irradiance of interior boxes -> (4 pi / 3) c cos a / h^2."""
    h, n = 0.02, 120
    xs = (np.arange(n) - n / 2) * h + h / 2
    X, Y = np.meshgrid(xs, xs)
    P = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    cosa = 0.8
    dirv = np.array([math.sqrt(1 - cosa ** 2), 0, cosa], np.float32)
    qd, _ = po.dir_roundtrip(dirv[None, :])
    c = np.array([0.2, 0.3, 0.4], np.float32)
    qc = po.rgbe_roundtrip(c[None, :])[0]
    boxes = np.zeros((len(P), 15), np.float32)
    boxes[:, 0:3] = P
    boxes[:, 5] = 1.0
    boxes[:, 9:12] = dirv
    boxes[:, 12:15] = c
    d = floor_under_dummy(search=400, radius=0.5)
    oc = po.OracleScene.from_desc(d)
    oc.gphoton_load(d, 1, boxes, stateless=True, recompute=True)
    irr = oc.gphoton_read()
    r = np.sqrt(np.sum(P[:, :2] ** 2, axis=1))
    inner = r < 0.6
    exp = (4 * np.pi / 3) * qc * float(qd[0] @ [0, 0, 1]) / h ** 2
    got = irr[inner, 6:9]
    assert np.allclose(got.mean(0), exp, rtol=0.02), (got.mean(0), exp)
    assert np.max(np.abs(got / exp - 1)) < 0.05


def test_path_vertex_reads_the_map():
    """Two-bounce scene without the sun, depth 1: the floor vertices of the
    ceiling pixels' paths get their light from a map of constant irradiance
    c0 over the floor: pixel = rho_c P (1/S) sum_i sqrt(z1_i) rho_f c0
    (getLight: irradiance x diffuse + lights, pathlight.cc:140-191)."""
    S_, rx, ry = 4, 40, 36
    d, _ = CF.two_bounce(samples=S_, depth=1, resx=rx, resy=ry)
    lights = [L for L in d.lights if L.type != D.YG_LIGHT_SUN]
    G = D.globalphotonlight(radius=100.0, depth=1, photons=1000, search=1)  # 100-unit boxes
    d2 = D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, d.cam, list(d.mats), [], [],
                     lights + [G], d.opts)
    li = len(lights)
    c0 = np.array([0.25, 0.5, 0.75], np.float32)  # exact in RGBE
    xs = np.arange(-22, 22) * 100.0 + 50.0
    X, Y = np.meshgrid(xs, xs)
    boxes = np.zeros((X.size, 15), np.float32)
    boxes[:, 0], boxes[:, 1] = X.ravel(), Y.ravel()
    boxes[:, 5] = 1.0   # N: up, towards the ceiling
    boxes[:, 6:9] = c0
    boxes[:, 11] = 1.0
    oc = po.OracleScene.from_desc(d2)
    oc.gphoton_load(d2, li, boxes, stateless=True)
    img, st = oc.render(d2)
    k = CF.area_pixel_index(rx, ry)
    exp = np.zeros((ry, rx, 3))
    for i in range(ry):
        for j in range(rx):
            z = [CF.radinv(int(k[i, j]) * S_ + s + 1, 2) for s in range(S_)]
            exp[i, j] = np.array(CF.RHO_C) * 1.5 * (np.sum(np.sqrt(z)) / S_) * np.array(CF.RHO_F) * c0
    assert np.max(np.abs(img - exp) / exp) < 3e-5
