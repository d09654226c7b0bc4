"""Texture coordinates of the modulators (modulator_t::doMapping,
texture.cc:125-165) and the orco matrix of meshObject_t::transform
(mesh.cc:422-493), pinned by geometry: the matrix's closed form for the
identity transform, its mapping of a transformed mesh's bound onto the unit
cube, and closed-form images of a wall seen through each texco
(tests/closed_form_scenes.py)."""
import math

import numpy as np
import pytest

import closed_form_scenes as CF
from oracle import pyoracle as po
from yafray_amd import descriptors as D


def _box(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [2, 3, 7], [2, 7, 6], [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]], np.uint32)
    return v, f


def _f32(x):
    return np.float32(x)


def test_orco_matrix_identity_closed_form():
    """M = identity: backOrco = scale(h) with translation c, so its inverse
    is x -> x / h - c / h per axis, with h = half the extent of the bound
    widened by BOX_MIN (1e-5, each bound rounded once from double) and c its
    centre; Gauss-Jordan leaves exactly fl(1/h) and -fl(c/h)."""
    lo, hi = (-1.25, 0.5, 2.0), (3.0, 0.75, 2.5)
    v, f = _box(lo, hi)
    got = po.orco_matrix(v, f)
    exp = np.zeros((3, 4), np.float32)
    for a in range(3):
        l = _f32(np.float64(_f32(lo[a])) - 1e-5)
        u = _f32(np.float64(_f32(hi[a])) + 1e-5)
        h = _f32(0.5 * np.float64(_f32(u - l)))
        c = _f32(_f32(0.5) * _f32(l + u))
        exp[a, a] = _f32(_f32(1.0) / h)
        exp[a, 3] = -_f32(c / h)
    assert np.array_equal(got, exp)


def _rot(axis, deg):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    if axis == 0:
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    if axis == 1:
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_orco_matrix_maps_the_object_box_to_the_unit_cube(seed):
    """For a transform M = T R S the orco matrix takes the transformed
    vertices back to the untransformed box scaled to [-1, 1]^3: each corner
    of the box lands on its sign vector (within float rounding and the
    BOX_MIN widening)."""
    rng = np.random.default_rng(seed)
    lo = rng.uniform(-2, 0, 3)
    hi = lo + rng.uniform(0.5, 3, 3)
    v, f = _box(lo, hi)
    R = _rot(0, rng.uniform(0, 360)) @ _rot(2, rng.uniform(0, 360)) @ np.diag(rng.uniform(0.5, 2, 3))
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = rng.uniform(-5, 5, 3)
    M = M.astype(np.float32)
    B = po.orco_matrix(v, f, M).astype(np.float64)
    w = (v.astype(np.float64) @ M[:3, :3].astype(np.float64).T) + M[:3, 3]
    orco = w @ B[:, :3].T + B[:, 3]
    signs = np.sign(v - (lo + hi) / 2)
    # the untransformed box's centre need not map to the origin (the
    # translation is the transformed bound's centre), but the box's extent
    # maps to 2 along each axis
    span = orco.max(0) - orco.min(0)
    assert np.allclose(span, 2.0, atol=1e-4)
    centred = orco - (orco.max(0) + orco.min(0)) / 2
    assert np.allclose(centred, signs, atol=1e-4)


def test_orco_matrix_library_equals_oracle():
    """yg_orco_matrix (the library's host code) and the oracle's restatement
    agree bit for bit on identity and general transforms."""
    from yafray_amd import lib as Y
    rng = np.random.default_rng(7)
    for k in range(6):
        v = rng.uniform(-3, 3, (40, 3)).astype(np.float32)
        f = rng.integers(0, 40, (30, 3)).astype(np.uint32)
        M = None
        if k:
            M = np.eye(4)
            M[:3, :3] = _rot(k % 3, rng.uniform(0, 360)) * rng.uniform(0.5, 2)
            M[:3, 3] = rng.uniform(-4, 4, 3)
            M = M.astype(np.float32)
        assert np.array_equal(Y.orco_matrix(v, f, M), po.orco_matrix(v, f, M))


@pytest.mark.parametrize("kind", list(CF.TEXCO_KINDS))
def test_texco_wall_closed_form(kind):
    """Every pixel of the wall away from cell edges shows the marble colour
    at the coordinates the texco gives (global, uv with faces with uv /
    vertex colours only / neither, orco from the matrix and from orco
    points, window, normal, reflect)."""
    d = CF.texco_wall(kind)
    img, _ = po.OracleScene.from_desc(d).render(d)
    exp, mask = CF.texco_wall_expected(kind)
    assert mask.mean() > 0.8
    err = np.abs(img[mask] - exp[mask]).max()
    assert err < 2e-5, (kind, err)


def test_texco_kinds_differ():
    """The closed forms are not vacuous: the seven coordinate kinds give
    seven different images."""
    imgs = [CF.texco_wall_expected(k)[0] for k in CF.TEXCO_KINDS]
    for a in range(len(imgs)):
        for b in range(a):
            assert np.abs(imgs[a] - imgs[b]).max() > 0.05


def test_texco_names():
    """string2texcotype (texture.h:135-149): any unknown name is uv."""
    assert D.texco_of("global") == D.YG_TEXCO_GLOBAL and D.texco_of("orco") == D.YG_TEXCO_ORCO
    assert D.texco_of("window") == D.YG_TEXCO_WIN and D.texco_of("normal") == D.YG_TEXCO_NOR
    assert D.texco_of("reflect") == D.YG_TEXCO_REFL
    assert D.texco_of("uv") == D.texco_of("sticky") == D.texco_of("") == D.YG_TEXCO_UV


def test_fp64_oracle_refuses_other_coordinates():
    d = CF.texco_wall("uv")
    with pytest.raises(RuntimeError):
        po.OracleScene.from_desc(d).render_fp64(d)
