"""Pins of the oracle's area light beyond the penumbra loop
(areaLight_t::illuminate and guessLight, arealight.cc:93-200), from the
compiler, the C library and geometry:

* draw order: the one-sample light's point is corner + toX*ourRandom() +
  toY*ourRandom(); compiled by g++ (as the reference is) the right operand
  of + is evaluated first, so toY takes the first draw
  (tests/evalorder_check.cpp, -O0 and -O2, gnu++98);
* the glibc rand() stream guessLight draws from continues after the block
  shuffle's draws: the oracle's incremental generator equals this host's
  libc rand() after srand(seed) and as many calls;
* fully lit: a floor under a 4 x 4-sample light with nothing between them
  gets guessLight's LIGHT verdict everywhere: the unjittered grid points
  summed without shadow rays, (plane_at / 16) sum_i rho E cos_i / d_i^2,
  and no ourRandom() draw at all;
* fully shadowed: an opaque plane between them gives the SHADOW verdict:
  black and no draw (with psamples 0 every point would draw 32 times);
* one sample: pixel = plane_at rho E cos / d^2 towards corner + toX r1 +
  toY r2, r2 and r1 the Park-Miller values 2k and 2k + 1 of the k-th lit
  visit in render order (areas in processing order, overlap included)."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

import bench
import closed_form_scenes as CF
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("opt", ["-O0", "-O2"])
def test_gxx_takes_the_right_operand_first(tmp_path, opt):
    exe = str(tmp_path / "evalorder")
    subprocess.run(["g++", "-std=gnu++98", opt, os.path.join(ROOT, "tests", "evalorder_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    assert subprocess.run([exe], capture_output=True, text=True, check=True).stdout.strip() == "toY first"


def test_rand_stream_continues_after_the_shuffle():
    libc = ctypes.CDLL("libc.so.6")
    for seed, skip in ((1, 20), (1, 240), (12345, 7)):
        libc.srand(seed)
        for _ in range(skip):
            libc.rand()
        ref = np.array([libc.rand() for _ in range(5000)], np.int32)
        got = np.zeros(5000, np.int32)
        po.lib().yo_grand_stream(seed, skip, 5000, got)
        assert np.array_equal(got, ref)


def _floor_points(d):
    """camera rays through pixel centres and their floor hits (z = 0), float64"""
    n = d.cam.resx * d.cam.resy
    o, dirs = bench._camera_rays(d.cam, np.arange(n))
    t = -o[:, 2] / dirs[:, 2]
    return o + dirs * t[:, None]


def test_fully_lit_floor_closed_form():
    d = CF.area_over_floor(samples=16, psamples=4)
    img, st = po.OracleScene.from_desc(d).render(d)
    L = d.lights[0]
    P = _floor_points(d)
    exp = CF.lambert_from_points(P, CF.area_grid_points(L), L, (0.7, 0.6, 0.5))
    np.testing.assert_allclose(img.reshape(-1, 3), exp, rtol=2e-4, atol=1e-7)
    assert int(st[4]) == 0  # no ourRandom() draw
    # the same light without psamples samples the penumbra: 32 draws per visit
    L0 = CF.area_over_floor(samples=16, psamples=0)
    _, st0 = po.OracleScene.from_desc(L0).render(L0)
    assert int(st0[4]) > 0 and int(st0[4]) % 32 == 0


def test_fully_shadowed_floor_is_black_without_draws():
    d = CF.area_over_floor(samples=16, psamples=3, slab=True)
    img, st = po.OracleScene.from_desc(d).render(d)
    # camera rays see the slab's underside?  no: the camera is below it, looking down
    assert np.all(img == 0) and int(st[4]) == 0


def _lcg(seed, n):
    m = 2147483647
    return (seed * pow(16807, n + 1, m)) % m / m


def test_single_sample_light_draw_order():
    d = CF.area_over_floor(samples=1, psamples=0, resx=40, resy=30)
    img, st = po.OracleScene.from_desc(d).render(d)
    L = d.lights[0]
    rx, ry = d.cam.resx, d.cam.resy
    regions = po.block_regions(rx, ry, 64, d.opts.rand_seed)
    a, b, dd = (np.array(L.a[:], np.float64), np.array(L.b[:], np.float64), np.array(L.d[:], np.float64))
    P = _floor_points(d).reshape(ry, rx, 3)
    exp = np.zeros((ry, rx, 3))
    k = 0
    seed = int(d.opts.lcg_seed)
    for (x, y, w, h, rxx, ryy, rw, rh) in regions:
        for i in range(y, y + h):
            for j in range(x, x + w):
                r2, r1 = _lcg(seed, 2 * k), _lcg(seed, 2 * k + 1)
                k += 1
                if not (rxx <= j < rxx + rw and ryy <= i < ryy + rh):
                    continue
                q = a + (b - a) * r1 + (dd - a) * r2
                exp[i, j] = CF.lambert_from_points(P[i, j][None, :], q[None, :], L, (0.7, 0.6, 0.5))[0]
    assert int(st[4]) == 2 * k
    np.testing.assert_allclose(img, exp, rtol=2e-4, atol=1e-7)
