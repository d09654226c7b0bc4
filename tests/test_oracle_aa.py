"""Pins for the oracle's antialiasing passes, scene_t::render(renderArea_t&)
(scene.cc:724-816) with renderArea_t::checkResample (renderblock.cc:22-69).

The oracle's render is compared bit for bit with a numpy restatement of the
pass loop (area split, jittered first pass, 8-neighbour contrast mask, RI_LP
sample offsets, running average) that takes its per-ray colours from the
oracle's camera-ray hook; the radical inverses it uses are pinned in
test_oracle_pins.py.  An infinite threshold must reproduce the plain render."""
import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import descriptors as D
from yafray_amd import scenes as S

F32, F64 = np.float32, np.float64


def sun_terrain(resx=96, resy=64, **aa):
    """Small fBm terrain under the sun and a constant sky (no pathlight)."""
    d = S.terrain_gi(po.heightfield, nx=96, ny=48, resx=resx, resy=resy, samples=1, depth=1)
    d.lights = D.carray(D.yg_light, [d.lights[1]])  # the sun only
    d.opts = D.render_opts(bias=2e-3, raydepth=3, background=(0.4, 0.5, 0.7), **aa)
    return d


def bri(dif):
    """abscol2bri of a float32 difference, summed in double (color.h:71)."""
    w = [F64(F32(0.299)), F64(F32(0.587)), F64(F32(0.114))]
    return (w[0] * np.abs(dif[..., 0].astype(F64)) + w[1] * np.abs(dif[..., 1].astype(F64)) +
            w[2] * np.abs(dif[..., 2].astype(F64))).astype(F32)


def restated_render(oc, d):
    L = po.lib()
    o = d.opts
    resx, resy = d.cam.resx, d.cam.resy
    passes, mins = o.aa_passes, o.aa_minsamples
    pw, thr = F64(F32(o.aa_pixelwidth)), F32(o.aa_threshold)
    tsd = F32(mins * passes)
    if tsd != 0:
        tsd = F32(1.0 / F64(tsd))
    img = np.zeros((resy, resx, 3), F32)
    resampled = 0
    for x, y, w, h, rx, ry, rw, rh in po.block_regions(resx, resy, 64, o.rand_seed):
        ii, jj = np.mgrid[y:y + h, x:x + w]
        if o.aa_jitterfirst and passes:
            n = range(1, w * h + 1)  # the area's pixel counter, row-major
            fx = np.array([L.yo_ri_vdc(k, 0) for k in n], F32)
            fy = np.array([L.yo_ri_s(k, 0) for k in n], F32)
        else:
            fx = fy = np.full(w * h, 0.5, F32)
        A = oc.trace_camera(d, jj.ravel().astype(F32) + fx, ii.ravel().astype(F32) + fy).reshape(h, w, 3)
        for p in range(passes):
            pad = np.pad(A, ((1, 1), (1, 1), (0, 0)), mode="edge")  # neighbours clamped to the area
            need = np.zeros((h, w), bool)
            for di in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    if di or dj:
                        need |= bri(A - pad[1 + di:1 + di + h, 1 + dj:1 + dj + w]) >= thr
            li, lj = np.nonzero(need)  # row-major within the area
            resampled += len(li)
            px, py = [], []
            for a, b in zip(li, lj):
                i, j = y + a, x + b
                pixnum = j + i * resx
                for ms in range(mins):
                    cursam = p * mins + ms
                    ox = F32(0.5 + pw * (F64(L.yo_ri_lp(cursam + pixnum, 0)) - 0.5))
                    oy = F32(0.5 + pw * (F64(F32(F32(cursam) * tsd)) - 0.5))
                    px.append(F32(j) + ox)
                    py.append(F32(i) + oy)
            if not px:
                continue
            cols = oc.trace_camera(d, np.array(px, F32), np.array(py, F32)).reshape(len(li), mins, 3)
            mf = F32(p * mins + 1)
            inv = F32(1.0 / F64(mf + F32(mins)))
            for q, (a, b) in enumerate(zip(li, lj)):
                tot = np.zeros(3, F32)
                for ms in range(mins):
                    tot = tot + cols[q, ms]
                A[a, b] = (mf * A[a, b] + tot) * inv
        img[ry:ry + rh, rx:rx + rw] = A[ry - y:ry - y + rh, rx - x:rx - x + rw]
    return img, resampled


@pytest.mark.parametrize("aa", [dict(aa_passes=2, aa_minsamples=3, aa_threshold=0.05),
                                dict(aa_passes=1, aa_minsamples=4, aa_threshold=0.02, aa_jitterfirst=True,
                                     aa_pixelwidth=2.0)])
def test_aa_render_matches_restatement(aa):
    d = sun_terrain(**aa)
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    img, stats = oc.render(d, threads=4)
    ref, resampled = restated_render(oc, d)
    assert resampled > 50  # the horizon and shadow edges are resampled
    assert np.array_equal(img, ref)
    # camera rays: every area pixel once, plus the AA samples
    n_area_px = sum(int(r[2]) * int(r[3]) for r in po.block_regions(d.cam.resx, d.cam.resy))
    assert int(stats[0]) == n_area_px + resampled * d.opts.aa_minsamples
    plain = sun_terrain()
    base, _ = oc.render(plain, threads=4)
    assert not np.array_equal(img, base)


def test_aa_unreachable_threshold_is_the_plain_render():
    d = sun_terrain(aa_passes=3, aa_minsamples=2, aa_threshold=1e30)
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    img, stats = oc.render(d, threads=4)
    base, bstats = oc.render(sun_terrain(), threads=4)
    assert np.array_equal(img, base)
    assert int(stats[0]) == int(bstats[0])


def tonemap_np(img, exposure, gamma, clamp):
    """color_t::expgam_Adjust (color.h:175-195) with scene_t's exposure = -e,
    gamma_R = 1/g; exp and pow in double, each step stored as float."""
    e = F32(-exposure)
    g = F32(1.0 / F64(F32(gamma))) if gamma != 0 else F32(0)
    c = img.astype(F32)
    if e == 0 and g == 1:
        return np.clip(c, 0, 1) if clamp else c
    if e != 0:
        c = (1.0 - np.exp((np.maximum(c, F32(0)) * e).astype(F64))).astype(F32)
    if g != 1:
        c = np.power(np.maximum(c, F32(0)).astype(F64), F64(g)).astype(F32)
    return c


@pytest.mark.parametrize("exposure,gamma,clamp", [(0.0, 1.0, True), (1.5, 1.0, False), (0.0, 2.2, False),
                                                  (0.8, 1.8, True)])
def test_tonemapping_is_expgam_adjust_of_each_colour(exposure, gamma, clamp):
    d = sun_terrain()
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    plain, _ = oc.render(d, threads=4)
    d.opts = D.render_opts(bias=2e-3, raydepth=3, background=(0.4, 0.5, 0.7), tonemap=True, exposure=exposure,
                           gamma=gamma, clamp_rgb=clamp)
    img, _ = oc.render(d, threads=4)
    ref = tonemap_np(plain, exposure, gamma, clamp)
    assert np.mean(img == ref) >= 0.9999
    assert np.max(np.abs(img - ref)) <= 1e-6
    if exposure or gamma != 1:
        assert not np.array_equal(img, plain)
