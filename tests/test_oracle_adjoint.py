"""FP64 reference gradient of the oracle against central finite differences
of its own forward render (h = 1e-3), on a small differentiable terrain."""
import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import scenes as S


@pytest.fixture(scope="module")
def setup():
    d = S.terrain_diff(po.heightfield, nx=48, ny=24, res=40, samples=4, depth=3, nmat=8, albedo_lo=0.35)
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    alb = np.array([list(m.color) for m in d.mats], np.float64)
    tgt_desc = S.with_params(d, albedo=alb * 0.9 + 0.04, amp=0.4, size=2.7, sun_power=2.5)
    target, _ = oc.render(tgt_desc)
    img, _ = oc.render(d)
    g = 2 * (img.astype(np.float64) - target)
    img2, gm, gmod, gl = oc.render_adjoint(d, g.astype(np.float32))
    assert np.array_equal(img, img2)

    def loss(dd):
        im, _ = oc.render(dd)
        return float(np.sum((im.astype(np.float64) - target) ** 2))

    return d, oc, alb, loss, gm, gmod, gl


H = 1e-3


def fd(loss, mk, x0):
    return (loss(mk(x0 + H)) - loss(mk(x0 - H))) / (2 * H)


def test_albedo_gradient_matches_fd(setup):
    d, oc, alb, loss, gm, gmod, gl = setup
    for (m, c) in [(0, 0), (2, 1), (4, 2), (6, 0), (7, 1)]:
        def mk(x, m=m, c=c):
            a = alb.copy()
            a[m, c] = x
            return S.with_params(d, albedo=a)
        f = fd(loss, mk, alb[m, c])
        assert abs(gm[m, c] - f) <= 1e-3 * max(abs(f), 1e-2 * np.max(np.abs(gm))), (m, c, gm[m, c], f)


def test_sun_power_gradient_matches_fd(setup):
    d, oc, alb, loss, gm, gmod, gl = setup
    f = fd(loss, lambda x: S.with_params(d, sun_power=x), 3.0)
    assert abs(gl[1] - f) <= 1e-3 * abs(f)
    assert gl[0] == 0.0  # the pathlight's power is not a free parameter here


def test_texture_gradients_match_fd(setup):
    d, oc, alb, loss, gm, gmod, gl = setup
    f_amp = fd(loss, lambda x: S.with_params(d, amp=x), 0.5)
    assert abs(gmod[0, 0] - f_amp) <= 1e-3 * abs(f_amp)
    # d/dlambda with every size scaled by lambda: d/dsize = d/dlambda / size
    f_size = fd(loss, lambda x: S.with_params(d, size=x), 3.0)
    assert abs(gmod[0, 1] / 3.0 - f_size) <= 2e-3 * abs(f_size)
