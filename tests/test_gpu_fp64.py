"""GPU images against the converged FP64 render (oracle yo_render_fp64 with
4096 samples of every sampled light, in double): the error of an image is its
sampling noise, and it must fall as the pathlight's sample count grows (the
GPU estimator converges to the same value as the double-precision one)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _pixels(d, n, seed=7):
    q = np.random.default_rng(seed).choice(d.cam.resx * d.cam.resy, size=n, replace=False)
    return (q % d.cam.resx).astype(np.int32), (q // d.cam.resx).astype(np.int32)


def _err(img, conv, px, py):
    v = img[py, px].astype(np.float64)
    return float(np.sqrt(np.mean((v - conv) ** 2)) / np.sqrt(np.mean(conv ** 2)))


def test_gi_error_falls_toward_fp64_render(torch_cuda):
    torch = torch_cuda
    from yafray_amd import lib as Y
    errs = []
    conv = None
    for spp in (4, 64, 512):
        d = S.terrain_gi(po.heightfield, nx=64, ny=32, resx=80, resy=72, samples=spp)
        sc = Y.Scene.from_desc(d)
        out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
        sc.render(out)
        torch.cuda.synchronize()
        img = out.cpu().numpy().reshape(d.cam.resy, d.cam.resx, 3)
        assert np.all(np.isfinite(img))
        px, py = _pixels(d, 512)
        if conv is None:
            # the converged value does not depend on the pathlight's own sample count
            oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
            conv = oc.render_fp64(d, samples=4096, px=px, py=py, threads=8)
        errs.append(_err(img, conv, px, py))
    # the FP32 oracle gives 0.098, 0.019 and 0.0056 on these pixels
    assert errs[0] > errs[1] > errs[2], errs
    assert errs[2] < 0.01 and errs[1] < 0.03 and errs[0] < 0.15, errs


def test_area_light_error_against_fp64_render(torch_cuda):
    torch = torch_cuda
    from yafray_amd import lib as Y
    errs = []
    for samples in (4, 64):
        d = S.direct_room(96, 72, samples=samples, subdiv=3)
        sc = Y.Scene.from_desc(d)
        out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
        sc.render(out)
        torch.cuda.synchronize()
        img = out.cpu().numpy().reshape(d.cam.resy, d.cam.resx, 3)
        px, py = _pixels(d, 768)
        oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
        conv = oc.render_fp64(d, samples=4096, px=px, py=py, threads=8)
        errs.append(_err(img, conv, px, py))
    # the FP32 oracle gives 0.039 and 0.0044 on these pixels
    assert errs[1] < errs[0] / 4 and errs[1] < 0.01, errs
