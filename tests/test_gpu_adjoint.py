"""Differentiable GI on the GPU: the path-replay adjoint (yg_render_adjoint)
against the oracle's FP64 reference gradient and central finite differences
(h = 1e-3) of the GPU forward render."""
import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _render(Y, torch, sc, desc):
    sc.set_desc(desc)
    out = torch.zeros(desc.cam.resx * desc.cam.resy * 3, dtype=torch.float32, device="cuda")
    sc.render(out)
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64).reshape(desc.cam.resy, desc.cam.resx, 3)


@pytest.fixture(scope="module")
def setup(torch_cuda):
    torch = torch_cuda
    from yafray_amd import lib as Y
    # albedos in [0.35, 0.9]: no path of depth 3 falls below the pathlight's
    # energy cut-off (pathlight.cc:315), so the loss is smooth in every
    # parameter and central differences apply
    d = S.terrain_diff(po.heightfield, nx=96, ny=48, res=48, samples=4, depth=3, nmat=16, albedo_lo=0.35)
    sc = Y.Scene.from_desc(d)
    alb = np.array([list(m.color) for m in d.mats], np.float64)
    target = _render(Y, torch, sc, S.with_params(d, albedo=alb * 0.9 + 0.04, amp=0.4, size=2.7,
                                                 sun_power=2.5))
    img = _render(Y, torch, sc, d)
    g = (2 * (img - target)).astype(np.float32)
    gt = torch.from_numpy(g.reshape(-1)).cuda()
    out = torch.zeros_like(gt)
    gmat = torch.zeros(len(d.mats) * 3, dtype=torch.float64, device="cuda")
    gmod = torch.zeros(2, dtype=torch.float64, device="cuda")
    glight = torch.zeros(len(d.lights), dtype=torch.float64, device="cuda")
    sc.render_adjoint(gt, out, gmat, gmod, glight)
    torch.cuda.synchronize()
    res = dict(gmat=gmat.cpu().numpy().reshape(-1, 3), gmod=gmod.cpu().numpy(),
               glight=glight.cpu().numpy(), img_adj=out.cpu().numpy().reshape(img.shape))

    def loss(desc):
        im = _render(Y, torch, sc, desc)
        return float(np.sum((im - target) ** 2))

    return Y, torch, sc, d, alb, img, g, target, res, loss


def test_adjoint_forward_image_matches_render(setup):
    Y, torch, sc, d, alb, img, g, target, res, loss = setup
    assert np.array_equal(res["img_adj"], img.astype(np.float32))


def test_adjoint_matches_fp64_oracle(setup):
    Y, torch, sc, d, alb, img, g, target, res, loss = setup
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    _, gm, gmod, gl = oc.render_adjoint(d, g)
    for a, b in ((res["gmat"], gm), (res["gmod"], gmod.reshape(-1)), (res["glight"], gl)):
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert err <= 1e-3, err


H = 1e-3


def test_adjoint_matches_finite_differences(setup):
    Y, torch, sc, d, alb, img, g, target, res, loss = setup
    gm = res["gmat"]
    big = np.argsort(-np.abs(gm).reshape(-1))[:4]
    for flat in big:
        m, c = divmod(int(flat), 3)
        a = alb.copy()
        a[m, c] += H
        lp = loss(S.with_params(d, albedo=a))
        a[m, c] -= 2 * H
        lm = loss(S.with_params(d, albedo=a))
        f = (lp - lm) / (2 * H)
        assert abs(gm[m, c] - f) <= 1e-3 * abs(f), (m, c, gm[m, c], f)
    f = (loss(S.with_params(d, sun_power=3.0 + H)) - loss(S.with_params(d, sun_power=3.0 - H))) / (2 * H)
    assert abs(res["glight"][1] - f) <= 1e-3 * abs(f)
    f = (loss(S.with_params(d, amp=0.5 + H)) - loss(S.with_params(d, amp=0.5 - H))) / (2 * H)
    assert abs(res["gmod"][0] - f) <= 1e-3 * abs(f)
    f = (loss(S.with_params(d, size=3.0 + H)) - loss(S.with_params(d, size=3.0 - H))) / (2 * H)
    assert abs(res["gmod"][1] / 3.0 - f) <= 2e-3 * abs(f)


def test_autograd_function(setup):
    Y, torch, sc, d, alb, img, g, target, res, loss = setup
    from yafray_amd.autograd import render_gi
    albedo = torch.tensor(alb, dtype=torch.float64, requires_grad=True)
    amp = torch.tensor(0.5, requires_grad=True)
    size = torch.tensor(3.0, requires_grad=True)
    power = torch.tensor(3.0, requires_grad=True)
    tgt = torch.from_numpy(target).float().cuda()
    im = render_gi(sc, albedo, amp, size, power)
    L = ((im - tgt) ** 2).sum()
    L.backward()
    assert np.allclose(albedo.grad.numpy(), res["gmat"], rtol=1e-4, atol=1e-6)
    assert abs(float(power.grad) - res["glight"][1]) <= 1e-4 * abs(res["glight"][1])
    assert abs(float(amp.grad) - res["gmod"][0]) <= 1e-4 * abs(res["gmod"][0])
    assert abs(float(size.grad) - res["gmod"][1] / 3.0) <= 1e-4 * abs(res["gmod"][1] / 3.0)


def test_adjoint_zero_albedo_channel(setup):
    """A vertex albedo with a zero channel: the replay cannot divide the
    path's suffix radiance by it, so the per-path product-rule replay runs
    instead; the gradient still matches the FP64 reference."""
    Y, torch, sc, d, alb, img, g, target, res, loss = setup
    a = alb.copy()
    a[:, 1] = 0.0
    d0 = S.with_params(d, albedo=a)
    sc.set_desc(d0)
    gt = torch.from_numpy(g.reshape(-1)).cuda()
    out = torch.zeros_like(gt)
    gmat = torch.zeros(len(d.mats) * 3, dtype=torch.float64, device="cuda")
    gmod = torch.zeros(2, dtype=torch.float64, device="cuda")
    glight = torch.zeros(len(d.lights), dtype=torch.float64, device="cuda")
    sc.render_adjoint(gt, out, gmat, gmod, glight)
    torch.cuda.synchronize()
    sc.set_desc(d)
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    _, gm, gmo, gl = oc.render_adjoint(d0, g)
    for x, y in ((gmat.cpu().numpy().reshape(-1, 3), gm), (gmod.cpu().numpy(), gmo.reshape(-1)),
                 (glight.cpu().numpy(), gl)):
        err = np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-30)
        assert err <= 1e-3, err
    assert np.any(gm[:, 1] != 0)


def test_adjoint_at_workload_size_matches_fp64_oracle(torch_cuda):
    """The differentiable-GI workload at its own size: 1,045,506-triangle
    terrain, 512x512, 16 spp, 3 bounces, 256 materials with albedos drawn
    U[0.1, 0.9], so paths do end at the pathlight's energy cut-off (the
    adjoint holds each path's length fixed, as it holds visibility).  The
    GPU gradient of an L2 loss against a target rendered from other
    parameters equals the oracle's double-precision gradient of the same
    loss: the whole gradient within 1e-3 relative L2 error, and each of 64
    sampled albedo components among the larger ones, the modulator's
    amplitude and frequency and the sun's power within 1e-2."""
    torch = torch_cuda
    from yafray_amd import lib as Y
    d = S.terrain_diff(po.heightfield)
    assert (d.cam.resx, d.cam.resy, len(d.mats), d.n_tris) == (512, 512, 256, 1045506)
    sc = Y.Scene.from_desc(d)
    alb = np.array([list(m.color) for m in d.mats], np.float64)
    target = _render(Y, torch, sc, S.with_params(d, albedo=alb * 0.9 + 0.04, amp=0.4, size=2.7, sun_power=2.5))
    img = _render(Y, torch, sc, d)
    g = (2 * (img - target)).astype(np.float32)
    gt = torch.from_numpy(g.reshape(-1)).cuda()
    out = torch.zeros_like(gt)
    gmat = torch.zeros(len(d.mats) * 3, dtype=torch.float64, device="cuda")
    gmod = torch.zeros(2, dtype=torch.float64, device="cuda")
    glight = torch.zeros(len(d.lights), dtype=torch.float64, device="cuda")
    sc.render_adjoint(gt, out, gmat, gmod, glight)
    torch.cuda.synchronize()
    gm_gpu, gmod_gpu, gl_gpu = gmat.cpu().numpy(), gmod.cpu().numpy(), glight.cpu().numpy()
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    _, gm, gmo, gl = oc.render_adjoint(d, g, threads=16)
    gm, gmo = gm.reshape(-1), gmo.reshape(-1)
    full_gpu = np.concatenate([gm_gpu, gmod_gpu, gl_gpu])
    full_ref = np.concatenate([gm, gmo, gl])
    err = np.linalg.norm(full_gpu - full_ref) / np.linalg.norm(full_ref)
    print("gradient relative L2 error %.3g" % err)
    assert err <= 1e-3, err
    big = np.flatnonzero(np.abs(gm) >= 0.05 * np.abs(gm).max())
    pick = np.random.default_rng(7).choice(big, size=min(64, len(big)), replace=False)
    rel = np.abs(gm_gpu[pick] - gm[pick]) / np.abs(gm[pick])
    print("sampled albedo components: max relative error %.3g (of %d)" % (rel.max(), len(big)))
    assert rel.max() <= 1e-2, rel.max()
    for a, b in ((gmod_gpu[0], gmo[0]), (gmod_gpu[1], gmo[1]), (gl_gpu[1], gl[1])):
        assert abs(a - b) <= 1e-2 * abs(b), (a, b)
