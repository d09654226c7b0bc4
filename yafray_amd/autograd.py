"""torch.autograd bridge for the differentiable GI render.

`render_gi(scene, albedo, amp, size, sun_power)` renders the scene with the
given continuous shading parameters and returns the image (resy, resx, 3).
Its backward calls yg_render_adjoint, which replays every pathlight path of
the forward render with the same samples (visibility held fixed) and returns
dLoss/d albedo, d amplitude, d frequency (modulator size), d sun power.

The parameters map onto the reference's shading inputs: albedo = the generic
shader's "color" (basicshaders.cc:238-367), amp = the modulator's "color"
blend factor and size = its "size" (texture.cc:122-320, 499-534), sun_power =
sunlight "power" (sunlight.cc:59-73).
"""
import ctypes as C

import torch

from . import descriptors as D
from . import lib as Y


def _set_params(scene, albedo, amp, size, sun_power):
    d = scene.desc
    a = albedo.detach().double().cpu().numpy()
    mats = [D.material(tuple(float(x) for x in a[k]), d.mats[k].mod_first, d.mats[k].mod_count)
            for k in range(len(d.mats))]
    mods = [D.yg_modulator.from_buffer_copy(m) for m in d.mods]
    if len(mods):
        mods[0].color = float(amp)
        mods[0].size = D.F3(float(size), float(size), float(size))
    lights = [D.yg_light.from_buffer_copy(L) for L in d.lights]
    for L in lights:
        if L.type == D.YG_LIGHT_SUN:
            L.power = float(sun_power)
    nd = D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, d.cam, mats, mods,
                     list(d.texs), lights, d.opts)
    scene.set_desc(nd)


class GIRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, scene, albedo, amp, size, sun_power):
        _set_params(scene, albedo, amp, size, sun_power)
        d = scene.desc
        out = torch.empty(d.cam.resy * d.cam.resx * 3, dtype=torch.float32, device="cuda")
        scene.render(out)
        ctx.scene = scene
        ctx.size = float(size)
        ctx.meta = [(t.device, t.dtype) if torch.is_tensor(t) else None
                    for t in (albedo, amp, size, sun_power)]
        ctx.nsun = [i for i, L in enumerate(d.lights) if L.type == D.YG_LIGHT_SUN]
        return out.view(d.cam.resy, d.cam.resx, 3)

    @staticmethod
    def backward(ctx, grad):
        scene = ctx.scene
        d = scene.desc
        g = grad.contiguous().float().view(-1)
        out = torch.empty_like(g)
        gmat = torch.zeros(len(d.mats) * 3, dtype=torch.float64, device="cuda")
        gmod = torch.zeros(max(len(d.mods), 1) * 2, dtype=torch.float64, device="cuda")
        glight = torch.zeros(len(d.lights), dtype=torch.float64, device="cuda")
        o = D.yg_render_opts.from_buffer_copy(d.opts)
        o.flags |= 4  # YG_RENDER_REUSE: replay the forward render just made
        Y.check(Y.lib().yg_render_adjoint(scene.h, C.byref(o), Y._ptr(g), Y._ptr(out), Y._ptr(gmat),
                                          Y._ptr(gmod), Y._ptr(glight), None, Y._stream()),
                "yg_render_adjoint")
        grads = [gmat.view(-1, 3),
                 gmod[0] if len(d.mods) else None,
                 (gmod[1] / ctx.size) if len(d.mods) else None,
                 glight[ctx.nsun[0]] if ctx.nsun else None]
        out_g = []
        for gr, meta in zip(grads, ctx.meta):
            if gr is None or meta is None:
                out_g.append(None)
            else:
                out_g.append(gr.to(device=meta[0], dtype=meta[1]))
        return (None, *out_g)


def render_gi(scene, albedo, amp, size, sun_power):
    return GIRender.apply(scene, albedo, amp, size, sun_power)
