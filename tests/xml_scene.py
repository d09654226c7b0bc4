"""A SceneDesc written as a YafRay XML scene file (the element grammar of
src/gram.yafray, parameter names of the reference factories), so that scene
files can be checked against the same scene built through the C ABI.  Floats
are written with 9 significant digits, which reads back to the same float32."""
import numpy as np

from yafray_amd import descriptors as D


def _g(x):
    return "%.9g" % float(x)


def _rgb(tag, c):
    return '<%s r="%s" g="%s" b="%s"/>' % (tag, _g(c[0]), _g(c[1]), _g(c[2]))


def _xyz(tag, p):
    return '<%s x="%s" y="%s" z="%s"/>' % (tag, _g(p[0]), _g(p[1]), _g(p[2]))


def _onoff(b):
    return "on" if b else "off"


def light_xml(k, L):
    name = "L%02d" % k  # lights in name order = descriptor order
    if L.type == D.YG_LIGHT_SUN:
        return ('<light type="sunlight" name="%s" power="%s" cast_shadows="%s">%s%s</light>'
                % (name, _g(L.power), _onoff(L.cast_shadows), _xyz("from", L.dir), _rgb("color", L.color)))
    if L.type == D.YG_LIGHT_PATH:
        return ('<light type="pathlight" name="%s" power="%s" depth="%d" samples="%d" use_QMC="%s" '
                'threshold="%s" cache="off"/>' % (name, _g(L.power), L.depth, L.samples, _onoff(L.use_qmc),
                                                  _g(L.threshold)))
    if L.type == D.YG_LIGHT_AREA:
        return ('<light type="arealight" name="%s" power="%s" samples="%d" psamples="%d">%s%s%s%s%s</light>'
                % (name, _g(L.power), L.samples, L.psamples, _xyz("a", L.a), _xyz("b", L.b), _xyz("c", L.c),
                   _xyz("d", L.d), _rgb("color", L.color)))
    raise ValueError("light type %d" % L.type)


def scene_xml(d, outfile, output_type="tga", render_attrs=None):
    out = ['<?xml version="1.0"?>', "<!-- written by tests/xml_scene.py -->", "<scene>"]
    for m, M in enumerate(d.mats):
        if M.mod_count:
            raise ValueError("modulators are not exported")
        out.append('<shader type="generic" name="M%03d"><attributes>%s</attributes></shader>'
                   % (m, _rgb("color", M.color)))
    for o, flags in enumerate(d.obj_flags):
        sel = np.nonzero(d.tri_obj == o)[0]
        used = np.unique(d.idx[sel].reshape(-1))
        remap = {int(v): k for k, v in enumerate(used)}
        out.append('<object name="O%04d" shader_name="M%03d" shadow="%s" recv_rad="%s" emit_rad="%s">'
                   % (o, int(d.tri_mat[sel[0]]) if len(sel) else 0, _onoff(flags & 1), _onoff(flags & 2),
                      _onoff(flags & 4)))
        out.append('<attributes></attributes><mesh autosmooth="%s"><points>' % _g(d.smooth.get(o, 0)))
        out.extend('<p x="%s" y="%s" z="%s"/>' % tuple(_g(c) for c in d.verts[v]) for v in used)
        out.append("</points><faces>")
        out.extend('<f a="%d" b="%d" c="%d" shader_name="M%03d"/>'
                   % (remap[int(t[0])], remap[int(t[1])], remap[int(t[2])], int(m))
                   for t, m in zip(d.idx[sel], d.tri_mat[sel]))
        out.append("</faces></mesh></object>")
    for k, L in enumerate(d.lights):
        out.append(light_xml(k, L))
    c = d.cam
    out.append('<camera name="cam" resx="%d" resy="%d" focal="%s" aspect_ratio="%s">%s%s%s</camera>'
               % (c.resx, c.resy, _g(c.focal), _g(c.aspect), _xyz("from", c.pos), _xyz("to", c.look),
                  _xyz("up", c.up)))
    o = d.opts
    bg = ""
    if o.has_background:
        out.append('<background type="constant" name="bg">%s</background>' % _rgb("color", o.background))
        bg = ' background_name="bg"'
    attrs = dict(raydepth=o.raydepth, bias=_g(o.self_bias), AA_passes=o.aa_passes,
                 AA_minsamples=o.aa_minsamples, AA_pixelwidth=_g(o.aa_pixelwidth), AA_threshold=_g(o.aa_threshold),
                 AA_jitterfirst=_onoff(o.aa_jitterfirst), exposure=_g(o.exposure), gamma=_g(o.gamma),
                 clamp_rgb=_onoff(o.clamp_rgb))
    attrs.update(render_attrs or {})
    out.append('<render camera_name="cam"%s %s><outfile value="%s"/><output_type value="%s"/></render>'
               % (bg, " ".join('%s="%s"' % kv for kv in attrs.items()), outfile, output_type))
    out.append("</scene>")
    return "\n".join(out) + "\n"


def read_tga(path):
    """Uncompressed 24-bit top-left targa -> (header bytes, [h, w, 3] RGB uint8)."""
    b = open(path, "rb").read()
    w, h = b[12] | (b[13] << 8), b[14] | (b[15] << 8)
    px = np.frombuffer(b[18:18 + 3 * w * h], np.uint8).reshape(h, w, 3)
    return b[:18], px[:, :, ::-1].copy()


def read_hdr(path):
    """Radiance RGBE file (flat or run-length scanlines) -> [h, w, 4] uint8 RGBE."""
    b = open(path, "rb").read()
    head_end = b.index(b"\n\n") + 2
    res_end = b.index(b"\n", head_end)
    ys, h, xs, w = b[head_end:res_end].split()
    assert ys == b"-Y" and xs == b"+X"
    h, w = int(h), int(w)
    pos, img = res_end + 1, np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if b[pos] == 2 and b[pos + 1] == 2 and 8 <= w <= 0x7fff:
            assert (b[pos + 2] << 8 | b[pos + 3]) == w
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = b[pos]
                    pos += 1
                    if n > 128:
                        img[y, x:x + n - 128, c] = b[pos]
                        pos += 1
                        x += n - 128
                    else:
                        img[y, x:x + n, c] = np.frombuffer(b[pos:pos + n], np.uint8)
                        pos += n
                        x += n
        else:
            img[y] = np.frombuffer(b[pos:pos + 4 * w], np.uint8).reshape(w, 4)
            pos += 4 * w
    assert pos == len(b)
    return b[:head_end], img


def tga_bytes(rgb):
    """color.cc:35-40: 8-bit channels of float colours."""
    c = np.asarray(rgb, np.float32)
    v = np.floor(np.float32(255.0) * c).clip(0, 255)
    return np.where(c < 0, 0, np.where(c >= 1, 255, v)).astype(np.uint8)


def rgbe(rgb):
    """HDR_io.cc:79-93: shared exponent of the largest channel."""
    c = np.asarray(rgb, np.float32)
    d = c.max(axis=-1)
    m, e = np.frexp(d.astype(np.float64))
    s = np.where(d > 0, np.ldexp(1.0, 8 - e), 0).astype(np.float32)
    out = np.zeros(c.shape[:-1] + (4,), np.uint8)
    big = d > np.float32(1e-32)
    for k in range(3):
        out[..., k] = np.where(big, np.floor(c[..., k] * s[...]), 0).clip(0, 255).astype(np.uint8)
    out[..., 3] = np.where(big, e + 128, 0).astype(np.uint8)
    return out
