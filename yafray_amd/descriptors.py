"""ctypes mirrors of the POD descriptors in include/yafgpu.h, plus helpers that
fill them with the reference factories' parameter defaults.

Defaults follow the reference factories: pathLight_t::factory
(src/lights/pathlight.cc:753-807), sunLight_t::factory (src/lights/sunlight.cc:59-73),
areaLight_t::factory (src/lights/arealight.cc:257-282), photonLight_t::factory
(src/lights/photonlight.cc:692-732), genericShader_t::factory
(src/shaders/basicshaders.cc:238-367) and the texture factories of
src/shaders/basictex.cc.
"""
import ctypes as C

import numpy as np

F3 = C.c_float * 3


class yg_ray(C.Structure):
    _fields_ = [("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("tmin", C.c_float),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float), ("tmax", C.c_float)]


class yg_hit(C.Structure):
    _fields_ = [("prim", C.c_int32), ("t", C.c_float), ("u", C.c_float), ("v", C.c_float)]


class yg_camera(C.Structure):
    _fields_ = [("pos", F3), ("look", F3), ("up", F3), ("resx", C.c_int32), ("resy", C.c_int32),
                ("aspect", C.c_float), ("focal", C.c_float)]


class yg_texture(C.Structure):
    _fields_ = [("type", C.c_int32), ("noise", C.c_int32), ("color1", F3), ("color2", F3),
                ("depth", C.c_int32), ("size", C.c_float), ("hard", C.c_int32),
                ("bias", C.c_int32), ("turb", C.c_float), ("sharpness", C.c_float),
                ("shape", C.c_int32), ("rings", C.c_int32), ("H", C.c_float),
                ("lacunarity", C.c_float), ("octaves", C.c_float), ("iscale", C.c_float),
                ("mtype", C.c_int32), ("offset", C.c_float), ("gain", C.c_float)]


YG_MUSGRAVE_FBM, YG_MUSGRAVE_MULTIFRACTAL, YG_MUSGRAVE_HETEROTERRAIN, YG_MUSGRAVE_HYBRIDMF, YG_MUSGRAVE_RIDGEDMF = \
    0, 1, 2, 3, 4
MUSGRAVE_TYPES = {"fBm": YG_MUSGRAVE_FBM, "multifractal": YG_MUSGRAVE_MULTIFRACTAL,
                  "heteroterrain": YG_MUSGRAVE_HETEROTERRAIN, "hybridmf": YG_MUSGRAVE_HYBRIDMF,
                  "ridgedmf": YG_MUSGRAVE_RIDGEDMF}


class yg_modulator(C.Structure):
    _fields_ = [("tex", C.c_int32), ("mode", C.c_int32), ("color", C.c_float), ("size", F3),
                ("ofs", F3), ("proj", C.c_int32 * 3), ("has_mtx", C.c_int32), ("mtx", C.c_float * 12),
                ("texco", C.c_int32)]


# modulator texture coordinates (YG_TEXCO_*); TEXCO_NAMES maps the plugin's
# "texco" strings as modulator_t::string2texcotype (texture.h:135-149): any
# other name is "uv"
YG_TEXCO_GLOBAL, YG_TEXCO_ORCO, YG_TEXCO_UV, YG_TEXCO_WIN, YG_TEXCO_NOR, YG_TEXCO_REFL = range(6)
TEXCO_NAMES = {"global": YG_TEXCO_GLOBAL, "orco": YG_TEXCO_ORCO, "uv": YG_TEXCO_UV, "window": YG_TEXCO_WIN,
               "normal": YG_TEXCO_NOR, "reflect": YG_TEXCO_REFL}


def texco_of(name):
    return TEXCO_NAMES.get(name, YG_TEXCO_UV)


class yg_material(C.Structure):
    _fields_ = [("color", F3), ("mod_first", C.c_int32), ("mod_count", C.c_int32)]


class yg_light(C.Structure):
    _fields_ = [("type", C.c_int32), ("use_in_render", C.c_int32),
                ("use_in_indirect", C.c_int32), ("color", F3), ("power", C.c_float),
                ("dir", F3), ("cast_shadows", C.c_int32),
                ("a", F3), ("b", F3), ("c", F3), ("d", F3), ("samples", C.c_int32),
                ("psamples", C.c_int32), ("depth", C.c_int32), ("use_qmc", C.c_int32),
                ("threshold", C.c_float), ("from_", F3), ("to", F3), ("angle", C.c_float),
                ("photons", C.c_int32), ("search", C.c_int32), ("mindepth", C.c_int32),
                ("bias", C.c_float), ("fixedradius", C.c_float), ("cluster", C.c_float),
                ("dummy", C.c_int32), ("cache", C.c_int32), ("cache_size", C.c_float),
                ("shadow_threshold", C.c_float)]


class yg_render_opts(C.Structure):
    _fields_ = [("min_raydis", C.c_float), ("self_bias", C.c_float), ("raydepth", C.c_int32),
                ("background", F3), ("has_background", C.c_int32), ("block", C.c_int32),
                ("rand_seed", C.c_uint32), ("lcg_seed", C.c_uint32), ("flags", C.c_int32),
                ("aa_passes", C.c_int32), ("aa_minsamples", C.c_int32), ("aa_pixelwidth", C.c_float),
                ("aa_threshold", C.c_float), ("aa_jitterfirst", C.c_int32),
                ("tonemap", C.c_int32), ("exposure", C.c_float), ("gamma", C.c_float),
                ("clamp_rgb", C.c_int32)]


class yg_render_stats(C.Structure):
    _fields_ = [("rays_camera", C.c_uint64), ("rays_extension", C.c_uint64),
                ("rays_shadow", C.c_uint64), ("adaptive_extra", C.c_uint64),
                ("lcg_draws", C.c_uint64), ("ms", C.c_float), ("ms_trace", C.c_float),
                ("ms_bounce", C.c_float), ("bounce_launches", C.c_uint32),
                ("rays_bounce", C.c_uint64), ("ms_ext", C.c_float), ("ms_shadow", C.c_float),
                ("graph", C.c_int32)]


class yg_build_opts(C.Structure):
    _fields_ = [("max_leaf", C.c_int32), ("sah_bins", C.c_int32), ("threads", C.c_int32),
                ("width", C.c_int32), ("stack_lds", C.c_int32), ("rays_per_lane", C.c_int32),
                ("compact_below", C.c_float), ("builder", C.c_int32)]


YG_BUILDER_DEVICE = 0
YG_BUILDER_HOST = 1
YG_BUILDER_DEVICE_PLOC = 2


YG_TUNE_MAX = 32


class yg_tune_opts(C.Structure):
    _fields_ = [("widths", C.c_int32), ("leaves", C.c_int32), ("bins", C.c_int32),
                ("repeats", C.c_int32), ("threads", C.c_int32), ("stacks", C.c_int32),
                ("rays", C.c_int32), ("compacts", C.c_int32)]


class yg_tune_result(C.Structure):
    _fields_ = [("best", yg_build_opts), ("best_ms", C.c_float), ("ncand", C.c_int32),
                ("cand", yg_build_opts * YG_TUNE_MAX), ("cand_ms", C.c_float * YG_TUNE_MAX),
                ("cand_build_ms", C.c_float * YG_TUNE_MAX)]


class yg_photon_stats(C.Structure):
    _fields_ = [("emitted", C.c_uint64), ("stored", C.c_uint64), ("clustered", C.c_uint64),
                ("ms", C.c_float), ("rays", C.c_uint64), ("ms_trace", C.c_float), ("ms_irr", C.c_float)]


YG_TEX_CLOUDS, YG_TEX_MARBLE, YG_TEX_WOOD, YG_TEX_MUSGRAVE = 1, 2, 3, 4
YG_MOD_MIX, YG_MOD_MUL, YG_MOD_ADD, YG_MOD_SUB = 0, 1, 2, 3
YG_LIGHT_SUN, YG_LIGHT_AREA, YG_LIGHT_PATH, YG_LIGHT_PHOTON, YG_LIGHT_GLOBALPHOTON = 1, 2, 3, 4, 5
YG_EVAL_NOISE, YG_EVAL_FBM, YG_EVAL_TURBULENCE, YG_EVAL_TEX_FLOAT, YG_EVAL_TEX_COLOR = range(5)

MIN_RAYDIST = 0.00005  # configio.py:95-100


def _f3(v):
    return F3(*[float(x) for x in v])


def camera(pos, look, up, resx, resy, aspect=1.0, focal=1.0):
    return yg_camera(_f3(pos), _f3(look), _f3(up), resx, resy, aspect, focal)


def sunlight(direction, color=(1, 1, 1), power=1.0, cast_shadows=True):
    L = yg_light()
    L.type = YG_LIGHT_SUN
    L.use_in_render = 1
    L.use_in_indirect = 1
    L.dir = _f3(direction)
    L.color = _f3(color)
    L.power = power
    L.cast_shadows = int(cast_shadows)
    return L


def arealight(a, b, c, d, samples=50, color=(1, 1, 1), power=1.0, psamples=0, dummy=False):
    """dummy: shoots global photons only (arealight.cc:98, arealight.h:90-95)."""
    L = yg_light()
    L.type = YG_LIGHT_AREA
    L.use_in_render = 1
    L.use_in_indirect = 1
    L.a, L.b, L.c, L.d = _f3(a), _f3(b), _f3(c), _f3(d)
    L.samples = samples
    L.psamples = psamples
    L.color = _f3(color)
    L.power = power
    L.dummy = int(dummy)
    return L


def globalphotonlight(radius=1.0, depth=2, photons=50000, search=200):
    """globalPhotonLight_t::factory (globalphotonlight.cc:280-294); no
    caustics (caus_depth is not used)."""
    L = yg_light()
    L.type = YG_LIGHT_GLOBALPHOTON
    L.use_in_render = 1
    L.use_in_indirect = 1
    L.fixedradius = radius
    L.depth = depth
    L.photons = photons
    L.search = search
    return L


def pathlight(samples=16, power=1.0, depth=3, use_qmc=False, threshold=0.1, cache=False, cache_size=0.01,
              shadow_threshold=0.3, search=50):
    """pathLight_t::factory (pathlight.cc:756-812): cache parameters are read
    only with cache on, search at least 3."""
    L = yg_light()
    L.type = YG_LIGHT_PATH
    L.use_in_render = 1
    L.use_in_indirect = 0   # pathLight_t::init sets use_in_indirect = false (pathlight.cc:100)
    L.samples = max(1, samples)
    L.power = power
    L.depth = depth
    L.use_qmc = int(use_qmc)
    L.threshold = threshold
    L.cache = int(cache)
    L.cache_size = cache_size
    L.shadow_threshold = shadow_threshold
    L.search = max(3, search)
    return L


def photonlight(frm=(0, 0, 1), to=(0, 0, 0), color=(1, 1, 1), power=1.0, angle=45.0,
                photons=5000, search=50, depth=3, mindepth=1, bias=0.001, fixedradius=1.0,
                cluster=None, use_qmc=False):
    L = yg_light()
    L.type = YG_LIGHT_PHOTON
    L.use_in_render = 1
    L.use_in_indirect = 0   # photonLight_t ctor (photonlight.cc:102)
    L.from_ = _f3(frm)
    L.to = _f3(to)
    L.color = _f3(color)
    L.power = power
    L.angle = angle
    L.photons = photons
    L.search = search
    L.depth = depth
    L.mindepth = mindepth
    L.bias = bias
    L.fixedradius = fixedradius
    L.cluster = fixedradius / 4 if cluster is None else cluster
    L.use_qmc = int(use_qmc)
    return L


def musgrave(color1=(0, 0, 0), color2=(1, 1, 1), H=1.0, lacunarity=2.0, octaves=2.0, size=1.0,
             intensity=1.0, mtype="fBm", offset=1.0, gain=1.0, noise=None):
    """textureMusgrave_t (basictex.cc:657-709); mtype: fBm, multifractal,
    heteroterrain, hybridmf or ridgedmf."""
    T = yg_texture()
    T.type = YG_TEX_MUSGRAVE
    if noise is not None:
        T.noise = noise
    T.color1, T.color2 = _f3(color1), _f3(color2)
    T.H, T.lacunarity, T.octaves, T.size, T.iscale = H, lacunarity, octaves, size, intensity
    T.mtype = MUSGRAVE_TYPES[mtype] if isinstance(mtype, str) else int(mtype)
    T.offset, T.gain = offset, gain
    return T


def clouds(color1=(0, 0, 0), color2=(1, 1, 1), depth=2, size=1.0, hard=False, bias=0):
    T = yg_texture()
    T.type = YG_TEX_CLOUDS
    T.color1, T.color2 = _f3(color1), _f3(color2)
    T.depth, T.size, T.hard, T.bias = depth, size, int(hard), bias
    return T


def marble(color1=(0, 0, 0), color2=(1, 1, 1), depth=2, turbulence=1.0, sharpness=1.0,
           size=1.0, hard=False, shape=0):
    T = yg_texture()
    T.type = YG_TEX_MARBLE
    T.color1, T.color2 = _f3(color1), _f3(color2)
    T.depth, T.turb, T.size, T.hard, T.shape = depth, turbulence, size, int(hard), shape
    T.sharpness = 1.0 / sharpness if sharpness > 1 else 1.0
    return T


def wood(color1=(0, 0, 0), color2=(1, 1, 1), depth=2, turbulence=1.0, size=1.0, hard=False,
         rings=False, shape=0):
    T = yg_texture()
    T.type = YG_TEX_WOOD
    T.color1, T.color2 = _f3(color1), _f3(color2)
    T.depth, T.turb, T.size, T.hard, T.rings, T.shape = depth, turbulence, size, int(hard), int(rings), shape
    return T


def modulator(tex, mode=YG_MOD_MIX, color=0.0, size=1.0, sizexyz=(1, 1, 1), ofs=(0, 0, 0), proj=(0, 0, 0),
              mtx=None, texco=YG_TEXCO_GLOBAL):
    """proj: per output axis 0 = same axis, 1..3 = x/y/z, -1 = none ("n");
    mtx: 3x4 (or 4x4) texture matrix applied to the texture coordinates, or
    None; texco: YG_TEXCO_* or a plugin name ("global", "orco", "uv",
    "window", "normal", "reflect").  The plugin's default is "orco"; this
    constructor's is the hit point (YG_TEXCO_GLOBAL)."""
    M = yg_modulator()
    M.tex = tex
    M.mode = mode
    M.texco = texco_of(texco) if isinstance(texco, str) else int(texco)
    M.color = color
    s = (size, size, size) if size != 1.0 else sizexyz
    M.size = _f3(s)
    M.ofs = _f3(ofs)
    for a in range(3):
        M.proj[a] = int(proj[a])
    if mtx is not None:
        m = np.asarray(mtx, np.float32).reshape(-1)[:12]
        M.has_mtx = 1
        for k in range(12):
            M.mtx[k] = float(m[k])
    return M


def material(color, mod_first=0, mod_count=0):
    return yg_material(_f3(color), mod_first, mod_count)


# yg_render_opts.flags (include/yafgpu.h)
YG_RENDER_TIMING, YG_RENDER_KTIMING, YG_RENDER_REUSE, YG_RENDER_NO_GRAPH, YG_RENDER_ROULETTE = 1, 2, 4, 8, 16


def render_opts(bias=0.1, raydepth=3, background=None, block=64, rand_seed=1, lcg_seed=123212,
                aa_passes=0, aa_minsamples=1, aa_pixelwidth=1.5, aa_threshold=0.05, aa_jitterfirst=False,
                tonemap=False, exposure=0.0, gamma=1.0, clamp_rgb=False, roulette=False):
    """Render parameters with the reference's defaults; the AA values are
    clamped as interface_impl.cc:501-533 clamps them.  roulette: the
    non-reference Russian roulette mode (YG_RENDER_ROULETTE)."""
    o = yg_render_opts()
    if roulette:
        o.flags |= YG_RENDER_ROULETTE
    if aa_passes < 0:
        aa_passes = 1
    if aa_passes and aa_minsamples < 1:
        aa_minsamples = 1
    aa_pixelwidth = min(max(aa_pixelwidth, 1.0), 2.0)
    o.aa_passes = aa_passes
    o.aa_minsamples = aa_minsamples
    o.aa_pixelwidth = aa_pixelwidth
    o.aa_threshold = aa_threshold
    o.aa_jitterfirst = int(bool(aa_jitterfirst))
    o.tonemap = int(bool(tonemap))
    o.exposure = exposure
    o.gamma = gamma
    o.clamp_rgb = int(bool(clamp_rgb))
    o.min_raydis = MIN_RAYDIST
    o.self_bias = bias
    o.raydepth = raydepth
    if background is not None:
        o.background = _f3(background)
        o.has_background = 1
    o.block = block
    o.rand_seed = rand_seed
    o.lcg_seed = lcg_seed
    return o


def carray(ctype, items):
    arr = (ctype * max(len(items), 1))()
    for i, it in enumerate(items):
        arr[i] = it
    if not items:
        return (ctype * 0)()
    return arr


class SceneDesc:
    """Geometry + shading + lights + camera + render options of one scene."""

    def __init__(self, verts, idx, tri_obj, tri_mat, obj_flags, cam, mats, mods, texs, lights,
                 opts, smooth=None, tex=None, orco=None):
        self.verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        self.idx = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
        self.tri_obj = np.ascontiguousarray(tri_obj, np.int32)
        self.tri_mat = np.ascontiguousarray(tri_mat, np.int32)
        self.obj_flags = np.ascontiguousarray(obj_flags, np.int32)
        self.cam = cam
        self.mats = carray(yg_material, mats)
        self.mods = carray(yg_modulator, mods)
        self.texs = carray(yg_texture, texs)
        self.lights = carray(yg_light, lights)
        self.opts = opts
        # autosmooth angle (degrees) per object id; objects absent are flat
        self.smooth = dict(smooth or {})
        # texture-coordinate data per object id: {"uv": [nt_obj, 6] or None,
        # "flags": [nt_obj] (bit0 uv, bit1 vertex colours) or None, "orco":
        # bool (the object's vertices have orco points in self.orco), "mtx":
        # 12 floats (orco matrix) or None}; orco: [nv, 3] orco point per vertex
        self.tex = dict(tex or {})
        self.orco = None if orco is None else np.ascontiguousarray(orco, np.float32).reshape(-1, 3)

    @property
    def n_tris(self):
        return len(self.idx)

    def objects(self):
        """Yield (obj_id, verts, idx, mats, flags) per object in object order."""
        for o in range(len(self.obj_flags)):
            sel = np.nonzero(self.tri_obj == o)[0]
            yield o, sel
