"""ctypes binding of libyafgpu.so (the C ABI declared in include/yafgpu.h).

Device buffers are torch tensors on the current HIP device; their data
pointers cross the ABI.  There is no fallback: if the library is missing,
import fails loudly.
"""
import ctypes as C
import os

import numpy as np

from . import descriptors as D

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libyafgpu.so")
_LIB = None

vp = C.c_void_p
u32 = C.c_uint32
i32 = C.c_int32
f32 = C.c_float


class YgError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError("libyafgpu.so not built: run `python -m yafray_amd.build`")
    L = C.CDLL(LIB_PATH)
    sigs = {
        "yg_abi_version": (C.c_int, []),
        "yg_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
        "yg_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "yg_scene_create": (C.c_int, [C.POINTER(vp)]),
        "yg_scene_destroy": (C.c_int, [vp]),
        "yg_mesh_add": (C.c_int, [vp, vp, u32, vp, u32, vp, i32, C.POINTER(i32)]),
        "yg_mesh_autosmooth": (C.c_int, [vp, i32, f32]),
        "yg_mesh_texcoords": (C.c_int, [vp, i32, vp, vp, vp, vp]),
        "yg_orco_matrix": (C.c_int, [vp, vp, u32, vp, u32, vp]),
        "yg_build": (C.c_int, [vp, vp]),
        "yg_scene_info": (C.c_int, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64), C.POINTER(i32)]),
        "yg_trace_closest": (C.c_int, [vp, vp, u32, vp, vp]),
        "yg_trace_closest_counted": (C.c_int, [vp, vp, u32, vp, vp, vp]),
        "yg_trace_any": (C.c_int, [vp, vp, u32, vp, vp]),
        "yg_trace_any_noalloc": (C.c_int, [vp, vp, u32, vp, vp]),
        "yg_first_hit": (C.c_int, [vp, vp, vp, u32, f32, vp, vp]),
        "yg_is_shadowed": (C.c_int, [vp, vp, vp, vp, u32, i32, f32, f32, vp, vp]),
        "yg_noise_eval": (C.c_int, [i32, vp, vp, u32, vp, vp]),
        "yg_gen_heightfield": (C.c_int, [i32, i32] + [f32] * 7 + [vp, vp]),
        "yg_halton": (C.c_int, [i32, u32, u32, vp, vp]),
        "yg_set_camera": (C.c_int, [vp, vp]),
        "yg_set_shading": (C.c_int, [vp, vp, u32, vp, u32, vp, u32]),
        "yg_set_lights": (C.c_int, [vp, vp, u32]),
        "yg_render": (C.c_int, [vp, vp, vp, vp, vp]),
        "yg_photon_build": (C.c_int, [vp, u32, vp, vp, vp]),
        "yg_photon_read": (C.c_int, [vp, vp, C.c_uint64, C.POINTER(C.c_uint64), vp]),
        "yg_photon_gather": (C.c_int, [vp, vp, vp, vp, vp, vp, u32, vp, vp]),
        "yg_gphoton_build": (C.c_int, [vp, u32, vp, vp, vp]),
        "yg_gphoton_read": (C.c_int, [vp, vp, C.c_uint64, C.POINTER(C.c_uint64), vp]),
        "yg_gphoton_read_photons": (C.c_int, [vp, vp, C.c_uint64, C.POINTER(C.c_uint64), vp]),
        "yg_cache_read": (C.c_int, [vp, vp, C.c_uint64, C.POINTER(C.c_uint64), vp]),
    }
    sigs["yg_render_batch"] = (C.c_int, [C.POINTER(vp), u32, vp, C.POINTER(vp), vp, vp])
    sigs["yg_render_adjoint"] = (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp])
    sigs["yg_autotune"] = (C.c_int, [vp, vp, vp, vp, vp, vp])
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def exported_symbols():
    return lib()


def last_error():
    buf = C.create_string_buffer(1024)
    lib().yg_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc, what=""):
    if rc != 0:
        raise YgError("%s failed (%d): %s" % (what, rc, last_error()))


def _ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return vp(t.data_ptr())


def _stream():
    import torch
    return vp(torch.cuda.current_stream().cuda_stream)


class Scene:
    """A yg_scene handle: meshes (addObject_trimesh), BVH, shading, lights."""

    def __init__(self):
        self.h = vp()
        check(lib().yg_scene_create(C.byref(self.h)), "yg_scene_create")
        self._keep = []

    def close(self):
        if self.h:
            lib().yg_scene_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autosmooth(self, obj, angle):
        """meshObject_t::autoSmooth of object obj (vertex normals; before build)."""
        check(lib().yg_mesh_autosmooth(self.h, int(obj), float(angle)), "yg_mesh_autosmooth")

    def texcoords(self, obj, face_uv=None, face_flags=None, orco=None, orco_mtx=None):
        """Face uv [nt_obj, 6], face flags [nt_obj] (bit0 uv, bit1 vertex
        colours), orco points [nv_obj, 3] and the orco matrix (12 floats) of
        object obj (before build)."""
        def arr(a, dt):
            return None if a is None else np.ascontiguousarray(a, dt).reshape(-1)
        uv, fl, oc, mt = arr(face_uv, np.float32), arr(face_flags, np.uint8), arr(orco, np.float32), arr(orco_mtx, np.float32)
        ptr = lambda a: None if a is None else a.ctypes.data_as(vp)
        check(lib().yg_mesh_texcoords(self.h, int(obj), ptr(uv), ptr(fl), ptr(oc), ptr(mt)), "yg_mesh_texcoords")

    def add_mesh(self, verts, idx, mat=None, flags=7):
        v = np.ascontiguousarray(verts, np.float32).reshape(-1)
        f = np.ascontiguousarray(idx, np.uint32).reshape(-1)
        m = None if mat is None else np.ascontiguousarray(mat, np.int32)
        oid = i32()
        check(lib().yg_mesh_add(self.h, v.ctypes.data_as(vp), len(v) // 3, f.ctypes.data_as(vp),
                                len(f) // 3, None if m is None else m.ctypes.data_as(vp), flags,
                                C.byref(oid)), "yg_mesh_add")
        return oid.value

    def build(self, max_leaf=4, bins=32, threads=0, width=0, stack_lds=0, rays_per_lane=0, compact_below=0.0,
              builder=0):
        """stack_lds: stack entries per lane the render launches keep in LDS
        (0 = library default, -1 = the whole stack); rays_per_lane: rays each
        lane of the GI level launches traces (0 = 1); compact_below: survivor
        fraction below which a GI level's paths are compacted (0 = 1, always;
        -1 = never); builder: D.YG_BUILDER_DEVICE (default: binned SAH on
        the GPU), D.YG_BUILDER_HOST (the same binned SAH on host threads) or
        D.YG_BUILDER_DEVICE_PLOC (clustering on the GPU)."""
        o = D.yg_build_opts(max_leaf, bins, threads, width, stack_lds, rays_per_lane, compact_below, builder)
        check(lib().yg_build(self.h, C.byref(o)), "yg_build")

    def info(self):
        nt, nn, nb = C.c_uint64(), C.c_uint64(), C.c_uint64()
        md = i32()
        check(lib().yg_scene_info(self.h, C.byref(nt), C.byref(nn), C.byref(nb), C.byref(md)),
              "yg_scene_info")
        return {"tris": nt.value, "nodes": nn.value, "dev_bytes": nb.value, "max_depth": md.value}

    @classmethod
    def from_desc(cls, desc, build=True, **kw):
        """Ingest a SceneDesc object by object (object order = id order)."""
        s = cls()
        order = np.argsort(desc.tri_obj, kind="stable")
        bounds = np.searchsorted(desc.tri_obj[order], np.arange(len(desc.obj_flags) + 1))
        contiguous = bool(np.all(order == np.arange(len(order))))
        for o in range(len(desc.obj_flags)):
            sel = order[bounds[o]:bounds[o + 1]]
            tri = desc.idx[sel]
            if contiguous and len(tri):
                lo, hi = int(tri.min()), int(tri.max()) + 1
                v, f = desc.verts[lo:hi], (tri - lo).astype(np.uint32)
            else:
                used, inv = np.unique(tri.reshape(-1), return_inverse=True)
                v, f = desc.verts[used], inv.reshape(-1, 3).astype(np.uint32)
            oid = s.add_mesh(v, f, desc.tri_mat[sel], int(desc.obj_flags[o]))
            if desc.smooth.get(o, 0) >= 1:
                s.autosmooth(oid, desc.smooth[o])
            t = desc.tex.get(o)
            if t:
                oc = None
                if t.get("orco"):
                    oc = desc.orco[lo:hi] if (contiguous and len(tri)) else desc.orco[used]
                s.texcoords(oid, t.get("uv"), t.get("flags"), oc, t.get("mtx"))
        if build:
            s.build(**kw)
        s.set_desc(desc)
        return s

    def set_desc(self, desc):
        L = lib()
        check(L.yg_set_camera(self.h, C.byref(desc.cam)), "yg_set_camera")
        check(L.yg_set_shading(self.h, C.cast(desc.mats, vp), len(desc.mats),
                               C.cast(desc.mods, vp) if len(desc.mods) else None, len(desc.mods),
                               C.cast(desc.texs, vp) if len(desc.texs) else None, len(desc.texs)),
              "yg_set_shading")
        check(L.yg_set_lights(self.h, C.cast(desc.lights, vp) if len(desc.lights) else None,
                              len(desc.lights)), "yg_set_lights")
        self.desc = desc

    # --- queries on device tensors ---
    def trace_closest(self, rays, hits, counters=None):
        n = rays.shape[0]
        if counters is None:
            check(lib().yg_trace_closest(self.h, _ptr(rays), n, _ptr(hits), _stream()),
                  "yg_trace_closest")
        else:
            check(lib().yg_trace_closest_counted(self.h, _ptr(rays), n, _ptr(hits),
                                                 _ptr(counters), _stream()),
                  "yg_trace_closest_counted")

    def trace_any(self, rays, occ):
        check(lib().yg_trace_any(self.h, _ptr(rays), rays.shape[0], _ptr(occ), _stream()),
              "yg_trace_any")

    def trace_any_noalloc(self, rays, occ):
        check(lib().yg_trace_any_noalloc(self.h, _ptr(rays), rays.shape[0], _ptr(occ), _stream()),
              "yg_trace_any_noalloc")

    def first_hit(self, frm, d, hits, min_raydis=D.MIN_RAYDIST):
        check(lib().yg_first_hit(self.h, _ptr(frm), _ptr(d), frm.shape[0], min_raydis, _ptr(hits),
                                 _stream()), "yg_first_hit")

    def is_shadowed(self, p, l, self_obj, out, to_point, min_raydis=D.MIN_RAYDIST, self_bias=0.1):
        check(lib().yg_is_shadowed(self.h, _ptr(p), _ptr(l), _ptr(self_obj), p.shape[0],
                                   int(to_point), min_raydis, self_bias, _ptr(out), _stream()),
              "yg_is_shadowed")

    def render(self, out_rgb, opts=None, timing=False):
        o = opts if opts is not None else self.desc.opts
        if timing:
            o = D.yg_render_opts.from_buffer_copy(o)
            o.flags |= 1
        st = D.yg_render_stats()
        check(lib().yg_render(self.h, C.byref(o), _ptr(out_rgb), C.byref(st), _stream()),
              "yg_render")
        return st

    def autotune(self, out_rgb, opts=None, widths=(2, 4, 8), leaves=(2, 4, 8), bins=(32,), repeats=2,
                 threads=0, stacks=(), rays=(1,), compacts=(1.0,)):
        """Rebuild with every (width, max_leaf, bins) candidate and try every
        LDS stack part in `stacks` (entries, -1 = whole stack; () = the
        default), rays per lane in `rays` and compaction threshold in
        `compacts` (-1 = never), time the frame of `opts` into out_rgb, keep
        the fastest.  Returns (best yg_build_opts, best ms,
        [(width, leaf, bins, ms, build_ms, stack_lds, rays_per_lane, compact_below), ...])."""
        o = opts if opts is not None else self.desc.opts
        t = D.yg_tune_opts()
        t.widths = sum(1 << w for w in widths)
        t.leaves = sum(1 << n for n in leaves)
        t.bins = sum(1 << (int(b).bit_length() - 1) for b in bins)
        if any(b & (b - 1) for b in bins):
            raise ValueError("SAH bin counts must be powers of two")
        t.repeats = repeats
        t.threads = threads
        for v in stacks:
            if v == -1:
                t.stacks |= 1
            elif v in (8, 16, 32, 64):
                t.stacks |= 1 << (v.bit_length() - 1)
            else:
                raise ValueError("stack_lds candidates: -1, 8, 16, 32, 64")
        for v in rays:
            if v not in (1, 2, 4):
                raise ValueError("rays_per_lane candidates: 1, 2, 4")
            t.rays |= 1 << (v.bit_length() - 1)
        for v in compacts:
            if v == -1:
                t.compacts |= 1
            elif v in (0.25, 0.5, 0.75, 1.0):
                t.compacts |= 1 << int(round(v * 4))
            else:
                raise ValueError("compact_below candidates: -1, 0.25, 0.5, 0.75, 1")
        if tuple(compacts) == (1.0,):
            t.compacts = 0
        r = D.yg_tune_result()
        check(lib().yg_autotune(self.h, C.byref(o), _ptr(out_rgb), C.byref(t), C.byref(r), _stream()),
              "yg_autotune")
        cands = [(r.cand[k].width, r.cand[k].max_leaf, r.cand[k].sah_bins, r.cand_ms[k], r.cand_build_ms[k],
                  r.cand[k].stack_lds, r.cand[k].rays_per_lane, r.cand[k].compact_below) for k in range(r.ncand)]
        return r.best, r.best_ms, cands

    def render_adjoint(self, grad_rgb, out_rgb, gmat, gmod, glight, opts=None):
        """Forward render into out_rgb and dLoss/dparams from dLoss/dpixel."""
        o = opts if opts is not None else self.desc.opts
        st = D.yg_render_stats()
        check(lib().yg_render_adjoint(self.h, C.byref(o), _ptr(grad_rgb), _ptr(out_rgb), _ptr(gmat),
                                      _ptr(gmod), _ptr(glight), C.byref(st), _stream()),
              "yg_render_adjoint")
        return st

    def photon_build(self, li, opts=None):
        o = opts if opts is not None else self.desc.opts
        st = D.yg_photon_stats()
        check(lib().yg_photon_build(self.h, li, C.byref(o), C.byref(st), _stream()),
              "yg_photon_build")
        return st

    def gphoton_build(self, li, opts=None):
        """globalPhotonLight_t::init on the device for light li (yg_gphoton_build)."""
        o = opts if opts is not None else self.desc.opts
        st = D.yg_photon_stats()
        check(lib().yg_gphoton_build(self.h, li, C.byref(o), C.byref(st), _stream()), "yg_gphoton_build")
        return st

    def gphoton_read(self, out=None):
        """The irradiance boxes into a float32 CUDA tensor of 15 floats each;
        returns the box count (out=None: count only)."""
        cnt = C.c_uint64()
        check(lib().yg_gphoton_read(self.h, _ptr(out) if out is not None else None,
                                    out.numel() // 15 if out is not None else 0, C.byref(cnt), _stream()),
              "yg_gphoton_read")
        return cnt.value

    def gphoton_read_photons(self, out=None):
        """The raw photon map in store order into an int32 CUDA tensor of 5
        words per photon (position bits, direction code, RGBE); returns the
        photon count (out=None: count only)."""
        cnt = C.c_uint64()
        check(lib().yg_gphoton_read_photons(self.h, _ptr(out) if out is not None else None,
                                            out.numel() // 5 if out is not None else 0, C.byref(cnt), _stream()),
              "yg_gphoton_read_photons")
        return cnt.value

    def cache_read(self):
        """The irradiance cache of this scene's last cache render: [n, 18]
        numpy array (P, N, polar position, colour, adist, M, precision,
        rank, pixel, devaluated) in insertion order."""
        import numpy as np
        import torch
        cnt = C.c_uint64()
        check(lib().yg_cache_read(self.h, None, 0, C.byref(cnt), _stream()), "yg_cache_read")
        n = cnt.value
        out = torch.zeros(max(n, 1) * 18, dtype=torch.float32, device="cuda")
        check(lib().yg_cache_read(self.h, _ptr(out), n, C.byref(cnt), _stream()), "yg_cache_read")
        return out.cpu().numpy()[: n * 18].reshape(n, 18).astype(np.float32)

    def photon_read(self, out):
        cnt = C.c_uint64()
        check(lib().yg_photon_read(self.h, _ptr(out), out.numel() // 9 if out is not None else 0,
                                   C.byref(cnt), _stream()), "yg_photon_read")
        return cnt.value

    def photon_gather(self, p, n, ng, eye, mat, out):
        check(lib().yg_photon_gather(self.h, _ptr(p), _ptr(n), _ptr(ng), _ptr(eye), _ptr(mat),
                                     mat.shape[0], _ptr(out), _stream()), "yg_photon_gather")


def orco_matrix(verts, idx, m16=None):
    """meshObject_t::transform's orco matrix, rows 0..2 (yg_orco_matrix): the
    mesh's untransformed vertices [nv, 3], triangles [nt, 3], object matrix
    m16 (4x4 row-major, None = identity)."""
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    f = np.ascontiguousarray(idx, np.uint32).reshape(-1)
    m = None if m16 is None else np.ascontiguousarray(m16, np.float32).reshape(-1)
    out = np.zeros(12, np.float32)
    check(lib().yg_orco_matrix(None if m is None else m.ctypes.data_as(vp), v.ctypes.data_as(vp), len(v) // 3,
                               f.ctypes.data_as(vp), len(f) // 3, out.ctypes.data_as(vp)), "yg_orco_matrix")
    return out.reshape(3, 4)


def render_batch(scenes, outs, opts=None, timing=False):
    """One batched launch over several Scene objects; outs[k] gets scene k's image."""
    n = len(scenes)
    arr = (vp * n)(*[s.h for s in scenes])
    optarr = (D.yg_render_opts * n)()
    for k, s in enumerate(scenes):
        o = D.yg_render_opts.from_buffer_copy(opts[k] if opts else s.desc.opts)
        if timing:
            o.flags |= 1
        optarr[k] = o
    outp = (vp * n)(*[vp(t.data_ptr()) for t in outs])
    st = D.yg_render_stats()
    check(lib().yg_render_batch(arr, n, optarr, outp, C.byref(st), _stream()), "yg_render_batch")
    return st


def noise_eval(what, tex, pts, out):
    check(lib().yg_noise_eval(what, None if tex is None else C.byref(tex), _ptr(pts),
                              pts.shape[0], _ptr(out), _stream()), "yg_noise_eval")


def gen_heightfield(nx, ny, ex, ey, amp, freq, H, lac, octv, out):
    check(lib().yg_gen_heightfield(nx, ny, ex, ey, amp, freq, H, lac, octv, _ptr(out), _stream()),
          "yg_gen_heightfield")


def halton(base, i0, n, out):
    check(lib().yg_halton(base, i0, n, _ptr(out), _stream()), "yg_halton")


def device_heights(nx, ny, ex, ey, amp, freq, H, lac, octv):
    """Heightfield vertices computed on the device, returned as numpy."""
    import torch
    out = torch.empty(nx * ny * 3, dtype=torch.float32, device="cuda")
    gen_heightfield(nx, ny, ex, ey, amp, freq, H, lac, octv, out)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(-1, 3)
