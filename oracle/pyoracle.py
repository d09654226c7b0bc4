"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_LIB_FAST = None

F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
U32P = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
U8P = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
U64P = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def lib(fast=False):
    """The oracle library.  fast=True: the same restatement built with the
    reference's release flags (-O3 -ffast-math), for CPU-baseline timing
    only; the checker is always the strict build."""
    global _LIB, _LIB_FAST
    if fast:
        if _LIB_FAST is None:
            _LIB_FAST = _load(os.path.join(_HERE, "liboracle_fast.so"))
        return _LIB_FAST
    if _LIB is None:
        _LIB = _load(os.path.join(_HERE, "liboracle.so"))
    return _LIB


def _load(path):
    if True:
        if not os.path.exists(path):
            raise RuntimeError("%s is not built (run __graft_entry__.build())" % path)
        L = C.CDLL(path)
        vp = C.c_void_p
        L.yo_scene_new.restype = vp
        L.yo_scene_new.argtypes = [F32P, C.c_int64, U32P, C.c_int64, vp, vp]
        L.yo_log_begin.argtypes = [vp, C.c_int64, vp, C.c_int64]
        L.yo_log_end.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.yo_scene_free.argtypes = [vp]
        L.yo_mesh_autosmooth.restype = C.c_int
        L.yo_mesh_autosmooth.argtypes = [vp, C.c_int, C.c_float]
        L.yo_mesh_texcoords.restype = C.c_int
        L.yo_mesh_texcoords.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        L.yo_orco_matrix.restype = C.c_int
        L.yo_orco_matrix.argtypes = [vp, vp, C.c_int64, vp, C.c_int64, vp]
        L.yo_shading_normals.argtypes = [vp, I32P, F32P, C.c_int64, F32P]
        L.yo_scene_nodes.restype = C.c_int64
        L.yo_scene_nodes.argtypes = [vp]
        L.yo_trace_brute.argtypes = [vp, F32P, C.c_int64, C.c_int, I32P, F32P, F32P, F32P]
        L.yo_trace_bvh.argtypes = [vp, F32P, C.c_int64, C.c_int, I32P, F32P, F32P, F32P,
                                   U64P, U64P, C.c_int]
        L.yo_first_hit.argtypes = [vp, F32P, F32P, C.c_int64, C.c_float, I32P, F32P]
        L.yo_is_shadowed.argtypes = [vp, F32P, F32P, vp, C.c_int64, C.c_int, C.c_float,
                                     C.c_float, U8P]
        L.yo_halton_seq.argtypes = [C.c_int, C.c_int64, F32P]
        L.yo_halton_at.restype = C.c_float
        L.yo_halton_at.argtypes = [C.c_int, C.c_uint32]
        for n in ("yo_ri_vdc", "yo_ri_s", "yo_ri_lp"):
            getattr(L, n).restype = C.c_float
            getattr(L, n).argtypes = [C.c_uint32, C.c_uint32]
        L.yo_next_prime.restype = C.c_int
        L.yo_next_prime.argtypes = [C.c_int]
        L.yo_our_random.restype = C.c_float
        L.yo_our_random.argtypes = [C.POINTER(C.c_int32)]
        L.yo_lcg_jump.restype = C.c_int32
        L.yo_lcg_jump.argtypes = [C.c_int32, C.c_uint64]
        L.yo_glibc_rand.argtypes = [C.c_uint32, C.c_int64, I32P]
        L.yo_newperlin.restype = C.c_float
        L.yo_newperlin.argtypes = [C.c_float] * 3
        for f in ("yo_stdperlin", "yo_blendernoise"):
            getattr(L, f).restype = C.c_float
            getattr(L, f).argtypes = [C.c_float] * 3
        L.yo_grad_table.argtypes = [F32P]
        L.yo_fbm.restype = C.c_float
        L.yo_fbm.argtypes = [C.c_float] * 6
        L.yo_turbulence.restype = C.c_float
        L.yo_turbulence.argtypes = [C.c_float] * 3 + [C.c_int, C.c_float, C.c_int]
        L.yo_noise_eval.restype = C.c_int
        L.yo_noise_eval.argtypes = [C.c_int, vp, F32P, C.c_int64, F32P]
        L.yo_canonical_counts.restype = C.c_int64
        L.yo_canonical_counts.argtypes = [F32P, C.c_int64, U32P, C.c_int64, F32P, C.c_int64, C.c_int, C.c_int,
                                          U64P, U64P, U64P]
        L.yo_diffuse_at.argtypes = [vp, vp, vp, C.c_int, F32P, C.c_int64, F32P]
        L.yo_gen_heightfield.argtypes = [C.c_int, C.c_int] + [C.c_float] * 7 + [F32P]
        L.yo_tri_hit.restype = C.c_int
        L.yo_tri_hit.argtypes = [F32P] * 6
        L.yo_tri_intersect.restype = C.c_float
        L.yo_tri_intersect.argtypes = [F32P] * 4
        L.yo_tri_normal.argtypes = [F32P] * 4
        L.yo_render.restype = C.c_int
        L.yo_render.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_int,
                                vp, C.c_int, vp, F32P, U64P]
        L.yo_render_mt.restype = C.c_int
        L.yo_render_mt.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_int,
                                   vp, C.c_int, vp, F32P, U64P, C.c_int]
        L.yo_trace_camera.restype = C.c_int
        L.yo_trace_camera.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp, C.c_int, vp, C.c_int, vp,
                                      F32P, F32P, C.c_int64, F32P]
        L.yo_block_regions_into.restype = C.c_int
        L.yo_block_regions_into.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint32, I32P, C.c_int]
        L.yo_render_adjoint.restype = C.c_int
        L.yo_render_adjoint.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_int,
                                        vp, C.c_int, vp, F32P, F32P, vp, vp, vp, C.c_int]
        L.yo_render_fp64.restype = C.c_int
        L.yo_render_fp64.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp, C.c_int, vp, C.c_int, vp,
                                     C.c_int, C.c_int, vp, vp, C.c_int64, C.c_uint64, vp, C.c_int]
        L.yo_photon_build.restype = C.c_int
        L.yo_photon_build.argtypes = [vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_int,
                                      C.c_int, vp, C.c_int, vp, U64P]
        L.yo_photon_load.restype = C.c_int
        L.yo_photon_load.argtypes = [vp, F32P, C.c_int64, vp]
        L.yo_photon_read.restype = C.c_int64
        L.yo_photon_read.argtypes = [vp, vp, C.c_int64]
        L.yo_photon_gather.argtypes = [vp, vp, vp, vp, vp, F32P, F32P, F32P, F32P, I32P,
                                       C.c_int64, F32P]
        L.yo_gphoton_build.restype = C.c_int
        L.yo_gphoton_build.argtypes = [vp, vp, C.c_int, vp, vp, vp, C.c_int, vp, C.c_int, vp, C.c_int,
                                       C.c_int, U64P]
        L.yo_gphoton_read.restype = C.c_int64
        L.yo_gphoton_read.argtypes = [vp, vp, C.c_int64]
        L.yo_gphoton_load.restype = C.c_int
        L.yo_gphoton_load.argtypes = [vp, F32P, C.c_int64, vp, C.c_int, C.c_int, C.c_int]
        L.yo_gphoton_read_photons.restype = C.c_int64
        L.yo_gphoton_read_photons.argtypes = [vp, vp, C.c_int64]
        L.yo_gphoton_load_photons.restype = C.c_int
        L.yo_gphoton_load_photons.argtypes = [vp, vp, C.c_int64]
        L.yo_psamp_plan.restype = C.c_int
        L.yo_psamp_plan.argtypes = [vp, vp, vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp]
        L.yo_cache_read.restype = C.c_int64
        L.yo_cache_read.argtypes = [vp, vp, C.c_int64, U64P]
        L.yo_cache_record.argtypes = [vp, C.c_int]
        L.yo_cache_visits.restype = C.c_int64
        L.yo_cache_visits.argtypes = [vp, vp, C.c_int64]
        L.yo_grand_stream.argtypes = [C.c_uint32, C.c_int64, C.c_int64, I32P]
        L.yo_heap_offer_trace.restype = C.c_int64
        L.yo_heap_offer_trace.argtypes = [F32P, C.c_int64, C.c_uint, F32P]
        L.yo_rgbe_roundtrip.argtypes = [F32P, F32P, C.c_int64]
        L.yo_dir_roundtrip.argtypes = [F32P, F32P, vp, C.c_int64]
        L.yo_dirconv_tables.argtypes = [F32P]
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def heap_offer_trace(weights, limit):
    """The oracle's found heap after offering weights[i] (id i) in order, as
    lightCache_t::gatherSamples offers candidates: [(id, weight)] in array order."""
    w = np.ascontiguousarray(weights, np.float32)
    out = np.zeros(2 * (limit + 2), np.float32)
    n = lib().yo_heap_offer_trace(w, len(w), limit, out)
    return [(int(out[2 * i]), float(out[2 * i + 1])) for i in range(n)]


def orco_matrix(verts, idx, m16=None, fast=False):
    """meshObject_t::transform's orco matrix (rows 0..2) of a mesh with
    untransformed vertices verts [nv, 3], triangles idx [nt, 3] and object
    matrix m16 (4x4, None = identity)."""
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    f = np.ascontiguousarray(idx, np.uint32).reshape(-1)
    m = None if m16 is None else np.ascontiguousarray(m16, np.float32).reshape(-1)
    out = np.zeros(12, np.float32)
    vp = C.c_void_p
    rc = lib(fast).yo_orco_matrix(None if m is None else m.ctypes.data_as(vp), v.ctypes.data_as(vp), len(v) // 3,
                                  f.ctypes.data_as(vp), len(f) // 3, out.ctypes.data_as(vp))
    if rc:
        raise RuntimeError("yo_orco_matrix failed (%d)" % rc)
    return out.reshape(3, 4)


class OracleScene:
    """Triangle soup with per-triangle object and material ids."""

    @classmethod
    def from_desc(cls, desc, fast=False):
        """The scene of a SceneDesc, its autosmooth angles applied."""
        oc = cls(desc.verts, desc.idx, desc.tri_obj, desc.tri_mat, fast=fast)
        for o, a in getattr(desc, "smooth", {}).items():
            oc.autosmooth(o, a)
        for o, t in getattr(desc, "tex", {}).items():
            sel = np.nonzero(desc.tri_obj == o)[0]
            corner = desc.orco[desc.idx[sel]].reshape(-1, 9) if t.get("orco") else None
            oc.texcoords(o, t.get("uv"), t.get("flags"), corner, t.get("mtx"))
        return oc

    def __init__(self, verts, idx, tri_obj=None, tri_mat=None, fast=False):
        self.fast = fast
        L = lib(fast)
        self.verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
        self.idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3)
        self.tri_obj = None if tri_obj is None else np.ascontiguousarray(tri_obj, dtype=np.int32)
        self.tri_mat = None if tri_mat is None else np.ascontiguousarray(tri_mat, dtype=np.int32)
        self.h = L.yo_scene_new(self.verts.reshape(-1), len(self.verts), self.idx.reshape(-1),
                                len(self.idx), _ptr(self.tri_obj), _ptr(self.tri_mat))

    def autosmooth(self, obj, angle):
        """meshObject_t::autoSmooth of object obj (vertex normals)."""
        lib(self.fast).yo_mesh_autosmooth(self.h, int(obj), float(angle))

    def texcoords(self, obj, face_uv=None, face_flags=None, corner_orco=None, orco_mtx=None):
        """Texture-coordinate data of object obj: face uv [nt_obj, 6], face
        flags [nt_obj], corner orco points [nt_obj, 9], orco matrix (12)."""
        def arr(a, dt):
            return None if a is None else np.ascontiguousarray(a, dt).reshape(-1)
        self._tex_keep = getattr(self, "_tex_keep", [])
        a = [arr(face_uv, np.float32), arr(face_flags, np.uint8), arr(corner_orco, np.float32),
             arr(orco_mtx, np.float32)]
        self._tex_keep.append(a)
        p = [None if x is None else x.ctypes.data_as(C.c_void_p) for x in a]
        rc = lib(self.fast).yo_mesh_texcoords(self.h, int(obj), *p)
        if rc:
            raise RuntimeError("yo_mesh_texcoords failed (%d)" % rc)

    def shading_normals(self, prim, h):
        """getSurface's shading normals at hit points h [n, 3] of triangles prim."""
        prim = np.ascontiguousarray(prim, np.int32)
        h = np.ascontiguousarray(h, np.float32).reshape(-1)
        out = np.zeros(len(prim) * 3, np.float32)
        lib(self.fast).yo_shading_normals(self.h, prim, h, len(prim), out)
        return out.reshape(-1, 3)

    def close(self):
        if self.h:
            lib(self.fast).yo_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_nodes(self):
        return lib(self.fast).yo_scene_nodes(self.h)

    def trace(self, rays, mode=0, brute=False, threads=1):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = len(rays)
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        u = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        if brute:
            lib(self.fast).yo_trace_brute(self.h, rays.reshape(-1), n, mode, prim, t, u, v)
            return prim, t, u, v
        nodes = np.zeros(1, np.uint64)
        tris = np.zeros(1, np.uint64)
        lib(self.fast).yo_trace_bvh(self.h, rays.reshape(-1), n, mode, prim, t, u, v, nodes, tris, threads)
        self.last_counts = (int(nodes[0]), int(tris[0]))
        return prim, t, u, v

    def first_hit(self, frm, d, min_raydis=5e-5):
        frm = np.ascontiguousarray(frm, np.float32).reshape(-1)
        d = np.ascontiguousarray(d, np.float32).reshape(-1)
        n = len(frm) // 3
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        lib(self.fast).yo_first_hit(self.h, frm, d, n, min_raydis, prim, t)
        return prim, t

    def is_shadowed(self, p, l, self_obj, to_point, min_raydis=5e-5, self_bias=0.1):
        p = np.ascontiguousarray(p, np.float32).reshape(-1)
        l = np.ascontiguousarray(l, np.float32).reshape(-1)
        so = None if self_obj is None else np.ascontiguousarray(self_obj, np.int32)
        n = len(p) // 3
        out = np.empty(n, np.uint8)
        lib(self.fast).yo_is_shadowed(self.h, p, l, _ptr(so), n, int(to_point), min_raydis, self_bias, out)
        return out

    def render(self, desc, threads=1):
        """desc: yafray_amd.descriptors.SceneDesc (its ctypes arrays)."""
        L = lib(self.fast)
        rgb = np.zeros(desc.cam.resx * desc.cam.resy * 3, np.float32)
        stats = np.zeros(5, np.uint64)
        rc = L.yo_render_mt(self.h, C.addressof(desc.cam), _addr(desc.mats), len(desc.mats),
                            _addr(desc.mods), len(desc.mods), _addr(desc.texs), len(desc.texs),
                            _addr(desc.lights), len(desc.lights), _ptr(desc.obj_flags),
                            len(desc.obj_flags), C.addressof(desc.opts), rgb, stats, threads)
        if rc != 0:
            raise RuntimeError("yo_render failed: %d" % rc)
        return rgb.reshape(desc.cam.resy, desc.cam.resx, 3), stats

    def trace_camera(self, desc, px, py):
        """Colours of camera rays through camera-plane points (px, py) (sun
        lights only): [n, 3] float32."""
        px = np.ascontiguousarray(px, np.float32)
        py = np.ascontiguousarray(py, np.float32)
        rgb = np.zeros((len(px), 3), np.float32)
        rc = lib(self.fast).yo_trace_camera(self.h, C.addressof(desc.cam), _addr(desc.mats), len(desc.mats),
                                   _addr(desc.mods), _addr(desc.texs), _addr(desc.lights), len(desc.lights),
                                   _ptr(desc.obj_flags), len(desc.obj_flags), C.addressof(desc.opts),
                                   px, py, len(px), rgb)
        if rc != 0:
            raise RuntimeError("yo_trace_camera failed: %d" % rc)
        return rgb

    def render_logged(self, desc, threads=1, ext_cap=0, sh_cap=0):
        """render() that also returns every first-hit and shadow query the
        render made: (img, stats, ext[n,8], sh[n,8]) with the record layouts
        of yo_log_begin (int fields are int32 bit patterns in the floats)."""
        L = lib(self.fast)
        ext = np.zeros((max(ext_cap, 1), 8), np.float32)
        sh = np.zeros((max(sh_cap, 1), 8), np.float32)
        L.yo_log_begin(_ptr(ext), ext_cap, _ptr(sh), sh_cap)
        try:
            img, stats = self.render(desc, threads)
        finally:
            ne, ns = C.c_int64(), C.c_int64()
            L.yo_log_end(C.byref(ne), C.byref(ns))
        if ne.value > ext_cap or ns.value > sh_cap:
            raise RuntimeError("query log overflow: %d/%d first-hit, %d/%d shadow"
                               % (ne.value, ext_cap, ns.value, sh_cap))
        return img, stats, ext[:ne.value], sh[:ns.value]

    def render_adjoint(self, desc, grad_rgb, threads=1):
        """FP64 reference gradient: returns (image, gmat[nmat,3], gmod[nmod,2], glight[nl])."""
        L = lib(self.fast)
        rgb = np.zeros(desc.cam.resx * desc.cam.resy * 3, np.float32)
        g = np.ascontiguousarray(grad_rgb, np.float32).reshape(-1)
        gmat = np.zeros(len(desc.mats) * 3, np.float64)
        gmod = np.zeros(max(len(desc.mods), 1) * 2, np.float64)
        glight = np.zeros(len(desc.lights), np.float64)
        rc = L.yo_render_adjoint(self.h, C.addressof(desc.cam), _addr(desc.mats), len(desc.mats),
                                 _addr(desc.mods), len(desc.mods), _addr(desc.texs), len(desc.texs),
                                 _addr(desc.lights), len(desc.lights), _ptr(desc.obj_flags),
                                 len(desc.obj_flags), C.addressof(desc.opts), g, rgb, _ptr(gmat),
                                 _ptr(gmod), _ptr(glight), int(threads))
        if rc != 0:
            raise RuntimeError("yo_render_adjoint failed: %d" % rc)
        return (rgb.reshape(desc.cam.resy, desc.cam.resx, 3), gmat.reshape(-1, 3),
                gmod[: 2 * len(desc.mods)].reshape(-1, 2), glight)

    def render_fp64(self, desc, samples=None, px=None, py=None, seed=1, threads=1):
        """FP64 render through pixel centres.  samples=None: the whole image
        with the reference's sample streams, [resy, resx, 3] float64.  Else
        the pixels (px[q], py[q]) with `samples` samples of every sampled
        light (the converged estimate), [n, 3] float64."""
        L = lib(self.fast)
        args = (self.h, C.addressof(desc.cam), _addr(desc.mats), len(desc.mats), _addr(desc.mods),
                _addr(desc.texs), _addr(desc.lights), len(desc.lights), _ptr(desc.obj_flags),
                len(desc.obj_flags), C.addressof(desc.opts))
        if samples is None:
            rgb = np.zeros(desc.cam.resx * desc.cam.resy * 3, np.float64)
            rc = L.yo_render_fp64(*args, 0, 0, None, None, 0, 0, _ptr(rgb), threads)
            out = rgb.reshape(desc.cam.resy, desc.cam.resx, 3)
        else:
            px = np.ascontiguousarray(px, np.int32)
            py = np.ascontiguousarray(py, np.int32)
            rgb = np.zeros(len(px) * 3, np.float64)
            rc = L.yo_render_fp64(*args, 1, int(samples), _ptr(px), _ptr(py), len(px), seed,
                                  _ptr(rgb), threads)
            out = rgb.reshape(-1, 3)
        if rc != 0:
            raise RuntimeError("yo_render_fp64 failed: %d" % rc)
        return out

    def photon_build(self, desc, li):
        L = lib(self.fast)
        stats = np.zeros(5, np.uint64)
        rc = L.yo_photon_build(self.h, _addr(desc.mats), len(desc.mats), _addr(desc.mods),
                               len(desc.mods), _addr(desc.texs), len(desc.texs),
                               _addr(desc.lights), len(desc.lights), li, _ptr(desc.obj_flags),
                               len(desc.obj_flags), C.addressof(desc.opts), stats)
        if rc != 0:
            raise RuntimeError("yo_photon_build failed: %d" % rc)
        return stats

    def photon_load(self, desc, li, marks):
        """Use `marks` ([n, 9]: dir, pos, colour) as light li's clustered map."""
        m = np.ascontiguousarray(marks, np.float32).reshape(-1)
        rc = lib(self.fast).yo_photon_load(self.h, m, len(m) // 9, C.addressof(desc.lights[li]))
        if rc != 0:
            raise RuntimeError("yo_photon_load failed")

    def photon_read(self):
        L = lib(self.fast)
        n = L.yo_photon_read(self.h, None, 0)
        out = np.zeros(max(n, 1) * 9, np.float32)
        L.yo_photon_read(self.h, _ptr(out), n)
        return out[: n * 9].reshape(n, 9)

    def gphoton_build(self, desc, li, stateless=False):
        """globalPhotonLight_t::init for light li (serial, from opts.lcg_seed);
        returns [emitted, stored, boxes, rays, draws]."""
        L = lib(self.fast)
        stats = np.zeros(5, np.uint64)
        rc = L.yo_gphoton_build(self.h, _addr(desc.mats), len(desc.mats), _addr(desc.mods), _addr(desc.texs),
                                _addr(desc.lights), len(desc.lights), _ptr(desc.obj_flags), len(desc.obj_flags),
                                C.addressof(desc.opts), li, int(stateless), stats)
        if rc != 0:
            raise RuntimeError("yo_gphoton_build failed: %d" % rc)
        return stats

    def gphoton_read(self):
        """[n, 15]: position, N, irradiance, merged direction, merged colour per box."""
        L = lib(self.fast)
        n = L.yo_gphoton_read(self.h, None, 0)
        out = np.zeros(max(n, 1) * 15, np.float32)
        L.yo_gphoton_read(self.h, _ptr(out), n)
        return out[: n * 15].reshape(n, 15)

    def gphoton_read_photons(self):
        """[n, 5] uint32: the raw photon map in store order (position as float
        bits, direction code theta | phi << 8, RGBE r | g << 8 | b << 16 | e << 24)."""
        L = lib(self.fast)
        n = L.yo_gphoton_read_photons(self.h, None, 0)
        out = np.zeros(max(n, 1) * 5, np.uint32)
        L.yo_gphoton_read_photons(self.h, out.ctypes.data, n)
        return out[: n * 5].reshape(n, 5)

    def gphoton_load_photons(self, recs):
        """Replace the raw photon map by [n, 5] uint32 records (after gphoton_load)."""
        r = np.ascontiguousarray(recs, np.uint32)
        if lib(self.fast).yo_gphoton_load_photons(self.h, r.ctypes.data, len(r)) != 0:
            raise RuntimeError("yo_gphoton_load_photons: no map loaded")

    def psamp_plan(self, P, N, Ru, Rv, samples, grid=36):
        """photonSampler_t::samplingFrom at one point (stateless gather):
        (hits[sectors], weight[sectors], multi, photons found)."""
        v = [np.ascontiguousarray(x, np.float32) for x in (P, N, Ru, Rv)]
        hits = np.zeros(64, np.int32)
        w = np.zeros(64, np.float32)
        m = np.zeros(1, np.float32)
        nf = np.zeros(1, np.int32)
        ns = lib(self.fast).yo_psamp_plan(self.h, *[x.ctypes.data for x in v], int(samples), int(grid),
                                          hits.ctypes.data, w.ctypes.data, m.ctypes.data, nf.ctypes.data)
        if ns < 0:
            raise RuntimeError("yo_psamp_plan failed: %d" % ns)
        return hits[:ns], w[:ns], float(m[0]), int(nf[0])

    def cache_read(self):
        """Irradiance cache of the last render with a cache = on pathlight:
        ([n, 18] samples in insertion order -- P, N, pP, colour, adist, M,
        precision, generic-tree rank, fake-pass pixel, devaluated -- and the
        number of samples its render pass had to take)."""
        L = lib(self.fast)
        nn = np.zeros(1, np.uint64)
        n = L.yo_cache_read(self.h, None, 0, nn)
        out = np.zeros(max(n, 1) * 18, np.float32)
        L.yo_cache_read(self.h, out.ctypes.data, n, nn)
        return out[: n * 18].reshape(n, 18), int(nn[0])

    def cache_record(self, on=True):
        lib(self.fast).yo_cache_record(self.h, int(on))

    def cache_visits(self):
        """[n, 13] fake-pass visits of the last recorded cache render: pixel,
        area, qualifies, P, face-forwarded N, pP, material."""
        L = lib(self.fast)
        n = L.yo_cache_visits(self.h, None, 0)
        out = np.zeros(max(n, 1) * 13, np.float32)
        L.yo_cache_visits(self.h, out.ctypes.data, n)
        return out[: n * 13].reshape(n, 13)

    def gphoton_load(self, desc, li, boxes, stateless=True, recompute=False):
        """Use `boxes` ([n, 15], as gphoton_read gives them) as light li's
        irradiance boxes; recompute: pre-gather their irradiance here."""
        b = np.ascontiguousarray(boxes, np.float32).reshape(-1)
        rc = lib(self.fast).yo_gphoton_load(self.h, b, len(b) // 15, C.addressof(desc.lights[li]), li,
                                            int(stateless), int(recompute))
        if rc != 0:
            raise RuntimeError("yo_gphoton_load failed")

    def photon_gather(self, desc, li, p, n, ng, eye, mat):
        L = lib(self.fast)
        cnt = len(mat)
        rgb = np.zeros(cnt * 3, np.float32)
        light = desc.lights[li]
        L.yo_photon_gather(self.h, _addr(desc.mats), _addr(desc.mods), _addr(desc.texs),
                           C.addressof(light), np.ascontiguousarray(p, np.float32).reshape(-1),
                           np.ascontiguousarray(n, np.float32).reshape(-1),
                           np.ascontiguousarray(ng, np.float32).reshape(-1),
                           np.ascontiguousarray(eye, np.float32).reshape(-1),
                           np.ascontiguousarray(mat, np.int32), cnt, rgb)
        return rgb.reshape(cnt, 3)


def _addr(arr):
    if arr is None or len(arr) == 0:
        return None
    return C.addressof(arr)


def block_regions(width, height, block=64, seed=1):
    """blockSpliter_t regions in render order: [n, 8] int32 (x y w h rx ry rw rh)."""
    cap = ((width + block - 1) // block) * ((height + block - 1) // block)
    out = np.zeros((cap, 8), np.int32)
    n = lib().yo_block_regions_into(width, height, block, seed, out, cap)
    return out[:n]


def halton_seq(base, n):
    out = np.empty(n, np.float32)
    lib().yo_halton_seq(base, n, out)
    return out


def glibc_rand(seed, n):
    out = np.empty(n, np.int32)
    lib().yo_glibc_rand(seed, n, out)
    return out


def noise_eval(what, tex, pts):
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1)
    n = len(pts) // 3
    out = np.empty(n * (3 if what == 4 else 1), np.float32)
    rc = lib().yo_noise_eval(what, None if tex is None else C.addressof(tex), pts, n, out)
    if rc != 0:
        raise RuntimeError("yo_noise_eval failed")
    return out


def diffuse_at(desc, mat, pts):
    """Albedo of material `mat` of a SceneDesc at points [n, 3] after its
    modulators: [n, 3] float32."""
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1)
    n = len(pts) // 3
    out = np.empty(n * 3, np.float32)
    lib().yo_diffuse_at(_addr(desc.mats), _addr(desc.mods), _addr(desc.texs), mat, pts, n, out)
    return out.reshape(n, 3)


def canonical_counts(verts, idx, rays, mode, threads=8):
    """Mean node visits, triangle tests and box tests per ray of the
    canonical tight BVH2 of the byte model (yo_bytemodel.c):
    (n_node, n_tri, n_box_tested, tree nodes)."""
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    f = np.ascontiguousarray(idx, np.uint32).reshape(-1)
    r = np.ascontiguousarray(rays, np.float32).reshape(-1)
    n = len(r) // 8
    vis = np.zeros(1, np.uint64)
    t = np.zeros(1, np.uint64)
    b = np.zeros(1, np.uint64)
    nodes = lib().yo_canonical_counts(v, len(v) // 3, f, len(f) // 3, r, n, mode, threads, vis, t, b)
    return float(vis[0]) / n, float(t[0]) / n, float(b[0]) / n, int(nodes)


def heightfield(nx, ny, ex, ey, amp, freq, H, lac, octv):
    out = np.empty(nx * ny * 3, np.float32)
    lib().yo_gen_heightfield(nx, ny, ex, ey, amp, freq, H, lac, octv, out)
    return out.reshape(-1, 3)


def rgbe_roundtrip(c):
    c = np.ascontiguousarray(c, np.float32).reshape(-1)
    out = np.zeros_like(c)
    lib().yo_rgbe_roundtrip(c, out, len(c) // 3)
    return out.reshape(-1, 3)


def dir_roundtrip(d):
    """storedPhoton_t direction quantisation: (decoded [n,3], codes [n,2] theta/phi)."""
    d = np.ascontiguousarray(d, np.float32).reshape(-1)
    out = np.zeros_like(d)
    code = np.zeros(len(d) // 3 * 2, np.int32)
    lib().yo_dir_roundtrip(d, out, code.ctypes.data_as(C.c_void_p), len(d) // 3)
    return out.reshape(-1, 3), code.reshape(-1, 2)


def dirconv_tables():
    t = np.zeros(1022, np.float32)
    lib().yo_dirconv_tables(t)
    return t[:255], t[255:510], t[510:766], t[766:]
