"""Synthetic scenes for the five benchmark workloads.

All randomness comes from a project-owned SplitMix64 stream (base seed
0x59414652) so the GPU path and the CPU oracle see identical inputs.  Objects
are numbered in alphabetical order of their names and lights are listed in
alphabetical order, which is the order the reference uses when it fills
scene_t (src/interface/interface_impl.cc:571-576, src/loader/render.cc:1158-1164).
"""
import math

import numpy as np

from . import descriptors as D

BASE_SEED = 0x59414652
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class SplitMix64:
    """Counter form of SplitMix64: output i = mix(seed + (i+1)*gamma)."""

    def __init__(self, seed):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def u64(self, n):
        with np.errstate(over="ignore"):
            k = np.arange(self.count + 1, self.count + 1 + n, dtype=np.uint64)
            out = _mix(self.state + k * _GAMMA)
        self.count += n
        return out

    def uniform(self, n, lo=0.0, hi=1.0):
        x = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * x


def sub_seed(*keys):
    s = np.uint64(BASE_SEED)
    with np.errstate(over="ignore"):
        for k in keys:
            s = _mix(s + np.uint64(k + 1) * _GAMMA)
    return int(s)


# ---------------------------------------------------------------- meshes --

def icosphere(subdiv):
    """Unit icosphere: 20*4^subdiv triangles, outward counter-clockwise winding."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for (a, b, c) in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts, np.float64), np.array(faces, np.uint32)


def cube():
    """Unit cube [-0.5,0.5]^3: 8 vertices, 12 triangles."""
    v = np.array([[x, y, z] for z in (-0.5, 0.5) for y in (-0.5, 0.5) for x in (-0.5, 0.5)],
                 np.float64)
    f = np.array([[0, 2, 3], [0, 3, 1], [4, 5, 7], [4, 7, 6], [0, 1, 5], [0, 5, 4],
                  [2, 6, 7], [2, 7, 3], [0, 4, 6], [0, 6, 2], [1, 3, 7], [1, 7, 5]], np.uint32)
    return v, f


def grid_indices(nx, ny):
    """Two triangles per grid cell, row-major cells: (v00,v10,v11), (v00,v11,v01)."""
    i = np.arange(nx - 1, dtype=np.uint32)
    j = np.arange(ny - 1, dtype=np.uint32)
    jj, ii = np.meshgrid(j, i, indexing="ij")
    v00 = (jj * nx + ii).reshape(-1)
    v10 = v00 + 1
    v01 = v00 + nx
    v11 = v01 + 1
    t = np.empty((len(v00), 2, 3), np.uint32)
    t[:, 0] = np.stack([v00, v10, v11], 1)
    t[:, 1] = np.stack([v00, v11, v01], 1)
    return t.reshape(-1, 3)


def rotation(u1, u2, u3):
    """Uniform random rotation matrix from three uniforms (Shoemake quaternion)."""
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
                  math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3),
                  math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def rotations(u):
    """Vectorised rotation() over rows of u (n,3)."""
    u1, u2, u3 = u[:, 0], u[:, 1], u[:, 2]
    x = np.sqrt(1 - u1) * np.sin(2 * np.pi * u2)
    y = np.sqrt(1 - u1) * np.cos(2 * np.pi * u2)
    z = np.sqrt(u1) * np.sin(2 * np.pi * u3)
    w = np.sqrt(u1) * np.cos(2 * np.pi * u3)
    R = np.empty((len(u), 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


class MeshBuilder:
    def __init__(self):
        self.verts, self.idx, self.obj, self.mat = [], [], [], []
        self.nv = 0
        self.nobj = 0

    def add(self, v, f, mat):
        """Append one object; mat is a scalar or per-face array."""
        v = np.asarray(v, np.float32).reshape(-1, 3)
        f = np.asarray(f, np.uint32).reshape(-1, 3)
        self.verts.append(v)
        self.idx.append(f + np.uint32(self.nv))
        self.obj.append(np.full(len(f), self.nobj, np.int32))
        m = np.asarray(mat, np.int32)
        self.mat.append(np.broadcast_to(m, (len(f),)).astype(np.int32))
        self.nv += len(v)
        self.nobj += 1

    def arrays(self):
        return (np.concatenate(self.verts), np.concatenate(self.idx), np.concatenate(self.obj),
                np.concatenate(self.mat))


# --------------------------------------------------------------- configs --

def albedos(rng, n):
    return rng.uniform(3 * n, 0.1, 0.9).reshape(n, 3)


def direct_room(resx=640, resy=480, samples=16, subdiv=5):
    """Plain direct lighting: box room (10 tris), icosphere r=1 at (0,0,1),
    two unit cubes on the floor, 1x1 area light (power 10) facing -z."""
    rng = SplitMix64(sub_seed(1))
    alb = albedos(rng, 4)
    mb = MeshBuilder()
    # objects in name order: box0, box1, room, sphere
    cube_v, cube_f = cube()
    for k in range(2):
        u = rng.uniform(4)
        Rm = rotation(0.0, 0.0, 0.0)
        ang = 2 * math.pi * u[2]
        Rm = np.array([[math.cos(ang), -math.sin(ang), 0], [math.sin(ang), math.cos(ang), 0], [0, 0, 1]])
        cx = -2.2 + 4.4 * u[0]
        cy = -1.0 + 3.0 * u[1]
        if abs(cx) < 1.7 and abs(cy) < 1.7:
            cx = math.copysign(1.9, cx if cx != 0 else 1.0)
        mb.add(cube_v @ Rm.T + np.array([cx, cy, 0.5]), cube_f, k)
    room_v = np.array([[-3, -3, 0], [3, -3, 0], [3, 3, 0], [-3, 3, 0],
                       [-3, -3, 6], [3, -3, 6], [3, 3, 6], [-3, 3, 6]], np.float64)
    room_f = np.array([[0, 1, 2], [0, 2, 3],       # floor
                       [4, 6, 5], [4, 7, 6],       # ceiling
                       [3, 2, 6], [3, 6, 7],       # back (y=3)
                       [0, 3, 7], [0, 7, 4],       # left (x=-3)
                       [1, 5, 6], [1, 6, 2]], np.uint32)  # right (x=3)
    mb.add(room_v, room_f, 2)
    sv, sf = icosphere(subdiv)
    mb.add(sv + np.array([0.0, 0.0, 1.0]), sf, 3)
    verts, idx, obj, mat = mb.arrays()
    mats = [D.material(a) for a in alb]
    h = 5.9
    light = D.arealight((-0.5, -0.5, h), (-0.5, 0.5, h), (0.5, 0.5, h), (0.5, -0.5, h),
                        samples=samples, power=10.0)
    cam = D.camera((0, -2.9, 2), (0, 0, 1.2), (0, -2.9, 3), resx, resy, 1.0, 1.0)
    opts = D.render_opts(bias=1e-3, raydepth=3)
    return D.SceneDesc(verts, idx, obj, mat, [7] * mb.nobj, cam, mats, [], [], [light], opts)


def terrain_mesh(nx, ny, heights_fn, ex=2.0, ey=1.0, nmat=64, seed_key=2, amp=None, freq=4.0):
    """Heightfield over [0,ex]x[0,ey]; z = 0.2*E*fBm(freq*p) (E = ex), fBm H=1,
    lacunarity 2, 6 octaves over newPerlin_t.  heights_fn(nx,ny,ex,ey,amp,freq,H,lac,oct)
    returns the (nx*ny,3) vertex array.  Materials: nmat height bands."""
    amp = 0.2 * ex if amp is None else amp
    verts = np.asarray(heights_fn(nx, ny, ex, ey, amp, freq, 1.0, 2.0, 6.0), np.float32).reshape(-1, 3)
    idx = grid_indices(nx, ny)
    zc = verts[idx, 2].astype(np.float64).mean(axis=1)
    lo, hi = float(zc.min()), float(zc.max())
    band = np.clip(((zc - lo) / max(hi - lo, 1e-30) * nmat).astype(np.int64), 0, nmat - 1)
    return verts, idx, band.astype(np.int32)


def terrain_camera(resx, resy, ex=2.0, ey=1.0):
    return D.camera((0.5 * ex, -0.45 * ey, 0.55 * ex), (0.5 * ex, 0.45 * ey, 0.0),
                    (0.5 * ex, -0.45 * ey, 0.55 * ex + 1.0), resx, resy, 1.0, 1.2)


def terrain_gi(heights_fn, nx=1024, ny=512, resx=1280, resy=720, samples=4, depth=3,
               nmat=64, threshold=1e3, seed_key=2):
    """Path-traced GI on the fBm terrain: sun (1,0.5,1.5) power 3, constant sky
    (0.4,0.5,0.7), pathlight samples/depth, use_QMC on."""
    ex, ey = 2.0, 1.0
    verts, idx, band = terrain_mesh(nx, ny, heights_fn, ex, ey, nmat)
    rng = SplitMix64(sub_seed(seed_key))
    alb = albedos(rng, nmat)
    mats = [D.material(a) for a in alb]
    sun_dir = np.array([1.0, 0.5, 1.5])
    sun_dir /= np.linalg.norm(sun_dir)
    # lights in name order: path, sun
    lights = [D.pathlight(samples=samples, power=1.0, depth=depth, use_qmc=True,
                          threshold=threshold),
              D.sunlight(sun_dir, power=3.0)]
    cam = terrain_camera(resx, resy, ex, ey)
    opts = D.render_opts(bias=2e-3, raydepth=3, background=(0.4, 0.5, 0.7))
    return D.SceneDesc(verts, idx, np.zeros(len(idx), np.int32), band, [7], cam, mats, [], [],
                       lights, opts)


def photon_field(n_side=64, subdiv=4, extent=100.0, resx=1920, resy=1080,
                 photons=3998792, search=50, nmat=16):
    """Photon mapping: n_side^2 icospheres (rotated, scaled U[0.5,1.5]) on a
    jittered grid over [0,E]^2 plus a 2-triangle ground plane; a 45-degree
    spot photon light, diffuse mode, depth 1, mindepth 0, K = search,
    fixedradius 0.05 E."""
    E = extent
    rng = SplitMix64(sub_seed(3))
    alb = albedos(rng, nmat + 1)
    sv, sf = icosphere(subdiv)
    n = n_side * n_side
    cell = E / n_side
    jit = rng.uniform(2 * n).reshape(n, 2)
    gx, gy = np.meshgrid(np.arange(n_side), np.arange(n_side), indexing="xy")
    cx = (gx.reshape(-1) + 0.15 + 0.7 * jit[:, 0]) * cell
    cy = (gy.reshape(-1) + 0.15 + 0.7 * jit[:, 1]) * cell
    scale = rng.uniform(n, 0.5, 1.5)
    R = rotations(rng.uniform(3 * n).reshape(n, 3))
    mids = (rng.u64(n) % np.uint64(nmat)).astype(np.int32) + 1
    # object order: ground, sph0000 ... (alphabetical)
    gv = np.array([[-0.5 * E, -0.5 * E, 0], [1.5 * E, -0.5 * E, 0], [1.5 * E, 1.5 * E, 0],
                   [-0.5 * E, 1.5 * E, 0]], np.float32)
    gf = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    nsv, nsf = len(sv), len(sf)
    verts = np.empty((4 + n * nsv, 3), np.float32)
    verts[:4] = gv
    # the sphere radius is 1*scale; rest it on the ground
    inst = np.einsum("nij,vj->nvi", R, sv) * scale[:, None, None]
    inst += np.stack([cx, cy, scale], 1)[:, None, :]
    verts[4:] = inst.reshape(-1, 3).astype(np.float32)
    idx = np.empty((2 + n * nsf, 3), np.uint32)
    idx[:2] = gf
    idx[2:] = (sf[None, :, :] + (4 + np.arange(n, dtype=np.uint32) * nsv)[:, None, None]).reshape(-1, 3)
    tri_obj = np.empty(len(idx), np.int32)
    tri_obj[:2] = 0
    tri_obj[2:] = np.repeat(np.arange(1, n + 1, dtype=np.int32), nsf)
    tri_mat = np.empty(len(idx), np.int32)
    tri_mat[:2] = 0
    tri_mat[2:] = np.repeat(mids, nsf)
    mats = [D.material(a) for a in alb]
    light = D.photonlight(frm=(0.5 * E, 0.5 * E, 0.8 * E), to=(0.5 * E, 0.5 * E, 0.0),
                          power=float(E * E) * 0.5, angle=45.0, photons=photons, search=search,
                          depth=1, mindepth=0, fixedradius=0.05 * E, use_qmc=True)
    cam = D.camera((0.5 * E, -0.35 * E, 0.55 * E), (0.5 * E, 0.5 * E, 0.0),
                   (0.5 * E, -0.35 * E, 0.55 * E + 1.0), resx, resy, 1.0, 1.2)
    opts = D.render_opts(bias=1e-3 * E, raydepth=3)
    return D.SceneDesc(verts, idx, tri_obj, tri_mat, [7] * (n + 1), cam, mats, [], [], [light], opts)


def terrain_diff(heights_fn, nx=1024, ny=512, res=512, samples=16, depth=3, nmat=256,
                 amp=0.5, size=3.0, sun_power=3.0, albedo_lo=0.1, albedo_hi=0.9, seed_key=5):
    """Differentiable GI: the terrain at res x res with nmat height-band
    materials whose albedo is modulated (mode mul) by a Musgrave fBm texture;
    amplitude = modulator colour, frequency = modulator size.  Albedos are
    drawn U[albedo_lo, albedo_hi] per channel, by default the GI terrain's
    U[0.1, 0.9].  With albedo_lo = 0.35 no path of depth 3 falls below the
    pathlight's energy cut-off (pathlight.cc:315), so the image is smooth in
    every parameter (what central differences need); below that the cut-off
    ends some paths, a step that the adjoint, like visibility, holds fixed."""
    ex, ey = 2.0, 1.0
    verts, idx, band = terrain_mesh(nx, ny, heights_fn, ex, ey, nmat)
    rng = SplitMix64(sub_seed(seed_key))
    alb = rng.uniform(3 * nmat, albedo_lo, albedo_hi).reshape(nmat, 3)
    mats = [D.material(a, 0, 1) for a in alb]
    texs = [D.musgrave(color1=(0.6, 0.6, 0.6), color2=(1.0, 1.0, 1.0), H=1.0, lacunarity=2.0,
                       octaves=4.0, size=1.0, intensity=0.5)]
    mods = [D.modulator(0, D.YG_MOD_MUL, color=amp, size=size)]
    sun_dir = np.array([1.0, 0.5, 1.5])
    sun_dir /= np.linalg.norm(sun_dir)
    lights = [D.pathlight(samples=samples, power=1.0, depth=depth, use_qmc=True, threshold=1e3),
              D.sunlight(sun_dir, power=sun_power)]
    cam = terrain_camera(res, res, ex, ey)
    opts = D.render_opts(bias=2e-3, raydepth=3, background=(0.4, 0.5, 0.7))
    return D.SceneDesc(verts, idx, np.zeros(len(idx), np.int32), band, [7], cam, mats, mods, texs,
                       lights, opts)


def with_params(desc, albedo=None, amp=None, size=None, sun_power=None):
    """Copy of a terrain_diff SceneDesc with some parameters replaced."""
    mats = list(desc.mats)
    if albedo is not None:
        mats = [D.material(tuple(float(x) for x in a), m.mod_first, m.mod_count)
                for a, m in zip(albedo, desc.mats)]
    mods = [D.yg_modulator.from_buffer_copy(m) for m in desc.mods]
    if amp is not None:
        mods[0].color = float(amp)
    if size is not None:
        mods[0].size = D.F3(float(size), float(size), float(size))
    lights = [D.yg_light.from_buffer_copy(L) for L in desc.lights]
    if sun_power is not None:
        for L in lights:
            if L.type == D.YG_LIGHT_SUN:
                L.power = float(sun_power)
    return D.SceneDesc(desc.verts, desc.idx, desc.tri_obj, desc.tri_mat, desc.obj_flags, desc.cam,
                       mats, mods, list(desc.texs), lights, desc.opts)


BATCH_RES = [(256, 256), (640, 480), (1280, 720), (1920, 1080)]


def batch_specs(n=64, lo=1e3, hi=2e6):
    """Per-scene (nx, ny, resx, resy) for the batched workload: triangle count
    log-uniform in [lo, hi] with nx = 2 ny, resolution uniform from BATCH_RES."""
    rng = SplitMix64(sub_seed(4))
    u = rng.uniform(2 * n).reshape(n, 2)
    specs = []
    for k in range(n):
        tris = math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * u[k, 0])
        # 2*(nx-1)*(ny-1) = tris with nx = 2 ny
        ny = max(2, int(round(0.5 + math.sqrt(tris / 4.0))))
        nx = 2 * ny
        rx, ry = BATCH_RES[min(3, int(u[k, 1] * 4))]
        specs.append((nx, ny, rx, ry))
    return specs


def gphoton_room(resx=80, resy=60, subdiv=3, photons=200000, depth=2, search=100, radius=1.0, samples=8,
                 path_depth=1, sun=None):
    """The direct-lighting room with its area light made a dummy emitter of
    global photons (arealight.cc:98), a global photon light
    (globalphotonlight.cc:280-294) and a pathlight whose path vertices read
    the irradiance map (pathlight.cc:140-185); optionally a sun."""
    r = direct_room(resx, resy, samples=4, subdiv=subdiv)
    A = r.lights[0]
    lights = [D.arealight(A.a, A.b, A.c, A.d, samples=1, color=tuple(A.color), power=A.power, dummy=True),
              D.globalphotonlight(radius=radius, depth=depth, photons=photons, search=search),
              D.pathlight(samples=samples, depth=path_depth, use_qmc=True, threshold=1e3)]
    if sun is not None:
        lights.append(D.sunlight(sun, power=0.8))
    return D.SceneDesc(r.verts, r.idx, r.tri_obj, r.tri_mat, r.obj_flags, r.cam, list(r.mats), [], [], lights,
                       r.opts)
