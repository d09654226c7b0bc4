"""Scenes whose renders, photon maps or albedos have closed forms, shared by
the oracle pins (tests/test_oracle_closed_forms.py) and the GPU tests
(tests/test_gpu_closed_forms.py).

Each builder returns a SceneDesc plus what its expected value needs.  The
expected values are written from the reference's definitions (file:line in
each docstring) and from geometry, never from the oracle's code.
"""
import math

import numpy as np

from yafray_amd import descriptors as D


def radinv(i, base):
    """Exact radical inverse of the integer i in `base`."""
    r, f = 0.0, 1.0 / base
    while i:
        r += (i % base) * f
        i //= base
        f /= base
    return r


def area_pixel_index(resx, resy, blk=64):
    """k[i, j]: the index of pixel (row i, column j) in its render area's
    row-major pixel order, the area grown by the 1-pixel overlap border of
    blockSpliter_t (renderblock.cc:71-101).  The pathlight's per-area Halton
    sampler (pathlight.cc:725-738, restarted per area: scene.cc:726) gives the
    s-th sample of that pixel the call index k * samples + s + 1 on every
    level a path of every earlier pixel reached."""
    k = np.zeros((resy, resx), np.int64)
    for by in range((resy + blk - 1) // blk):
        for bx in range((resx + blk - 1) // blk):
            rx, ry = bx * blk, by * blk
            rw, rh = min(blk, resx - rx), min(blk, resy - ry)
            x, y, w, h = rx, ry, rw, rh
            if x > 0:
                x -= 1; w += 1
            if y > 0:
                y -= 1; h += 1
            if x + w < resx - 1:
                w += 1
            if y + h < resy - 1:
                h += 1
            for i in range(ry, ry + rh):
                for j in range(rx, rx + rw):
                    k[i, j] = (i - y) * w + (j - x)
    return k


def _quad(z, E, flip=False):
    v = np.array([[-E, -E, z], [E, -E, z], [E, E, z], [-E, E, z]], np.float32)
    f = np.array([[0, 2, 1], [0, 3, 2]] if flip else [[0, 1, 2], [0, 2, 3]], np.uint32)
    return v, f


# ------------------------------------------------ two-bounce pathlight --

RHO_F = (0.7, 0.6, 0.5)   # floor albedo
RHO_C = (0.8, 0.7, 0.6)   # ceiling albedo
SKY = (0.4, 0.5, 0.7)


def two_bounce(samples=4, depth=2, resx=80, resy=72, power=1.5, sun_power=2.0):
    """Camera between a floor (z = 0, casts shadows) and a ceiling (z = 1,
    castShadows off) looking up at the ceiling, sun from above, constant sky,
    and a basement (z = -1) below the floor that casts shadows.

    * Camera rays see the ceiling (only shadow queries skip it: scene.cc:698,
      107-185).  Its own sun term is 0: its normal faces the camera, away
      from the sun (sunlight.cc:35-57, basicshaders.cc:46-70).
    * Every level-0 pathlight ray leaves the ceiling downwards and hits the
      floor.  There the sun is unshadowed, because the ceiling casts no
      shadow: getLight = rho_f (n.s) E (pathlight.cc:186-191), weighted by
      raycolour 1 and startray.N = sqrt(z1) (pathlight.cc:328-331,
      pathtools.cc:32-43).
    * raycolour becomes rho_f; the level-1 ray leaves the floor upwards around
      its face-forwarded normal HN, passes the ceiling (firstHit with the
      shadow flag skips it, pathlight.cc:317) and escapes: + rho_f sqrt(z1) B.
      A flipped HN would send it into the basement, where the floor hides
      the sun and the loop ends at depth 2: the sky term would be missing.
    So pixel = rho_c P (1/S) sum_i sqrt(z1_i) rho_f ((n.s) E + B), z1_i the
    base-2 radical inverse of sample i's per-area call index (Halton
    dimension 0, mcqmc.h:11-54), and each sample traces exactly 2 extension
    rays."""
    F = 2000.0
    vf, ff = _quad(0.0, F)
    vc, fc = _quad(1.0, F)
    vb, fb = _quad(-1.0, F)
    v = np.vstack([vb, vc, vf])
    f = np.vstack([fb, fc + 4, ff + 8])
    # objects in name order: basement (0), ceiling (1), floor (2)
    tri_obj = [0, 0, 1, 1, 2, 2]
    tri_mat = [2, 2, 1, 1, 0, 0]
    flags = [7, 6, 7]
    sun = np.array([0.3, 0.2, 1.0])
    sun /= np.linalg.norm(sun)
    lights = [D.pathlight(samples=samples, power=power, depth=depth, use_qmc=True, threshold=1e3),
              D.sunlight(sun, power=sun_power)]
    cam = D.camera((0.1, -0.2, 0.3), (0.1, -0.2, 1.3), (0.1, 0.8, 0.3), resx, resy, 1.0, 1.0)
    opts = D.render_opts(bias=1e-3, background=SKY)
    mats = [D.material(RHO_F), D.material(RHO_C), D.material((0.5, 0.5, 0.5))]
    d = D.SceneDesc(v, f, tri_obj, tri_mat, flags, cam, mats, [], [], lights, opts)
    return d, sun


def two_bounce_expected(samples, resx, resy, sun, power=1.5, sun_power=2.0):
    k = area_pixel_index(resx, resy)
    rf, rc, B = np.array(RHO_F), np.array(RHO_C), np.array(SKY)
    light = sun[2] * sun_power + B
    exp = np.zeros((resy, resx, 3))
    for i in range(resy):
        for j in range(resx):
            z = [radinv(int(k[i, j]) * samples + s + 1, 2) for s in range(samples)]
            exp[i, j] = rc * power * (np.sum(np.sqrt(z)) / samples) * rf * light
    return exp


# -------------------------------------------------------- photon bounce --

PH_RHO = (0.7, 0.6, 0.5)
PH_COL = (1.0, 0.8, 0.6)


def photon_bounce(photons=60 * 60, angle=30.0, power=3.0, cluster=1e-6):
    """A spot light at z = 1 aiming at a floor (z = 0), a ceiling at z = 2,
    photonlight depth 1, mindepth 1.  Every photon hits the floor first
    (depth 1: not stored, since depth > mindepth fails) and bounces
    (useForRadiosity, depth <= maxdepth: photonlight.cc:222-241) along
    randomVectorCone(Ng, NU, NV, 0.05, r1, r2) (vector3d.cc:162-168), r1, r2
    the Halton dimensions 2 and 3 (bases 5 and 7) at the photon's bounce
    index.  The bounce reaches the ceiling (depth 2 > mindepth: stored, and
    depth > maxdepth: no further bounce).  The stored colour is
    fromRadiosity at the floor = colour x power x 2 pi (1 - cos angle) x
    rho_floor (photonlight.cc:74, basicshaders.cc:30-44); the mark's
    direction points back to the floor hit, so its z is -cos(theta) with
    cos(theta) = 1 - 0.95 r2, and its azimuth in the floor's createCS frame
    (vector3d.h:240-258: u = (1, 0, 0), v = (0, 1, 0) for n = +z) is
    2 pi r1 + pi."""
    vb, fb = _quad(0.0, 400.0)
    vc, fc = _quad(2.0, 400.0)
    v = np.vstack([vc, vb])
    f = np.vstack([fc, fb + 4])
    L = D.photonlight(frm=(0, 0, 1), to=(0, 0, 0), color=PH_COL, power=power, angle=angle,
                      photons=photons, search=20, depth=1, mindepth=1, fixedradius=1.0, cluster=cluster,
                      use_qmc=True)
    cam = D.camera((0, 0, 1.5), (0, 0.01, 0), (0, 1, 2.5), 8, 8, 1.0, 1.0)
    mats = [D.material((0.5, 0.5, 0.5)), D.material(PH_RHO)]
    return D.SceneDesc(v, f, [0, 0, 1, 1], [0, 0, 1, 1], [7, 7], cam, mats, [], [], [L],
                       D.render_opts(bias=1e-3))


def photon_power(angle, power):
    """Per-photon colour scale of photonLight_t's constructor (photonlight.cc:74)."""
    c = float(np.float32(math.cos(math.radians(angle))))
    return power * 2.0 * math.pi * (1.0 - c)


# -------------------------------------------------------- cluster merge --

def photon_spot_plane(photons=40 * 40, cluster=1e-6, at=(5.0, 5.0), z=0.5, angle=20.0, power=2.0):
    """A spot light one unit over a plane at height z, photons stored at their
    first hit (depth 0, mindepth 0): with a tiny cluster size every photon is
    its own mark, so these are the photons preGathering merges."""
    E = 50.0
    v = np.array([[-E, -E, z], [E, -E, z], [E, E, z], [-E, E, z]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    L = D.photonlight(frm=(at[0], at[1], z + 1.0), to=(at[0], at[1], z), color=PH_COL, power=power,
                      angle=angle, photons=photons, search=20, depth=0, mindepth=0, fixedradius=1.0,
                      cluster=cluster, use_qmc=True)
    cam = D.camera((0, 0, 3), (0, 0.01, 0), (0, 1, 4), 8, 8, 1.0, 1.0)
    return D.SceneDesc(v, f, [0, 0], [0, 0], [7], cam, [D.material(PH_RHO)], [], [], [L],
                       D.render_opts(bias=1e-3))


def merge_expected(marks, cell):
    """preGathering (photonlight.cc:397-411) of unmerged marks [n, 9] (dir,
    pos, colour) inserted into cells of size `cell` (hash3d_t::getBox:
    floor(p / cell), hash3d.h:218-228): per cell, in (x, y, z) cell order
    (std::map order), the normalised sum of directions, the mean position
    and the summed colour."""
    keys = np.floor(marks[:, 3:6].astype(np.float64) / cell).astype(np.int64)
    order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
    out = []
    cur, grp = None, []
    for i in list(order) + [None]:
        key = None if i is None else tuple(keys[i])
        if key != cur and grp:
            g = marks[grp].astype(np.float64)
            dsum = g[:, 0:3].sum(0)
            out.append(np.concatenate([dsum / np.linalg.norm(dsum), g[:, 3:6].mean(0), g[:, 6:9].sum(0)]))
            grp = []
        if i is not None:
            cur = key
            grp.append(i)
    return np.array(out)


# ------------------------------------------------------------ modulators --

def marble_points(fvals):
    """Points where marble without turbulence (basictex.cc:133-153) takes the
    value f: 1/2 + 1/2 sin(5 s) = f with s = x + y + z."""
    s = np.arcsin(2 * np.asarray(fvals, np.float64) - 1) / 5
    return np.stack([s / 3, s / 3, s / 3], 1)


def modulated_scene(mode_factors, base=(0.4, 0.5, 0.6), c1=(0.1, 0.2, 0.3), c2=(0.9, 0.6, 0.5)):
    """One material of colour `base` with modulators [(mode, factor), ...]
    applied in order, all over one marble texture without turbulence."""
    tex = [D.marble(color1=c1, color2=c2, turbulence=0.0)]
    mods = [D.modulator(0, mode, color=c) for mode, c in mode_factors]
    mats = [D.material(base, 0, len(mods))]
    return mats, mods, tex


def modulate_expected(C, tc, mode, c):
    """modulator_t::modulate's colour part (texture.cc:499-534) with
    color.cc:103-109's mix(a, b, p) = a p + (1 - p) b (b for p < 0, a for p > 1)."""
    C = np.asarray(C, np.float64)
    tc = np.asarray(tc, np.float64)
    if not c > 0:
        return C
    def mix(a, b, p):
        if p > 1:
            return a
        return a * p + (1 - p) * b
    if mode == D.YG_MOD_MIX:
        return mix(tc, C, c)
    if mode == D.YG_MOD_MUL:
        return mix(tc, np.ones(3), c) * C
    if mode == D.YG_MOD_ADD:
        return tc * c + C
    return tc * -c + C


def textured_plane(mode_factors, resx=64, resy=48):
    """A plane z = 0 seen from above, lit by a sun straight above it (power
    1): every pixel is the modulated albedo at its hit point (N.s = 1)."""
    E = 50.0
    v = np.array([[-E, -E, 0], [E, -E, 0], [E, E, 0], [-E, E, 0]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    mats, mods, tex = modulated_scene(mode_factors)
    cam = D.camera((0.2, 0.1, 2.0), (0.2, 0.11, 0.0), (0.2, 1.1, 2.0), resx, resy, 1.0, 1.0)
    return D.SceneDesc(v, f, [0, 0], [0, 0], [7], cam, mats, mods, tex, [D.sunlight((0, 0, 1), power=1.0)],
                       D.render_opts(bias=1e-3))


# ----------------------------------------------------------------- camera --

def camera_grid(resx=96, resy=60, aspect=1.3, focal=1.7, G=12, D_=4.0, X=3.0, Z=2.0):
    """A wall y = D_ of G x G cells, each its own material whose albedo
    encodes the cell, seen by a camera at the origin looking along +y with
    up +z, lit by a sun shining along +y (power 1): each pixel shows the
    albedo of the cell its camera ray meets."""
    xs = np.linspace(-X, X, G + 1)
    zs = np.linspace(-Z, Z, G + 1)
    verts, faces, mats_id = [], [], []
    for gz in range(G):
        for gx in range(G):
            b = len(verts)
            verts += [[xs[gx], D_, zs[gz]], [xs[gx + 1], D_, zs[gz]], [xs[gx + 1], D_, zs[gz + 1]],
                      [xs[gx], D_, zs[gz + 1]]]
            faces += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
            mats_id += [gz * G + gx] * 2
    mats = [D.material(((gx + 1) / (G + 1), (gz + 1) / (G + 1), 0.25)) for gz in range(G) for gx in range(G)]
    cam = D.camera((0, 0, 0), (0, 1, 0), (0, 0, 1), resx, resy, aspect, focal)
    d = D.SceneDesc(np.array(verts, np.float32), np.array(faces, np.uint32), [0] * len(faces), mats_id, [7],
                    cam, mats, [], [], [D.sunlight((0, -1, 0), power=1.0)], D.render_opts(bias=1e-3))
    return d


def camera_grid_expected(resx=96, resy=60, aspect=1.3, focal=1.7, G=12, D_=4.0, X=3.0, Z=2.0):
    """Pinhole projection: the image plane lies at distance `focal` in front
    of the eye, one unit wide and aspect x resy / resx high, columns running
    to the camera's right and rows downwards (row 0 at the top); pixel (i, j)
    is sampled at its centre (j + 1/2, i + 1/2) (scene.cc:736-756).  So the
    ray of pixel (i, j) meets the wall y = D_ at
        x = D_ ((j + 1/2) / resx - 1/2) / focal,
        z = -D_ aspect (resy / resx) ((i + 1/2) / resy - 1/2) / focal.
    Returns the expected albedo per pixel and a mask of pixels whose point
    lies at least 1e-3 cell away from every cell edge."""
    j = np.arange(resx) + 0.5
    i = np.arange(resy) + 0.5
    x = D_ * (j / resx - 0.5) / focal
    z = -D_ * aspect * (resy / resx) * (i / resy - 0.5) / focal
    fx = (x + X) / (2 * X) * G
    fz = (z + Z) / (2 * Z) * G
    gx = np.floor(fx).astype(int)
    gz = np.floor(fz).astype(int)
    okx = (gx >= 0) & (gx < G) & (np.abs(fx - np.round(fx)) > 1e-3)
    okz = (gz >= 0) & (gz < G) & (np.abs(fz - np.round(fz)) > 1e-3)
    exp = np.zeros((resy, resx, 3))
    exp[:, :, 0] = ((gx + 1) / (G + 1))[None, :]
    exp[:, :, 1] = ((gz + 1) / (G + 1))[:, None]
    exp[:, :, 2] = 0.25
    mask = okz[:, None] & okx[None, :]
    return exp, mask


# ------------------------------------- level-1 hits: floor, wall, ceiling --

RHO_W = (0.55, 0.75, 0.65)   # wall albedo
WALL_X = 1.5


def create_cs(n):
    """createCS (vector3d.h:240-258)."""
    n = np.asarray(n, np.float64)
    if n[0] == 0 and n[1] == 0:
        u = np.array([-1.0, 0, 0]) if n[2] < 0 else np.array([1.0, 0, 0])
        return u, np.array([0.0, 1, 0])
    d = 1.0 / math.sqrt(n[1] * n[1] + n[0] * n[0])
    u = np.array([n[1] * d, -n[0] * d, 0.0])
    return u, np.cross(n, u)


def wall_bounce(samples=4, resx=48, resy=40, power=1.5, sun_power=2.0):
    """two_bounce's floor and non-casting ceiling, the floor ending at a
    vertical wall x = WALL_X (casts shadows, 1000 high) and a sun from the -x
    side: s = normalize(-a, 0, b).  The wall faces the sun, so level-0 or
    level-1 rays that reach it find light rho_w a E there; the floor gets
    rho_f b E; rays that pass the ceiling and miss the wall see the sky.
    Depth 2: the loop ends after the level-1 vertex."""
    F = 2000.0
    W = WALL_X
    vf = np.array([[-F, -F, 0], [W, -F, 0], [W, F, 0], [-F, F, 0]], np.float32)
    vw = np.array([[W, -F, 0], [W, F, 0], [W, F, 1000], [W, -F, 1000]], np.float32)
    vc, fc = _quad(1.0, F)
    q = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    v = np.vstack([vc, vf, vw])
    f = np.vstack([fc, q + 4, q + 8])
    tri_obj = [0, 0, 1, 1, 2, 2]        # ceiling, floor, wall
    tri_mat = [1, 1, 0, 0, 2, 2]
    flags = [6, 7, 7]
    sun = np.array([-0.6, 0.0, 0.8])
    lights = [D.pathlight(samples=samples, power=power, depth=2, use_qmc=True, threshold=1e3),
              D.sunlight(sun, power=sun_power)]
    cam = D.camera((0.3, 0.1, 0.4), (0.3, 0.1, 1.4), (0.3, 1.1, 0.4), resx, resy, 1.0, 1.0)
    mats = [D.material(RHO_F), D.material(RHO_C), D.material(RHO_W)]
    d = D.SceneDesc(v, f, tri_obj, tri_mat, flags, cam, mats, [], [], lights, D.render_opts(bias=1e-3, background=SKY))
    return d, sun


def _tri_normal(a, b, c):
    n = np.cross(np.asarray(b, np.float64) - a, np.asarray(c, np.float64) - a)
    return n / np.linalg.norm(n)


def wall_bounce_expected(d, sun, samples=4, power=1.5, sun_power=2.0, halton=None):
    """The pathlight estimator of pathlight.cc:305-347 evaluated on the
    analytic geometry of wall_bounce, sample by sample: camera ray (pinhole,
    as camera_grid_expected) to the ceiling; level-0 direction around the
    ceiling's face-forwarded normal from Halton dimensions 0, 1 at index
    k S + s + 1 (nextDirection: (Ru cos 2 pi z2 + Rv sin 2 pi z2) sqrt(1 - z1)
    + N sqrt(z1), pathtools.cc:32-43, Ru, Rv = createCS of the raw normal);
    level-1 direction from dimensions 2, 3 at the same index (every path
    reaches level 1).  halton(base, n) supplies the sample values (the
    reference's incremental Halton; base 2 is the exact radical inverse).
    Returns (expected image, mask of pixels with no sample near an edge)."""
    cam = d.cam
    resx, resy = cam.resx, cam.resy
    W = WALL_X
    rf, rc, rw, B = (np.array(x, np.float64) for x in (RHO_F, RHO_C, RHO_W, SKY))
    a, b = -sun[0], sun[2]
    E = sun_power
    # raw face normals as the triangles are wound
    nc = _tri_normal(*d.verts[d.idx[0]])
    nf = _tri_normal(*d.verts[d.idx[2]])
    nw = _tri_normal(*d.verts[d.idx[4]])
    pos = np.array(cam.pos[:], np.float64)
    fwd = np.array(cam.look[:], np.float64) - pos
    up = np.array(cam.up[:], np.float64) - pos
    right = np.cross(fwd, up); right /= np.linalg.norm(right)
    upv = np.cross(right, fwd); upv /= np.linalg.norm(upv)
    fwd /= np.linalg.norm(fwd)
    asp = cam.aspect * resy / resx
    k = area_pixel_index(resx, resy)
    nmax = int(k.max() + 1) * samples + 1
    H = {bs: halton(bs, nmax) for bs in (2, 3, 5, 7)}

    def ndir(n_raw, n_ff, z1, z2):
        ru, rv = create_cs(n_raw)
        z1 = min(z1, 1.0)
        t = 2 * math.pi * z2
        return (ru * math.cos(t) + rv * math.sin(t)) * math.sqrt(1 - z1) + n_ff * math.sqrt(z1)

    def trace(o, dvec):
        """(kind, point, margin): kind 'floor', 'wall' or 'sky' for a
        shadow-flag firstHit (the ceiling is skipped); margin = distance of
        the decision from an edge case."""
        tf = -o[2] / dvec[2] if dvec[2] < 0 else math.inf
        tw = (W - o[0]) / dvec[0] if dvec[0] > 0 else math.inf
        if tw < tf:
            p = o + dvec * tw
            if p[2] < 1000:
                return "wall", p, min(p[2], 1000 - p[2])
            return "sky", None, abs(1000 - p[2])
        if tf < math.inf:
            p = o + dvec * tf
            return "floor", p, abs(W - p[0])
        return "sky", None, 1.0

    exp = np.zeros((resy, resx, 3))
    mask = np.ones((resy, resx), bool)
    for i in range(resy):
        for j in range(resx):
            ray = right * ((j + 0.5) / resx - 0.5) - upv * asp * ((i + 0.5) / resy - 0.5) + fwd * cam.focal
            ray /= np.linalg.norm(ray)
            t = (1.0 - pos[2]) / ray[2]
            P = pos + ray * t
            n0 = nc if np.dot(nc, -ray) >= 0 else -nc   # face-forwarded to the eye
            tot = np.zeros(3)
            for s in range(samples):
                idx = int(k[i, j]) * samples + s + 1
                d0 = ndir(nc, n0, H[2][idx - 1], H[3][idx - 1])
                c0 = float(np.dot(d0, n0))
                kind, p1, m = trace(P, d0)
                if m < 1e-3:
                    mask[i, j] = False
                if kind == "sky":
                    tot += c0 * B
                    continue
                if kind == "floor":
                    tot += c0 * rf * b * E
                    nraw, rho = nf, rf
                else:
                    tot += c0 * rw * a * E
                    nraw, rho = nw, rw
                hn = nraw if np.dot(nraw, -d0) >= 0 else -nraw
                d1 = ndir(nraw, hn, H[5][idx - 1], H[7][idx - 1])
                kind1, p2, m1 = trace(p1 + d1 * 5e-5, d1)
                if m1 < 1e-3:
                    mask[i, j] = False
                if kind1 == "sky":
                    tot += rho * c0 * B
                elif kind1 == "floor":
                    tot += rho * c0 * rf * b * E
                else:
                    tot += rho * c0 * rw * a * E
            exp[i, j] = rc * power * tot / samples
    return exp, mask


# ------------------------------------------------- area light over a floor --

def area_over_floor(samples=16, psamples=4, slab=False, resx=64, resy=48, h=2.0, power=3.0, rho=(0.7, 0.6, 0.5)):
    """A 1 x 1 area light at height h facing down over a floor (z = 0), the
    camera looking down at the floor from below the light; with `slab`, an
    opaque plane at h / 2 larger than the view hides the light from every
    visible floor point.  No other light, no background."""
    F = 50.0
    vf, ff = _quad(0.0, F)
    verts, idx, tri_obj, tri_mat, flags = [vf], [ff], [0, 0], [0, 0], [7]
    if slab:
        vs, fs = _quad(h / 2, F)
        verts.append(vs)
        idx.append(fs + 4)
        tri_obj += [1, 1]
        tri_mat += [1, 1]
        flags.append(7)
    v = np.vstack(verts)
    f = np.vstack(idx)
    a, b, c, d = (-0.5, -0.5, h), (-0.5, 0.5, h), (0.5, 0.5, h), (0.5, -0.5, h)  # (b-a)x(d-a) = -z
    L = D.arealight(a, b, c, d, samples=samples, color=(1.0, 0.9, 0.8), power=power, psamples=psamples)
    L.use_in_indirect = 0
    cam = D.camera((0.3, -0.6, 0.25 * h), (0.3, -0.55, 0.0), (0.3, 0.4, 0.25 * h), resx, resy, 1.0, 1.0)
    mats = [D.material(rho), D.material((0.5, 0.5, 0.5))]
    return D.SceneDesc(v, f, tri_obj, tri_mat, flags, cam, mats, [], [], [L], D.render_opts(bias=1e-3))


def area_grid_points(L):
    """Sample points of a square area light of n*n samples: the cell centres
    a + (j + 1/2)/n (b - a) + (i + 1/2)/n (d - a) (fillQuad's grid for equal
    sides, arealight.cc:31-73)."""
    n = int(round(math.sqrt(L.samples)))
    a, b, d = (np.array(L.a[:], np.float64), np.array(L.b[:], np.float64), np.array(L.d[:], np.float64))
    pts = [a + (j + 0.5) / n * (b - a) + (i + 0.5) / n * (d - a) for i in range(n) for j in range(n)]
    return np.array(pts)


def lambert_from_points(P, pts, L, rho, weights=None):
    """sum over points of generic diffuse fromLight (basicshaders.cc:30-70, no
    1/pi) of pow colour / |L|^2 arriving from each point at floor points P
    (normal +z), times plane_at (the light's height above P along its normal),
    divided by the number of points: the area light's LIGHT verdict."""
    col = np.array(L.color[:], np.float64) * float(L.power)
    out = np.zeros((len(P), 3))
    for q in pts:
        Lv = q[None, :] - P
        d2 = np.sum(Lv * Lv, axis=1)
        ndl = Lv[:, 2] / np.sqrt(d2)
        out += np.maximum(ndl, 0)[:, None] / d2[:, None] * col[None, :]
    plane_at = float(L.a[2]) - P[:, 2]
    return out * np.array(rho)[None, :] * plane_at[:, None] / len(pts)


# ------------------------------------------------- texture coordinates --

TEXCO_C1, TEXCO_C2 = (0.1, 0.2, 0.3), (0.9, 0.6, 0.5)
# per texco kind: (YG_TEXCO_*, projection of the modulator)
TEXCO_KINDS = {
    "global": (D.YG_TEXCO_GLOBAL, (0, 0, 0)),
    "uv": (D.YG_TEXCO_UV, (1, -1, -1)),
    "orco_matrix": (D.YG_TEXCO_ORCO, (1, -1, 3)),
    "orco_points": (D.YG_TEXCO_ORCO, (1, -1, 3)),
    "window": (D.YG_TEXCO_WIN, (0, 0, 0)),
    "normal": (D.YG_TEXCO_NOR, (0, 0, 0)),
    "reflect": (D.YG_TEXCO_REFL, (0, 0, 0)),
}


def texco_wall(kind, G=8, resx=96, resy=60, aspect=1.3, focal=1.7, D_=4.0, X=3.0, Z=2.0):
    """camera_grid's wall with one material: marble without turbulence,
    colours TEXCO_C1 -> TEXCO_C2, blended in fully (MIX, colour 1) through one
    modulator with texture coordinates `kind`, so that each pixel is
    c1 + (c2 - c1) (1/2 + 1/2 sin(5 s)) with s the sum of the modulator's
    texpt.  For "uv" a third of the cells have face uv u = (x + X) / 2X,
    v = (z + Z) / 2Z, a third vertex colours only and a third neither; for
    "orco_points" the vertices carry orco points (0.3 x + 0.1, 7, 0.2 z -
    0.05)."""
    xs = np.linspace(-X, X, G + 1)
    zs = np.linspace(-Z, Z, G + 1)
    verts, faces, uv, flags = [], [], [], []
    for gz in range(G):
        for gx in range(G):
            b = len(verts)
            q = [[xs[gx], D_, zs[gz]], [xs[gx + 1], D_, zs[gz]], [xs[gx + 1], D_, zs[gz + 1]], [xs[gx], D_, zs[gz + 1]]]
            verts += q
            faces += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
            quv = [((p[0] + X) / (2 * X), (p[2] + Z) / (2 * Z)) for p in q]
            kind_c = (gx + gz) % 3
            fl = {0: 0, 1: 1, 2: 2}[kind_c]
            for tri in ((0, 1, 2), (0, 2, 3)):
                uv.append([c for t in tri for c in quv[t]])
                flags.append(fl)
    verts = np.array(verts, np.float32)
    texco, proj = TEXCO_KINDS[kind]
    tex = [D.marble(color1=TEXCO_C1, color2=TEXCO_C2, turbulence=0.0)]
    mods = [D.modulator(0, D.YG_MOD_MIX, color=1.0, proj=proj, texco=texco)]
    mats = [D.material((0.4, 0.5, 0.6), 0, 1)]
    cam = D.camera((0, 0, 0), (0, 1, 0), (0, 0, 1), resx, resy, aspect, focal)
    texd, orco = {}, None
    if kind == "uv":
        texd = {0: {"uv": np.array(uv, np.float32), "flags": np.array(flags, np.uint8)}}
    elif kind == "orco_points":
        orco = np.stack([0.3 * verts[:, 0] + 0.1, np.full(len(verts), 7.0), 0.2 * verts[:, 2] - 0.05], 1)
        texd = {0: {"orco": True}}
    return D.SceneDesc(verts, np.array(faces, np.uint32), [0] * len(faces), [0] * len(faces), [7], cam, mats,
                       mods, tex, [D.sunlight((0, -1, 0), power=1.0)], D.render_opts(bias=1e-3), tex=texd,
                       orco=orco)


def texco_wall_expected(kind, G=8, resx=96, resy=60, aspect=1.3, focal=1.7, D_=4.0, X=3.0, Z=2.0):
    """The expected colour per pixel from camera_grid_expected's hit points
    (x, D_, z) and doMapping's coordinates (texture.cc:125-165):
      global   s = x + D_ + z
      uv       (2u - 1 projected to x): 2u - 1 on cells with uv, -3 on cells
               with nothing (u = -1), -1 on cells with vertex colours only
      orco     the orco matrix maps the bound [-X, X] x [-Z, Z] (widened by
               BOX_MIN = 1e-5) to [-1, 1]: s = x / (X + 1e-5) + z / (Z + 1e-5)
      orco points  s = (0.3 x + 0.1) + (0.2 z - 0.05)
      window   the pixel centre's screen position: s = (2 (j + 1/2) / resx - 1)
               + (1 - 2 (i + 1/2) / resy)
      normal   N = (0, -1, 0): s = -1
      reflect  eye = -P, N faces it: reflect(N, eye/|eye|) = (x, -D_, z) / |P|
    Returns the colours and a mask of pixels 1e-3 cell or more from every
    cell edge."""
    j = np.arange(resx) + 0.5
    i = np.arange(resy) + 0.5
    x = (D_ * (j / resx - 0.5) / focal)[None, :] * np.ones((resy, 1))
    z = (-D_ * aspect * (resy / resx) * (i / resy - 0.5) / focal)[:, None] * np.ones((1, resx))
    fx = (x + X) / (2 * X) * G
    fz = (z + Z) / (2 * Z) * G
    gx = np.floor(fx).astype(int)
    gz = np.floor(fz).astype(int)
    mask = ((gx >= 0) & (gx < G) & (np.abs(fx - np.round(fx)) > 1e-3) & (gz >= 0) & (gz < G) &
            (np.abs(fz - np.round(fz)) > 1e-3))
    if kind == "global":
        s = x + D_ + z
    elif kind == "uv":
        u = (x + X) / (2 * X)
        c = (gx + gz) % 3
        s = np.where(c == 1, 2 * u - 1, np.where(c == 0, -3.0, -1.0))
    elif kind == "orco_matrix":
        s = x / (X + 1e-5) + z / (Z + 1e-5)
    elif kind == "orco_points":
        s = (0.3 * x + 0.1) + (0.2 * z - 0.05)
    elif kind == "window":
        s = (2 * j / resx - 1)[None, :] + (1 - 2 * i / resy)[:, None]
    elif kind == "normal":
        s = np.full((resy, resx), -1.0)
    else:
        r = np.sqrt(x * x + D_ * D_ + z * z)
        s = (x - D_ + z) / r
    f = 0.5 + 0.5 * np.sin(5 * s)
    c1, c2 = np.array(TEXCO_C1), np.array(TEXCO_C2)
    return c1 + (c2 - c1) * f[..., None], mask
