"""Pins of the CPU oracle against closed forms, library routines, published
check values and invariants (the reference ships no golden vectors and its
build needs the generated config.h, so it cannot be linked here)."""
import ctypes
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import scenes as S


# ------------------------------------------------------------------ RNG --

def test_glibc_rand_matches_libc():
    """blockSpliter_t's shuffle uses rand() (renderblock.cc:82): the oracle's
    TYPE_3 restatement must equal this host's glibc rand() stream."""
    libc = ctypes.CDLL("libc.so.6")
    for seed in (1, 7, 123456789):
        libc.srand(seed)
        ref = np.array([libc.rand() for _ in range(3000)], np.int32)
        assert np.array_equal(po.glibc_rand(seed, 3000), ref)


def test_park_miller_check_value():
    """ourRandom() is the Park-Miller minimal standard: from seed 1 the
    10000th state is 1043618065 (Park & Miller, CACM 1988)."""
    L = po.lib()
    s = ctypes.c_int32(1)
    for _ in range(10000):
        L.yo_our_random(ctypes.byref(s))
    assert s.value == 1043618065
    assert L.yo_lcg_jump(1, 10000) == 1043618065


def test_lcg_jump_equals_stepping():
    L = po.lib()
    s = ctypes.c_int32(123212)   # myseed, vector3d.cc:160
    vals = []
    for _ in range(777):
        vals.append(L.yo_our_random(ctypes.byref(s)))
    assert L.yo_lcg_jump(123212, 777) == s.value
    # value = state / float(2^31 - 1) = state / 2^31 in float
    assert vals[-1] == np.float32(np.float32(s.value) / np.float32(2147483647))


# ------------------------------------------------------------------ QMC --

def radical_inverse_exact(b, i):
    from fractions import Fraction
    r, f = Fraction(0), Fraction(1, b)
    while i:
        r += (i % b) * f
        i //= b
        f /= b
    return r


def test_halton_base2_is_exact_van_der_corput():
    seq = po.halton_seq(2, 4096)
    exp = np.array([float(radical_inverse_exact(2, i + 1)) for i in range(4096)], np.float32)
    assert np.array_equal(seq, exp)
    assert list(seq[:3]) == [0.5, 0.25, 0.75]


@pytest.mark.parametrize("base", [3, 5, 7, 11, 13])
def test_halton_incremental_close_to_radical_inverse(base):
    """Incremental getNext (mcqmc.h:35-49) tracks the exact radical inverse to
    a few 1e-7 everywhere; it may differ in the last bits (carry tolerance)."""
    n = 60000
    seq = po.halton_seq(base, n).astype(np.float64)
    idx = np.arange(1, n + 1)
    sel = idx[:: 97]
    exact = np.array([float(radical_inverse_exact(base, int(i))) for i in sel])
    assert np.max(np.abs(seq[sel - 1] - exact)) < 2e-5
    # most indices agree bit for bit with the setStart digit loop
    L = po.lib()
    at = np.array([L.yo_halton_at(base, int(i)) for i in sel], np.float32)
    assert np.mean(at == seq[sel - 1].astype(np.float32)) > 0.99


def test_ri_vdc_is_bit_reversal():
    L = po.lib()
    for i in [1, 2, 3, 5, 1234567, 0xDEADBEEF]:
        rev = int("{:032b}".format(i)[::-1], 2)
        assert L.yo_ri_vdc(i, 0) == np.float32(rev / 2.0 ** 32)


def test_ri_sobol_and_lp_first_values():
    L = po.lib()
    assert [L.yo_ri_s(i, 0) for i in range(4)] == [0.0, 0.5, 0.75, 0.25]
    # Larcher-Pillichshammer: generator columns 1, 11, 111, ... (binary, MSB first)
    assert [L.yo_ri_lp(i, 0) for i in range(4)] == [0.0, 0.5, 0.75, 0.25]
    assert L.yo_ri_lp(4, 0) == np.float32(0.875)
    assert L.yo_ri_s(4, 0) == np.float32(0.625)


def test_next_prime_sequence():
    L = po.lib()
    p, out = 2, [2]
    for _ in range(10):
        p = L.yo_next_prime(p)
        out.append(p)
    assert out == [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31]


# ------------------------------------------------------------- triangle --

def _f(*v):
    return np.array(v, np.float32)


def test_triangle_closed_forms():
    L = po.lib()
    a, b, c = _f(0, 0, 0), _f(1, 0, 0), _f(0, 1, 0)
    n = np.zeros(3, np.float32)
    L.yo_tri_normal(a, b, c, n)
    assert list(n) == [0, 0, 1]
    frm, d = _f(0.25, 0.25, 2), _f(0, 0, -1)
    assert L.yo_tri_hit(a, b, c, n, frm, d) == 1
    assert L.yo_tri_intersect(a, n, frm, d) == 2.0
    # the test is a line test: a ray pointing away still "hits", t < 0
    assert L.yo_tri_hit(a, b, c, n, frm, -d) == 1
    assert L.yo_tri_intersect(a, n, frm, -d) == -2.0
    # outside the triangle
    assert L.yo_tri_hit(a, b, c, n, _f(0.75, 0.75, 2), d) == 0
    # a vertex and an edge point are inside (products == 0 pass)
    assert L.yo_tri_hit(a, b, c, n, _f(0, 0, 2), d) == 1
    assert L.yo_tri_hit(a, b, c, n, _f(0.5, 0.5, 2), d) == 1


def test_shared_edge_is_watertight():
    """Edge products of two consistently wound triangles sharing an edge are
    exact negatives, so a ray crossing both from the same side through the
    shared edge hits at least one of them (triangle.h:48-58)."""
    L = po.lib()
    rng = np.random.default_rng(1)
    for _ in range(2000):
        p = rng.normal(size=(4, 3)).astype(np.float32)
        n1, n2 = np.zeros(3, np.float32), np.zeros(3, np.float32)
        L.yo_tri_normal(p[0], p[1], p[2], n1)
        L.yo_tri_normal(p[0], p[2], p[3], n2)
        w = rng.uniform(0, 1)
        q = (p[0] * w + p[2] * (1 - w)).astype(np.float32)
        frm = (q + rng.normal(size=3)).astype(np.float32)
        d = (q - frm).astype(np.float32)
        if np.sign(np.dot(d, n1)) != np.sign(np.dot(d, n2)):
            continue
        h = L.yo_tri_hit(p[0], p[1], p[2], n1, frm, d) + L.yo_tri_hit(p[0], p[2], p[3], n2, frm, d)
        assert h >= 1


def _rays(n, lo, hi, seed, tmax=np.inf):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def test_bvh_equals_brute_force_room():
    d = S.direct_room(32, 24, subdiv=3)
    sc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    rays = _rays(20000, -2.5, 2.5, 3)
    rays[:, 2] += 3
    pb, tb, ub, vb = sc.trace(rays, brute=True)
    pv, tv, uv, vv = sc.trace(rays)
    assert np.array_equal(pb, pv)
    m = pb >= 0
    assert np.array_equal(tb[m], tv[m]) and np.array_equal(ub[m], uv[m])
    rays[:, 7] = 1.0
    ab, _, _, _ = sc.trace(rays, mode=1, brute=True)
    av, _, _, _ = sc.trace(rays, mode=1)
    assert np.array_equal(ab >= 0, av >= 0)
    # any(tmax) <=> closest t < tmax
    rays[:, 7] = np.inf
    pc, tc, _, _ = sc.trace(rays)
    assert np.array_equal(av >= 0, (pc >= 0) & (tc < 1.0) & (tc > 0))


def test_bvh_equals_brute_force_grazing_terrain():
    """Camera rays over the full 1,045,506-triangle fBm terrain: a ray that
    crosses a terrain cell near z = 0 at a shallow angle is accepted by the
    float edge test while passing a few 1e-8 outside the cell's exact box, so
    a box inflation relative to the box's own coordinates alone loses it (the
    oracle once answered "miss" here)."""
    d = S.terrain_gi(po.heightfield)
    sc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    h = [float.fromhex(x) for x in ("0x1p+0", "-0x1.cccccc0000000p-2", "0x1.19999a0000000p+0",
                                    "0x1.516f800000000p-4", "0x1.65bef60000000p-1",
                                    "-0x1.6bd8260000000p-1")]
    v = np.array(h, np.float32)
    p, t = sc.first_hit(v[:3], v[3:], d.opts.min_raydis)
    f = v[:3] + v[3:] * np.float32(d.opts.min_raydis)
    ray = np.array([[*f, 0, *v[3:], np.inf]], np.float32)
    pb, tb, _, _ = sc.trace(ray, brute=True)
    assert pb[0] == 700890 and p[0] == pb[0] and t[0] == tb[0]
    # a sample of grazing camera rays toward the far terrain
    rng = np.random.default_rng(5)
    n = 400
    dirs = np.stack([rng.uniform(-0.45, 0.45, n), np.full(n, 0.75), rng.uniform(-0.62, -0.5, n)], 1)
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = v[:3]
    rays[:, 4:7] = dirs
    rays[:, 7] = np.inf
    pb, tb, _, _ = sc.trace(rays, brute=True)
    pv, tv, _, _ = sc.trace(rays)
    assert np.mean(pb >= 0) > 0.5
    assert np.array_equal(pb, pv) and np.array_equal(tb[pb >= 0], tv[pb >= 0])


def test_barycentrics_reconstruct_hit_point():
    d = S.direct_room(32, 24, subdiv=2)
    sc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    rays = _rays(3000, -2, 2, 5)
    rays[:, 2] += 3
    p, t, u, v = sc.trace(rays)
    m = p >= 0
    tri = d.verts[d.idx[p[m]]].astype(np.float64)
    rec = tri[:, 0] * (1 - u[m] - v[m])[:, None] + tri[:, 1] * u[m][:, None] + tri[:, 2] * v[m][:, None]
    hit = rays[m, :3].astype(np.float64) + rays[m, 4:7] * t[m][:, None]
    assert np.max(np.abs(rec - hit)) < 1e-4
    assert np.all(u[m] > -1e-4) and np.all(v[m] > -1e-4) and np.all(u[m] + v[m] < 1 + 1e-4)


def test_first_hit_and_shadow_semantics():
    """firstHit offsets the origin by min_raydis; isShadowed uses self_bias for
    the shading point's own object only (scene.cc:107-185)."""
    # two parallel unit quads (objects 0 and 1) at z=0 and z=1
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]], np.uint32)
    obj = np.array([0, 0, 1, 1], np.int32)
    sc = po.OracleScene(v, f, obj, None)
    prim, t = sc.first_hit(np.array([0.5, 0.5, 0.5], np.float32), np.array([0, 0, 1], np.float32))
    assert prim[0] in (2, 3) and abs(t[0] - (0.5 - 5e-5)) < 1e-6
    p = np.array([[0.5, 0.5, 0.0]], np.float32)
    up = np.array([[0, 0, 1]], np.float32)
    assert sc.is_shadowed(p, up, [0], 0)[0] == 1          # blocked by object 1
    assert sc.is_shadowed(p, np.array([[0.5, 0.5, 0.9]], np.float32), [0], 1)[0] == 0  # point below it
    # own-object geometry closer than self_bias is skipped, other objects are not
    q = np.array([[0.5, 0.5, 0.95]], np.float32)
    assert sc.is_shadowed(q, up, [0], 0, self_bias=0.1)[0] == 1
    assert sc.is_shadowed(q, up, [1], 0, self_bias=0.1)[0] == 0


# ---------------------------------------------------------------- noise --

def perlin_grad(h, x, y, z):
    """Perlin's 12 gradient directions selected by the low 4 hash bits."""
    h &= 15
    u = x if h < 8 else y
    v = y if h < 4 else (x if h in (12, 14) else z)
    return (u if (h & 1) == 0 else -u) + (v if (h & 2) == 0 else -v)


def test_perlin_lattice_and_gradients():
    L = po.lib()
    # value 0.5 at every lattice point
    for p in [(0, 0, 0), (3, -7, 12), (255, 1, 2), (-1, -1, -1)]:
        assert L.yo_newperlin(*map(float, p)) == 0.5
    # period 256
    rng = np.random.default_rng(2)
    for _ in range(200):
        p = rng.uniform(-50, 50, 3)
        assert L.yo_newperlin(*p) == pytest.approx(L.yo_newperlin(p[0] + 256, p[1], p[2]), abs=2e-5)
    # derivative at a lattice corner = 0.5 * gradient of that corner's hash
    eps = 1e-3
    for p in [(0, 0, 0), (1, 2, 3), (10, 20, 30), (100, 7, 250)]:
        x, y, z = map(float, p)
        g = np.array([(L.yo_newperlin(x + eps, y, z) - L.yo_newperlin(x - eps, y, z)) / (2 * eps),
                      (L.yo_newperlin(x, y + eps, z) - L.yo_newperlin(x, y - eps, z)) / (2 * eps),
                      (L.yo_newperlin(x, y, z + eps) - L.yo_newperlin(x, y, z - eps)) / (2 * eps)])
        # gradient vector of this corner: evaluate grad() on unit axes
        hh = _corner_hash(int(x), int(y), int(z))
        exp = 0.5 * np.array([perlin_grad(hh, 1, 0, 0), perlin_grad(hh, 0, 1, 0), perlin_grad(hh, 0, 0, 1)])
        assert np.allclose(g, exp, atol=2e-3), (p, g, exp)


PERM_FIRST = [0xA2, 0xA0, 0x19, 0x3B, 0xF8, 0xEB, 0xAA, 0xEE]


def _perm():
    buf = (ctypes.c_ubyte * 512)()
    po.lib().yo_perm(buf)
    return list(buf)


def _corner_hash(X, Y, Z):
    P = _perm()
    X &= 255; Y &= 255; Z &= 255
    A = P[X] + Y
    AA = P[A] + Z
    return P[AA]


def test_perm_table_is_a_permutation():
    P = _perm()
    assert len(P) == 512 and P[:256] == P[256:]
    assert sorted(P[:256]) == list(range(256))
    assert P[:8] == PERM_FIRST


def test_fbm_and_turbulence_compose_noise():
    L = po.lib()
    rng = np.random.default_rng(4)
    for _ in range(100):
        x, y, z = rng.uniform(-10, 10, 3).astype(np.float32)
        n1 = L.yo_newperlin(x, y, z)
        n2 = L.yo_newperlin(2 * x, 2 * y, 2 * z)
        n3 = L.yo_newperlin(4 * x, 4 * y, 4 * z)
        assert L.yo_fbm(x, y, z, 1.0, 2.0, 1.0) == pytest.approx(2 * n1 - 1, abs=1e-6)
        assert L.yo_fbm(x, y, z, 1.0, 2.0, 2.0) == pytest.approx((2 * n1 - 1) + 0.5 * (2 * n2 - 1), abs=1e-6)
        # fractional octave: last term weighted by the remainder
        assert L.yo_fbm(x, y, z, 1.0, 2.0, 2.5) == pytest.approx(
            (2 * n1 - 1) + 0.5 * (2 * n2 - 1) + 0.5 * 0.25 * (2 * n3 - 1), abs=1e-6)
        # turbulence(oct=1, size=1) = (n1 + 0.5 n2) * 2/3
        assert L.yo_turbulence(x, y, z, 1, 1.0, 0) == pytest.approx((n1 + 0.5 * n2) * 2 / 3, abs=1e-6)
        assert L.yo_turbulence(x, y, z, 1, 1.0, 1) == pytest.approx(
            (abs(2 * n1 - 1) + 0.5 * abs(2 * n2 - 1)) * 2 / 3, abs=1e-6)


def test_heightfield_grid():
    hf = po.heightfield(5, 3, 2.0, 1.0, 0.4, 4.0, 1.0, 2.0, 6.0)
    assert hf.shape == (15, 3)
    assert np.allclose(hf[:5, 0], [0, 0.5, 1.0, 1.5, 2.0]) and np.allclose(hf[::5, 1], [0, 0.5, 1.0])
    L = po.lib()
    assert hf[7, 2] == np.float32(np.float32(0.4) * L.yo_fbm(np.float32(4.0), np.float32(2.0), 0.0, 1.0, 2.0, 6.0))


# ---------------------------------------------------- marble and wood --

def _tex_float(tex, pts):
    return po.noise_eval(3, tex, np.asarray(pts, np.float32))


def test_marble_wood_closed_forms():
    """Without turbulence the textures are closed forms of the point
    (basictex.cc:133-153, 194-216): marble = (1/2 + 1/2 sin(5 s))^(1/sharpness)
    with s = x + y + z; wood = 1/2 + 1/2 sin(20 |p|) with rings, sin(10 s)
    without; the saw shape is frac(w / 2 pi), the triangle |2 frac(w / 2 pi) - 1|.
    Points are chosen at the sine's peaks, zeros and quarter periods."""
    from yafray_amd import descriptors as D
    # marble, sine shape, sharpness 1 and 2 (exponent 1/2)
    s = np.array([math.pi / 10, -math.pi / 10, 0.0, math.pi / 20, 3 * math.pi / 10])
    pts = np.stack([s / 3, s / 3, s / 3], 1)
    w = 0.5 + 0.5 * np.sin(5 * s)
    for sharp, ex in ((1.0, 1.0), (2.0, 0.5)):
        got = _tex_float(D.marble(turbulence=0.0, sharpness=sharp), pts)
        assert np.allclose(got, w ** ex, atol=2e-6), (sharp, got, w ** ex)
    # marble saw and triangle: w = 5 s / (2 pi)
    s2 = np.array([0.1, 0.5, 1.3, 2.9, 4.4])
    pts2 = np.stack([s2, np.zeros(5), np.zeros(5)], 1)
    f = 5 * s2 / (2 * math.pi)
    fr = f - np.floor(f)
    assert np.allclose(_tex_float(D.marble(turbulence=0.0, shape=1), pts2), fr, atol=2e-6)
    assert np.allclose(_tex_float(D.marble(turbulence=0.0, shape=2), pts2), np.abs(2 * fr - 1), atol=4e-6)
    # wood with rings: the radius alone matters
    r = np.array([math.pi / 40, math.pi / 20, 3 * math.pi / 40, 0.3, 0.77])
    dirs = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0.6, 0.8, 0], [0, 0.6, -0.8]])
    got = _tex_float(D.wood(turbulence=0.0, rings=True), dirs * r[:, None])
    assert np.allclose(got, 0.5 + 0.5 * np.sin(20 * r), atol=2e-6)
    # wood without rings: 10 s
    got = _tex_float(D.wood(turbulence=0.0), pts)
    assert np.allclose(got, 0.5 + 0.5 * np.sin(10 * s), atol=2e-6)
    # the colour is color1 + f (color2 - color1) (basictex.cc:155-159)
    t = D.marble(color1=(0.1, 0.2, 0.3), color2=(0.9, 0.6, 0.5), turbulence=0.0)
    c = po.noise_eval(4, t, pts.astype(np.float32)).reshape(-1, 3)
    c1, c2 = np.array([0.1, 0.2, 0.3]), np.array([0.9, 0.6, 0.5])
    assert np.allclose(c, c1 + w[:, None] * (c2 - c1), atol=3e-6)


def test_marble_wood_turbulence_term():
    """With turbulence t the phase gains t x turbulence(p) (the turbulence
    itself is pinned above): at points with s = 0 marble's sine shape gives
    1/2 + 1/2 sin(t T(p)); scaling t scales the phase."""
    from yafray_amd import descriptors as D
    rng = np.random.default_rng(5)
    xy = rng.uniform(-3, 3, (64, 2))
    pts = np.c_[xy, -xy.sum(1)].astype(np.float32)   # x + y + z = 0 up to rounding
    assert np.max(np.abs(pts.astype(np.float64).sum(1))) < 1e-5
    # turbulence(p, depth, size, hard) of the same parameters (any texture carries them)
    T = po.noise_eval(2, D.clouds(depth=3, size=1.7), pts)
    for tb in (0.5, 2.0):
        got = _tex_float(D.marble(turbulence=tb, depth=3, size=1.7), pts)
        exp = 0.5 + 0.5 * np.sin(5 * pts.astype(np.float64).sum(1) + tb * T.astype(np.float64))
        assert np.allclose(got, exp, atol=5e-6), tb


# ------------------------------------------- stdPerlin and Blender noise --

def _grad_table():
    g = np.zeros(768, np.float32)
    po.lib().yo_grad_table(g)
    return g.reshape(256, 3).astype(np.float64)


def _perm_np():
    return np.array(_perm(), np.int64)


def test_gradient_table_is_256_unit_vectors():
    """Perlin's original gradients (noise.cc:200-370) are normalised random
    directions: 256 distinct rows of length 1 to the table's 6 digits."""
    g = _grad_table()
    assert np.allclose(np.linalg.norm(g, axis=1), 1.0, atol=1e-4)
    assert len(np.unique(g.round(6), axis=0)) == 256


def test_stdperlin_lattice_and_gradients():
    """stdPerlin_t (noise.cc:372-435) is 1/2 at lattice points and its
    gradient there is 0.75 x the unit vector of row
    (p[p[bx] + by] + bz) & 255, b = (coordinate + 10000) & 255."""
    L = po.lib()
    g, H = _grad_table(), _perm_np()
    # one ulp of the shifted coordinate (x + 10000 in float): the step stays
    # exact, and the other corners' O(h^2) weights bias the difference by ~1.5 h
    h = 2.0 ** -10
    rng = np.random.default_rng(11)
    for X, Y, Z in rng.integers(-40, 40, (12, 3)):
        assert L.yo_stdperlin(float(X), float(Y), float(Z)) == 0.5
        b = [(int(c) + 10000) & 255 for c in (X, Y, Z)]
        row = (H[H[b[0]] + b[1]] + b[2]) & 255
        d = [(L.yo_stdperlin(X + h * e[0], Y + h * e[1], Z + h * e[2])
              - L.yo_stdperlin(X - h * e[0], Y - h * e[1], Z - h * e[2])) / (2 * h)
             for e in np.eye(3)]
        assert np.allclose(d, 0.75 * g[row], atol=0.004), (X, Y, Z, d, 0.75 * g[row])


def test_blender_noise_lattice_and_gradients():
    """blenderNoise_t (noise.cc:439-509) is 1/2 at lattice points; the corner
    weights 1-3t^2+2|t|^3 have zero slope there, so the gradient is the unit
    vector hashvectf[p[(z & 255) + p[p[x & 255] + (y & 255)]]]."""
    L = po.lib()
    g, H = _grad_table(), _perm_np()
    h = 2.0 ** -12   # the other corners enter with weight 3h^2: a ~1.5 h bias
    rng = np.random.default_rng(12)
    for X, Y, Z in rng.integers(-40, 40, (12, 3)):
        assert L.yo_blendernoise(float(X), float(Y), float(Z)) == 0.5
        row = H[(Z & 255) + H[H[X & 255] + (Y & 255)]]
        d = [(L.yo_blendernoise(X + h * e[0], Y + h * e[1], Z + h * e[2])
              - L.yo_blendernoise(X - h * e[0], Y - h * e[1], Z - h * e[2])) / (2 * h)
             for e in np.eye(3)]
        assert np.allclose(d, g[row], atol=0.002), (X, Y, Z, d, g[row])


def test_turbulence_and_fbm_over_each_generator():
    """turbulence sums octaves of the texture's generator, Blender's at
    p + (1,1,1) (noise.h:60-61, noise.cc:808-819); fBm sums signed octaves
    without the offset (noise.cc:638-650)."""
    from yafray_amd import descriptors as D
    L = po.lib()
    gens = {1: L.yo_stdperlin, 2: L.yo_blendernoise}
    pts = np.random.default_rng(13).uniform(-5, 5, (40, 3)).astype(np.float32)
    for kind, fn in gens.items():
        t = D.clouds(depth=3, size=1.3, hard=False)
        t.noise = kind
        got = po.noise_eval(2, t, pts)
        for q, p in enumerate(pts):
            x = (p + 1.0).astype(np.float32) if kind == 2 else p.copy()
            x = (x * np.float32(1.3)).astype(np.float32)
            s, amp = np.float32(0), np.float32(1)
            for _ in range(4):
                s = np.float32(s + amp * np.float32(fn(*map(float, x))))
                amp, x = np.float32(amp * 0.5), (x * np.float32(2)).astype(np.float32)
            exp = np.float32(s * np.float32(np.float32(8) / np.float32(15)))
            assert got[q] == exp, (kind, q, got[q], exp)
        m = D.musgrave(H=0.8, lacunarity=2.1, octaves=3.5)
        m.noise = kind
        f = po.noise_eval(1, m, pts)
        pw = np.float32(np.power(2.1, -0.8))
        for q, p in enumerate(pts[:10]):
            x, val, pwr = p.copy(), np.float32(0), np.float32(1)
            for _ in range(3):
                val = np.float32(val + np.float32(np.float32(2) * np.float32(fn(*map(float, x))) - 1) * pwr)
                pwr, x = np.float32(pwr * pw), (x * np.float32(2.1)).astype(np.float32)
            val = np.float32(val + np.float32(0.5) * np.float32(np.float32(2) * np.float32(fn(*map(float, x))) - 1) * pwr)
            assert abs(f[q] - val) <= 1e-6 * max(1.0, abs(val)), (kind, q, f[q], val)
