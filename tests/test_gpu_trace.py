"""GPU parity of the ray queries (libyafgpu.so through its C ABI) against the
CPU oracle: primitive ids, distances and barycentrics bit for bit."""
import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import descriptors as D
from yafray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _lib(torch_cuda):
    from yafray_amd import lib as Y
    return Y


def _rays(n, lo, hi, seed, tmax=np.inf, zshift=0.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    o[:, 2] += zshift
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def _gpu_trace(Y, torch, scene, rays):
    dr = torch.from_numpy(rays).cuda()
    hits = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
    scene.trace_closest(dr, hits)
    torch.cuda.synchronize()
    h = hits.cpu().numpy()
    return h[:, 0].view(np.int32).copy(), h[:, 1].copy(), h[:, 2].copy(), h[:, 3].copy()


def _check(prim_g, t_g, u_g, v_g, prim_c, t_c, u_c, v_c):
    n = len(prim_c)
    match = np.mean(prim_g == prim_c)
    assert match >= 0.9999, "primID match %.6f" % match
    m = (prim_g == prim_c) & (prim_c >= 0)
    # identical float operations on both sides: distances are bit-exact
    rel = np.abs(t_g[m] - t_c[m]) / np.maximum(np.abs(t_c[m]), 1e-30)
    assert np.max(rel, initial=0) <= 1e-5
    assert np.mean(t_g[m] == t_c[m]) >= 0.9999
    assert np.mean((u_g[m] == u_c[m]) & (v_g[m] == v_c[m])) >= 0.9999
    return match, n


def test_library_loads(torch_cuda):
    Y = _lib(torch_cuda)
    assert Y.lib().yg_abi_version() == 14


@pytest.mark.parametrize("builder", [D.YG_BUILDER_HOST, D.YG_BUILDER_DEVICE, D.YG_BUILDER_DEVICE_PLOC])
@pytest.mark.parametrize("seed,width", [(0, 2), (1, 2), (0, 4), (1, 4), (0, 8), (1, 8)])
def test_closest_any_room(torch_cuda, seed, width, builder):
    torch = torch_cuda
    Y = _lib(torch)
    d = S.direct_room(64, 48, subdiv=5)
    sc = Y.Scene.from_desc(d, width=width, builder=builder)
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    rays = _rays(200000, -2.9, 2.9, seed, zshift=3.0)
    pc, tc, uc, vc = oc.trace(rays, threads=8)
    pg, tg, ug, vg = _gpu_trace(Y, torch, sc, rays)
    _check(pg, tg, ug, vg, pc, tc, uc, vc)
    # any-hit with a finite tmax
    rays[:, 7] = 0.75
    ac, _, _, _ = oc.trace(rays, mode=1, threads=8)
    occ = torch.empty(len(rays), dtype=torch.uint8, device="cuda")
    sc.trace_any(torch.from_numpy(rays).cuda(), occ)
    ag = occ.cpu().numpy()
    assert np.mean(ag == (ac >= 0)) >= 0.9999


@pytest.mark.parametrize("width", [2, 4, 8])
def test_tmin_and_empty(torch_cuda, width):
    torch = torch_cuda
    Y = _lib(torch)
    # one triangle, rays with tmin excluding the hit
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]], np.uint32)
    sc = Y.Scene()
    sc.add_mesh(v, f)
    sc.build(width=width)
    rays = np.array([[0.2, 0.2, 1, 0, 0, 0, -1, np.inf],
                     [0.2, 0.2, 1, 1.5, 0, 0, -1, np.inf],
                     [0.2, 0.2, 1, 0, 0, 0, -1, 0.5],
                     [0.9, 0.9, 1, 0, 0, 0, -1, np.inf]], np.float32)
    p, t, u, vv = _gpu_trace(Y, torch, sc, rays)
    assert list(p) == [0, -1, -1, -1]
    assert t[0] == 1.0 and abs(u[0] - 0.2) < 1e-6 and abs(vv[0] - 0.2) < 1e-6
    # an empty scene builds and misses everything
    e = Y.Scene()
    e.build(width=width)
    p, t, _, _ = _gpu_trace(Y, torch, e, rays)
    assert np.all(p == -1)


@pytest.mark.parametrize("builder", [D.YG_BUILDER_HOST, D.YG_BUILDER_DEVICE, D.YG_BUILDER_DEVICE_PLOC])
@pytest.mark.parametrize("width", [2, 4, 8])
def test_terrain_parity_full_size(torch_cuda, width, builder):
    """1,045,506-triangle fBm terrain: 1M incoherent rays on the GPU, a
    100k-ray sample checked against the oracle; the tree built on the host or
    on the device."""
    torch = torch_cuda
    Y = _lib(torch)
    verts, idx, band = S.terrain_mesh(1024, 512, Y.device_heights)
    # device heightfield == oracle heightfield, bit for bit
    hv = po.heightfield(1024, 512, 2.0, 1.0, 0.4, 4.0, 1.0, 2.0, 6.0)
    assert np.array_equal(hv, verts)
    sc = Y.Scene()
    sc.add_mesh(verts, idx, band)
    sc.build(width=width, builder=builder)
    info = sc.info()
    assert info["tris"] == 1045506
    rng = np.random.default_rng(7)
    n = 1 << 20
    o = np.stack([rng.uniform(0, 2, n), rng.uniform(0, 1, n), rng.uniform(0.3, 0.8, n)], 1)
    d = rng.normal(size=(n, 3))
    d[:, 2] = -np.abs(d[:, 2]) - 0.2
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    pg, tg, ug, vg = _gpu_trace(Y, torch, sc, rays)
    sub = rng.choice(n, 100000, replace=False)
    oc = po.OracleScene(verts, idx, None, band)
    pc, tc, uc, vc = oc.trace(rays[sub], threads=8)
    _check(pg[sub], tg[sub], ug[sub], vg[sub], pc, tc, uc, vc)
    # a good part of the downward rays leave the [0,2]x[0,1] terrain before landing
    assert np.mean(pg >= 0) > 0.3
    # any hit with a finite length: this tree's traversal stack (max depth
    # well past the 16 entries kept in LDS) spills into the overflow buffer
    assert info["max_depth"] > 16
    rays[:, 7] = 0.8
    occ = torch.empty(n, dtype=torch.uint8, device="cuda")
    sc.trace_any(torch.from_numpy(rays).cuda(), occ)
    ag = occ.cpu().numpy()
    ac, _, _, _ = oc.trace(rays[sub], mode=1, threads=8)
    assert np.mean(ag[sub] == (ac >= 0)) >= 0.9999
    assert 0.05 < np.mean(ag) < 0.95


@pytest.mark.parametrize("builder", [D.YG_BUILDER_HOST, D.YG_BUILDER_DEVICE, D.YG_BUILDER_DEVICE_PLOC])
@pytest.mark.parametrize("width", [2, 4, 8])
def test_first_hit_is_shadowed(torch_cuda, width, builder):
    torch = torch_cuda
    Y = _lib(torch)
    d = S.direct_room(64, 48, subdiv=4)
    sc = Y.Scene.from_desc(d, width=width, builder=builder)
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    rng = np.random.default_rng(3)
    n = 50000
    frm = rng.uniform(-2.5, 2.5, (n, 3)).astype(np.float32)
    frm[:, 2] += 3
    dirs = rng.normal(size=(n, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    pc, tc = oc.first_hit(frm, dirs)
    hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    sc.first_hit(torch.from_numpy(frm).cuda(), torch.from_numpy(dirs).cuda(), hits)
    h = hits.cpu().numpy()
    pg = h[:, 0].view(np.int32)
    assert np.mean(pg == pc) >= 0.9999
    m = (pg == pc) & (pc >= 0)
    assert np.array_equal(h[m, 1], tc[m])
    # shading points on surfaces, shadow rays to points and directions
    p = frm + dirs * tc[:, None]
    ok = pc >= 0
    p, so = p[ok].astype(np.float32), d.tri_obj[pc[ok]].astype(np.int32)
    m = len(p)
    lp = rng.uniform(-2.5, 2.5, (m, 3)).astype(np.float32)
    lp[:, 2] += 3
    ld = rng.normal(size=(m, 3)).astype(np.float32)
    for to_point, l in ((1, lp), (0, ld)):
        for bias in (1e-3, 0.1):
            oc_o = oc.is_shadowed(p, l, so, to_point, self_bias=bias)
            out = torch.empty(m, dtype=torch.uint8, device="cuda")
            sc.is_shadowed(torch.from_numpy(p).cuda(), torch.from_numpy(l).cuda(),
                           torch.from_numpy(so).cuda(), out, to_point, self_bias=bias)
            assert np.mean(out.cpu().numpy() == oc_o) >= 0.9999


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("what", [0, 1, 2, 3, 4])
def test_noise_parity(torch_cuda, what, kind):
    """newPerlin_t, stdPerlin_t and blenderNoise_t (kind 0, 1, 2) through every
    evaluation, the textures including the four Musgrave multifractals:
    bit-exact with the oracle."""
    torch = torch_cuda
    Y = _lib(torch)
    rng = np.random.default_rng(11)
    pts = rng.uniform(-40, 40, (100000, 3)).astype(np.float32)
    texs = [D.musgrave(H=1.0, lacunarity=2.0, octaves=6.0, size=0.37, intensity=1.3),
            D.clouds(depth=3, size=0.8, hard=True, bias=2),
            D.clouds(depth=4, size=1.7),
            D.marble(depth=3, turbulence=2.5, sharpness=2.0, size=0.6, shape=0),
            D.marble(depth=2, turbulence=1.0, size=1.1, shape=2),
            D.wood(depth=2, turbulence=3.0, size=0.9, rings=True, shape=1),
            D.wood(depth=2, turbulence=0.0, size=0.9, rings=False, shape=0)]
    texs += [D.musgrave(H=0.8, lacunarity=2.1, octaves=o, size=0.45, intensity=0.9, mtype=mt, offset=0.9, gain=1.6)
             for mt, o in (("multifractal", 5.5), ("heteroterrain", 4.5), ("hybridmf", 6.25), ("ridgedmf", 5.0))]
    if what <= 2:
        texs = [texs[0], texs[1]]
    for tex in texs:
        tex.noise = kind
        out = torch.empty(len(pts) * (3 if what == 4 else 1), dtype=torch.float32, device="cuda")
        Y.noise_eval(what, tex, torch.from_numpy(pts).cuda(), out)
        g = out.cpu().numpy()
        c = po.noise_eval(what, tex, pts)
        assert np.array_equal(g, c), (what, tex.type, np.max(np.abs(g - c)))


def test_halton_tables(torch_cuda):
    torch = torch_cuda
    Y = _lib(torch)
    for base in (2, 3, 5, 7, 11, 13, 17, 19):
        out = torch.empty(50000, dtype=torch.float32, device="cuda")
        Y.halton(base, 100, 50000, out)
        assert np.array_equal(out.cpu().numpy(), po.halton_seq(base, 50100)[100:])
