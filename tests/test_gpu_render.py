"""GPU renders (yg_render through the C ABI) against the CPU oracle's
restatement of scene_t::render: same Halton / ourRandom() sample streams, so
images agree pixel for pixel up to float rounding of transcendentals."""
import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import descriptors as D
from yafray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _Y():
    from yafray_amd import lib as Y
    return Y


def gpu_render(desc, torch, build_kw=None):
    Y = _Y()
    sc = Y.Scene.from_desc(desc, **(build_kw or {}))
    out = torch.zeros(desc.cam.resx * desc.cam.resy * 3, dtype=torch.float32, device="cuda")
    st = sc.render(out)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(desc.cam.resy, desc.cam.resx, 3), st, sc


def rel_rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)) / max(np.sqrt(np.mean(b.astype(np.float64) ** 2)), 1e-30))


def small_terrain(nx=192, ny=96, resx=160, resy=90, samples=4, depth=3, nmat=64, mods=None):
    return S.terrain_gi(po.heightfield, nx=nx, ny=ny, resx=resx, resy=resy, samples=samples,
                        depth=depth, nmat=nmat)


def test_gi_terrain_small_matches_oracle(torch_cuda):
    torch = torch_cuda
    d = small_terrain()
    img, st, _ = gpu_render(d, torch)
    ref, rs = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat).render(d, threads=8)
    assert np.all(np.isfinite(img))
    assert rel_rmse(img, ref) <= 1e-3
    assert np.mean(np.all(img == ref, axis=2)) > 0.99
    # identical control flow: same number of camera, extension and shadow rays
    assert st.rays_camera == rs[0]
    assert st.rays_extension == rs[1]
    assert st.rays_shadow == rs[2]
    assert st.adaptive_extra == 0


def test_direct_room_area_light_matches_oracle(torch_cuda):
    torch = torch_cuda
    d = S.direct_room(160, 120, subdiv=4)
    img, st, _ = gpu_render(d, torch)
    ref, rs = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat).render(d)
    assert rel_rmse(img, ref) <= 1e-3
    assert np.mean(np.all(img == ref, axis=2)) > 0.99
    assert st.rays_camera == rs[0] and st.rays_shadow == rs[2]


def test_plane_gi_closed_form(torch_cuda):
    """An infinite-looking plane under an empty sky: every path sample escapes,
    so each pixel is rho*[(1/S) sum_i (w_i.N) B] + rho (N.s) E with w_i the
    cosine-weighted Halton direction of sample i (pathlight.cc:305-347,
    pathtools.cc:32-43, sunlight.cc:35-57).  (w_i.N) = sqrt(z1) with z1 the
    base-2 Halton value of the pixel's per-area call index."""
    torch = torch_cuda
    E = 50.0
    v = np.array([[-E, -E, 0], [E, -E, 0], [E, E, 0], [-E, E, 0]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    rho = (0.6, 0.5, 0.3)
    S_ = 4
    sun = np.array([1.0, 0.5, 1.5])
    sun /= np.linalg.norm(sun)
    lights = [D.pathlight(samples=S_, depth=2, use_qmc=True, threshold=1e3), D.sunlight(sun, power=2.0)]
    resx, resy = 96, 80
    cam = D.camera((0, 0, 3), (0, 0.01, 0), (0, 1, 4), resx, resy, 1.0, 1.0)
    opts = D.render_opts(bias=1e-3, background=(0.4, 0.5, 0.7))
    d = D.SceneDesc(v, f, [0, 0], [0, 0], [7], cam, [D.material(rho)], [], [], lights, opts)
    img, st, _ = gpu_render(d, torch)
    # expected value from the block layout of renderblock.cc:71-101
    B = np.array([0.4, 0.5, 0.7])
    rho_ = np.array(rho, np.float64)
    ndots = float(np.dot([0, 0, 1], sun))
    exp = np.zeros((resy, resx, 3))
    blk = 64
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
                    k = (i - y) * w + (j - x)
                    z1 = [float(radinv2(k * S_ + s + 1)) for s in range(S_)]
                    path = rho_ * (np.sum(np.sqrt(z1)) / S_) * B
                    exp[i, j] = path + rho_ * ndots * 2.0
    assert np.max(np.abs(img - exp) / exp) < 2e-5


def radinv2(i):
    r, f = 0.0, 0.5
    while i:
        r += (i & 1) * f
        i >>= 1
        f *= 0.5
    return r


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_gi_textured_musgrave_matches_oracle(torch_cuda, kind):
    """Albedo modulated by a Musgrave fBm texture (mode mul) on the terrain, its
    fBm over newPerlin_t, stdPerlin_t or blenderNoise_t (kind 0, 1, 2)."""
    torch = torch_cuda
    d = small_terrain(nx=96, ny=48, resx=96, resy=64, samples=4, depth=3, nmat=8)
    texs = [D.musgrave(color1=(0.2, 0.2, 0.2), color2=(1, 1, 1), octaves=4.0, size=1.0)]
    texs[0].noise = kind
    mods = [D.modulator(0, D.YG_MOD_MUL, color=0.7, size=3.0)]
    mats = [D.material(tuple(m.color), 0, 1) for m in d.mats]
    d = D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, d.cam, mats, mods, texs,
                    list(d.lights), d.opts)
    img, st, _ = gpu_render(d, torch)
    ref, rs = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat).render(d, threads=8)
    assert rel_rmse(img, ref) <= 1e-3
    assert st.rays_extension == rs[1]


def test_gi_linear_in_sun_power(torch_cuda):
    """The estimator is linear in the sun's emission (every path vertex's light
    is the sun term): doubling the power doubles the sun-lit part exactly."""
    torch = torch_cuda
    d = small_terrain(nx=64, ny=32, resx=64, resy=48)
    d.opts.has_background = 0
    a, _, _ = gpu_render(d, torch)
    L = list(d.lights)
    L[1].power *= 2.0
    d2 = D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, d.cam, list(d.mats), [], [], L, d.opts)
    b, _, _ = gpu_render(d2, torch)
    assert np.allclose(b, 2 * a, rtol=1e-5, atol=1e-7)


def test_photon_small_matches_oracle(torch_cuda):
    torch = torch_cuda
    Y = _Y()
    d = S.photon_field(n_side=4, subdiv=2, extent=10.0, resx=96, resy=64, photons=2 * 150 * 150,
                       search=20, nmat=4)
    sc = Y.Scene.from_desc(d)
    ps = sc.photon_build(0)
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    os_ = oc.photon_build(d, 0)
    assert ps.emitted == os_[0] == 150 * 150
    assert ps.stored == os_[1]
    assert ps.clustered == os_[2]
    n = ps.clustered
    buf = torch.zeros(n * 9, dtype=torch.float32, device="cuda")
    assert sc.photon_read(buf) == n
    g = buf.cpu().numpy().reshape(n, 9)
    c = oc.photon_read()
    assert np.mean(np.all(g == c, axis=1)) > 0.99
    assert np.allclose(g, c, rtol=1e-5, atol=1e-6)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    sc.render(out)
    torch.cuda.synchronize()
    img = out.cpu().numpy().reshape(d.cam.resy, d.cam.resx, 3)
    ref, _ = oc.render(d)
    assert rel_rmse(img, ref) <= 1e-3


@pytest.mark.parametrize("rays,compact", [(2, 1.0), (4, 1.0), (1, -1.0), (1, 0.5), (2, 0.75)])
def test_level_launch_settings_keep_the_image(torch_cuda, rays, compact):
    """Rays per lane of the GI level launches and the compaction threshold
    change which lane traces which path, never a result: the image and the
    ray counts equal the default launch's bit for bit (also under graph
    replay and in the adjoint's level replay)."""
    torch = torch_cuda
    d = small_terrain(nx=96, ny=48, resx=120, resy=80, samples=4, depth=4, nmat=8)
    ref, st0, _ = gpu_render(d, torch)
    Y = _Y()
    sc = Y.Scene.from_desc(d, rays_per_lane=rays, compact_below=compact)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    for _ in range(3):  # eager, captured, replayed
        out.zero_()
        st = sc.render(out)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy().reshape(ref.shape), ref)
        assert (st.rays_camera, st.rays_extension, st.rays_shadow) == (st0.rays_camera, st0.rays_extension,
                                                                       st0.rays_shadow)


@pytest.mark.parametrize("width", [2, 4, 8])
def test_smoothed_sphere_room_matches_oracle(torch_cuda, width):
    """The room's icosphere autoSmooth'ed at 180 degrees (every corner
    smoothed) and at 8 (corners split where adjacent faces meet at more than
    8 degrees, about half of them at subdivision 3): shading
    normals interpolated at primary hits, at path vertices (their tangent
    frames) and in the area light's and sun's N.L; GPU = oracle."""
    torch = torch_cuda
    d = S.direct_room(96, 72, subdiv=3, samples=4)
    lights = list(d.lights) + [D.pathlight(samples=4, depth=3, use_qmc=True, threshold=1e3),
                               D.sunlight((0.2, -0.3, 1.0), power=0.6)]
    for angle in (180.0, 8.0):
        ds = D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, d.cam, list(d.mats), [], [], lights,
                         d.opts, smooth={3: angle})
        img, st, _ = gpu_render(ds, torch, build_kw={"width": width})
        ref, rs = po.OracleScene.from_desc(ds).render(ds)
        assert np.all(np.isfinite(img))
        assert rel_rmse(img, ref) <= 1e-3
        assert (st.rays_camera, st.rays_extension, st.rays_shadow) == tuple(int(x) for x in rs[:3])
        assert st.lcg_draws == rs[4]
    # smoothing changes the image (the pins above are not vacuous)
    dflat = D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, d.cam, list(d.mats), [], [], lights,
                        d.opts)
    flat, _, _ = gpu_render(dflat, torch)
    assert rel_rmse(img, flat) > 1e-3


def test_smoothed_spheres_photon_map_matches_oracle(torch_cuda):
    """Photon bounces leave smoothed spheres around their interpolated
    normals' tangent frames and test fromRadiosity's side with them."""
    torch = torch_cuda
    Y = _Y()
    d = S.photon_field(n_side=6, subdiv=2, extent=20.0, resx=96, resy=72, photons=2 * 200 * 200, search=50, nmat=4)
    d.smooth = {o: 180.0 for o in range(len(d.obj_flags))}
    sc = Y.Scene.from_desc(d)
    ps = sc.photon_build(0)
    oc = po.OracleScene.from_desc(d)
    os_ = oc.photon_build(d, 0)
    assert (ps.emitted, ps.stored, ps.clustered) == tuple(int(x) for x in os_[:3])
    buf = torch.zeros(ps.clustered * 9, dtype=torch.float32, device="cuda")
    assert sc.photon_read(buf) == ps.clustered
    g, c = buf.cpu().numpy().reshape(-1, 9), oc.photon_read()
    assert np.mean(np.all(g == c, axis=1)) > 0.99
    assert np.allclose(g, c, rtol=1e-5, atol=1e-6)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    sc.render(out)
    torch.cuda.synchronize()
    ref, _ = oc.render(d)
    assert rel_rmse(out.cpu().numpy().reshape(ref.shape), ref) <= 1e-3


@pytest.mark.parametrize("builder", [D.YG_BUILDER_DEVICE, D.YG_BUILDER_DEVICE_PLOC])
@pytest.mark.parametrize("width", [2, 4, 8])
def test_device_built_tree_renders_like_the_oracle(torch_cuda, width, builder):
    """Trees built on the GPU (binned SAH, or locally-ordered clustering) in
    place of the host's: the GI terrain, the smoothed-sphere room with its
    area light and pathlight (per-corner normals scattered to the device
    tree's leaf order) and the photon field render as the oracle does, with
    the same ray counts."""
    torch = torch_cuda
    cases = [small_terrain(nx=128, ny=64, resx=120, resy=80, samples=4, depth=3, nmat=16)]
    r = S.direct_room(80, 60, subdiv=3, samples=4)
    cases.append(D.SceneDesc(r.verts, r.idx, r.tri_obj, r.tri_mat, r.obj_flags, r.cam, list(r.mats), [], [],
                             list(r.lights) + [D.pathlight(samples=4, depth=2, use_qmc=True, threshold=1e3)],
                             r.opts, smooth={3: 180.0}))
    for d in cases:
        img, st, sc = gpu_render(d, torch, build_kw={"width": width, "builder": builder})
        ref, rs = po.OracleScene.from_desc(d).render(d)
        assert rel_rmse(img, ref) <= 1e-3
        assert (st.rays_camera, st.rays_extension, st.rays_shadow) == tuple(int(x) for x in rs[:3])
        assert sc.info()["nodes"] > 0
    d = S.photon_field(n_side=4, subdiv=2, extent=10.0, resx=96, resy=64, photons=2 * 150 * 150, search=20, nmat=4)
    sc = _Y().Scene.from_desc(d, width=width, builder=builder)
    ps = sc.photon_build(0)
    oc = po.OracleScene.from_desc(d)
    os_ = oc.photon_build(d, 0)
    assert (ps.emitted, ps.stored, ps.clustered) == tuple(int(x) for x in os_[:3])
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    sc.render(out)
    torch.cuda.synchronize()
    ref, _ = oc.render(d)
    assert rel_rmse(out.cpu().numpy().reshape(ref.shape), ref) <= 1e-3


@pytest.mark.parametrize("search", [50, 100])
def test_photon_k50_merged_clusters_match_oracle(torch_cuda, search):
    """The photon workload's settings at small scale: K = 50 (and 100, past
    the old 64-entry heap), depth 1 (one diffuse bounce), mindepth 0 and the
    default cluster size fixedradius / 4, so that stored photons merge into
    marks (preGathering): map and image equal the oracle's."""
    torch = torch_cuda
    Y = _Y()
    d = S.photon_field(n_side=8, subdiv=2, extent=20.0, resx=160, resy=120, photons=2 * 300 * 300,
                       search=search, nmat=4)
    assert abs(d.lights[0].cluster - d.lights[0].fixedradius / 4) < 1e-7 and d.lights[0].depth == 1
    sc = Y.Scene.from_desc(d)
    ps = sc.photon_build(0)
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    os_ = oc.photon_build(d, 0)
    assert (ps.emitted, ps.stored, ps.clustered) == tuple(int(x) for x in os_[:3])
    assert ps.clustered < 0.6 * ps.stored        # marks really merge
    buf = torch.zeros(ps.clustered * 9, dtype=torch.float32, device="cuda")
    sc.photon_read(buf)
    g = buf.cpu().numpy().reshape(-1, 9)
    assert np.allclose(g, oc.photon_read(), rtol=1e-5, atol=1e-5)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    sc.render(out)
    torch.cuda.synchronize()
    img = out.cpu().numpy().reshape(d.cam.resy, d.cam.resx, 3)
    ref, _ = oc.render(d)
    assert np.all(np.isfinite(img))
    assert rel_rmse(img, ref) <= 1e-3
    # the oracle's per-pixel entry (camera ray + photon estimate) gives the same pixels
    px = (np.arange(0, d.cam.resx, 7) + 0.5).astype(np.float32)
    py = np.full_like(px, 60.5)
    assert np.allclose(oc.trace_camera(d, px, py), img[60, np.arange(0, d.cam.resx, 7)], rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("width", [2, 4, 8])
def test_batched_equals_single_renders(torch_cuda, width):
    """yg_render_batch over scenes of uneven size and resolution: every image
    equals that scene's own yg_render bit for bit, and one of them the oracle."""
    torch = torch_cuda
    Y = _Y()
    specs = [(40, 20, 64, 48), (160, 80, 96, 64), (24, 12, 33, 17), (96, 48, 128, 72), (60, 30, 80, 80)]
    scenes, outs, singles, descs, counts = [], [], [], [], []
    for k, (nx, ny, rx, ry) in enumerate(specs):
        d = S.terrain_gi(po.heightfield, nx=nx, ny=ny, resx=rx, resy=ry, samples=2, depth=3, seed_key=300 + k)
        sc = Y.Scene.from_desc(d, width=width)
        out = torch.zeros(rx * ry * 3, dtype=torch.float32, device="cuda")
        s1 = sc.render(out)
        counts.append((s1.rays_camera, s1.rays_extension, s1.rays_shadow))
        singles.append(out.clone())
        outs.append(torch.full_like(out, -1.0))
        scenes.append(sc)
        descs.append(d)
    st = Y.render_batch(scenes, outs)
    torch.cuda.synchronize()
    for a, b in zip(outs, singles):
        assert torch.equal(a, b)
    assert (st.rays_camera, st.rays_extension, st.rays_shadow) == tuple(sum(c[i] for c in counts) for i in range(3))
    d = descs[1]
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    ref, _ = oc.render(d, threads=8)
    img = outs[1].cpu().numpy().reshape(ref.shape)
    assert rel_rmse(img, ref) <= 1e-3


def test_frame_graph_replays_equal_eager(torch_cuda):
    """Frames replayed from the captured hipGraph equal the frame launched
    kernel by kernel, on every replay, with other device work and a render
    of another launch shape in between."""
    torch = torch_cuda
    Y = _Y()
    d = small_terrain(nx=256, ny=128, resx=320, resy=180, samples=2)
    sc = Y.Scene.from_desc(d)
    n = d.cam.resx * d.cam.resy * 3
    eager = D.yg_render_opts.from_buffer_copy(d.opts)
    eager.flags |= 8  # YG_RENDER_NO_GRAPH
    ref = torch.zeros(n, dtype=torch.float32, device="cuda")
    st0 = sc.render(ref, eager)
    torch.cuda.synchronize()
    assert st0.graph == 0
    ref_h = ref.cpu()
    small = small_terrain(nx=64, ny=32, resx=64, resy=48, samples=2)
    sc2 = Y.Scene.from_desc(small)
    out2 = torch.zeros(small.cam.resx * small.cam.resy * 3, dtype=torch.float32, device="cuda")
    modes = []
    for k in range(5):
        if k == 3:  # another launch shape between two replays
            sc2.render(out2)
        out = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
        st = sc.render(out)
        torch.cuda.synchronize()
        modes.append(int(st.graph))
        assert torch.equal(out.cpu(), ref_h), "frame %d (graph mode %d) differs" % (k, st.graph)
        for f in ("rays_camera", "rays_extension", "rays_shadow", "lcg_draws"):
            assert getattr(st, f) == getattr(st0, f), f
        x = torch.randn(1 << 20, device="cuda")  # eager work between replays
        assert torch.isfinite(x.sum()).item()
    # captured on the second consecutive frame of the shape (the eager one
    # was the first), replayed from then on
    assert modes[0] == 2 and modes[1] == 1 and modes[2] == 1, modes
    assert all(m in (0, 1, 2) for m in modes[3:]), modes


def test_autotune_keeps_fastest_build(torch_cuda):
    """yg_autotune times every (width 2/4/8, leaf) build with every launch
    setting (rays per lane, compaction threshold) on the frame, leaves the
    fastest in place, and the image does not depend on the choice."""
    torch = torch_cuda
    Y = _Y()
    d = small_terrain(nx=256, ny=128, resx=160, resy=90, samples=2)
    sc = Y.Scene.from_desc(d)
    n = d.cam.resx * d.cam.resy * 3
    ref = torch.zeros(n, dtype=torch.float32, device="cuda")
    sc.render(ref)
    best, best_ms, cands = sc.autotune(ref, widths=(2, 4, 8), leaves=(2, 4), repeats=2, rays=(1, 2),
                                       compacts=(1.0, -1))
    assert len(cands) == 3 * 2 * 2 * 2
    assert sorted(set((c[0], c[1]) for c in cands)) == [(2, 2), (2, 4), (4, 2), (4, 4), (8, 2), (8, 4)]
    assert sorted(set((c[6], c[7]) for c in cands)) == [(1, -1.0), (1, 1.0), (2, -1.0), (2, 1.0)]
    assert all(c[3] > 0 and c[4] > 0 for c in cands)
    assert best_ms == min(c[3] for c in cands)
    assert (best.width, best.max_leaf) == next((c[0], c[1]) for c in cands if c[3] == best_ms)
    # the scene holds the chosen build: same tree as building it afresh
    other = Y.Scene.from_desc(d, max_leaf=best.max_leaf, bins=best.sah_bins, width=best.width)
    assert sc.info()["nodes"] == other.info()["nodes"]
    out = torch.zeros(n, dtype=torch.float32, device="cuda")
    sc.render(out)
    torch.cuda.synchronize()
    a, b = out.cpu().numpy(), ref.cpu().numpy()
    assert np.mean(a == b) >= 0.9999
    ref_img, _ = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat).render(d, threads=8)
    assert rel_rmse(a.reshape(ref_img.shape), ref_img) <= 1e-3


def test_stack_lds_split_changes_nothing_but_speed(torch_cuda):
    """However many stack entries per lane stay in LDS (the rest in device
    memory), traversal and image are the same; the tuner's choice is kept."""
    torch = torch_cuda
    Y = _Y()
    d = small_terrain(nx=256, ny=128, resx=160, resy=90, samples=2)
    n = d.cam.resx * d.cam.resy * 3
    imgs, stats = [], []
    for sl in (-1, 4, 8, 16):
        sc = Y.Scene.from_desc(d, stack_lds=sl)
        out = torch.zeros(n, dtype=torch.float32, device="cuda")
        st = sc.render(out)
        torch.cuda.synchronize()
        imgs.append(out.cpu().numpy())
        stats.append((st.rays_camera, st.rays_extension, st.rays_shadow))
    for im, s in zip(imgs[1:], stats[1:]):
        assert np.array_equal(im, imgs[0]) and s == stats[0]
    sc = Y.Scene.from_desc(d)
    out = torch.zeros(n, dtype=torch.float32, device="cuda")
    best, best_ms, cands = sc.autotune(out, widths=(4,), leaves=(4,), stacks=(-1, 8, 16), repeats=1)
    assert [c[5] for c in cands] == [-1, 8, 16]
    assert best_ms == min(c[3] for c in cands) and best.stack_lds in (-1, 8, 16)
    sc.render(out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), imgs[0])


def test_roulette_mode_keeps_the_mean(torch_cuda):
    """YG_RENDER_ROULETTE (not a reference feature) ends paths at random from
    the third bounce on and divides survivors by their survival probability:
    fewer extension rays, the same expected image, reproducible frames; the
    adjoint refuses it."""
    torch = torch_cuda
    Y = _Y()
    d = small_terrain(samples=64, depth=4)
    sc = Y.Scene.from_desc(d)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    st0 = sc.render(out)
    torch.cuda.synchronize()
    ref = out.cpu().numpy().astype(np.float64)
    o = D.yg_render_opts.from_buffer_copy(d.opts)
    o.flags |= D.YG_RENDER_ROULETTE
    st1 = sc.render(out, o)
    torch.cuda.synchronize()
    rr = out.cpu().numpy().astype(np.float64)
    st2 = sc.render(out, o)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), rr.astype(np.float32))
    assert st2.rays_extension == st1.rays_extension
    assert st1.rays_extension < st0.rays_extension
    assert abs(rr.mean() / ref.mean() - 1) < 0.01
    assert rel_rmse(rr, ref) < 0.05
    g = torch.zeros_like(out)
    gm = torch.zeros(len(d.mats) * 3, dtype=torch.float64, device="cuda")
    gl = torch.zeros(len(d.lights), dtype=torch.float64, device="cuda")
    with pytest.raises(Exception, match="roulette"):
        sc.render_adjoint(g, out, gm, None, gl, opts=o)


@pytest.mark.parametrize("cast", [False, True])
def test_objects_that_cast_no_shadows(torch_cuda, cast):
    """A quad above the camera that casts no shadows: isShadowed and the
    pathlight's shadow-flag firstHit skip it (GPU = oracle, same ray counts),
    while the raw any-hit query (kdTree_t::IntersectS semantics) still hits it."""
    import test_oracle_render as TR
    torch = torch_cuda
    d, _ = TR._occluded_plane(cast)
    img, st, sc = gpu_render(d, torch)
    ref, rs = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat).render(d)
    assert rel_rmse(img, ref) <= 1e-3
    assert st.rays_extension == rs[1] and st.rays_shadow == rs[2]
    p = torch.tensor([[0.3, 0.2, 0.0]], dtype=torch.float32, device="cuda")
    up = torch.tensor([[0.0, 0.0, 1.0]], dtype=torch.float32, device="cuda")
    so = torch.tensor([0], dtype=torch.int32, device="cuda")
    out = torch.empty(1, dtype=torch.uint8, device="cuda")
    sc.is_shadowed(p, up, so, out, 0, self_bias=1e-3)
    assert int(out.item()) == int(cast)
    ray = torch.tensor([[0.3, 0.2, 0.01, 0.0, 0.0, 0.0, 1.0, float("inf")]], dtype=torch.float32, device="cuda")
    occ = torch.empty(1, dtype=torch.uint8, device="cuda")
    sc.trace_any(ray, occ)
    assert int(occ.item()) == 1


@pytest.mark.parametrize("path_first", [False, True])
def test_area_light_inside_gi_paths(torch_cuda, path_first):
    """A pathlight whose path vertices light an area light (use_in_indirect,
    pathlight.cc:186-191): the vertices' ourRandom() draws sit in the stream
    between the primary hit's lights in light order, sample by sample and
    bounce by bounce (scene.cc:260-277, pathlight.cc:305-347).  GPU = oracle,
    same shadow rays and draws, with the pathlight before or after the area
    light."""
    torch = torch_cuda
    d = S.direct_room(64, 48, subdiv=2)
    path = D.pathlight(samples=4, depth=2, use_qmc=True, threshold=1e3)
    lights = [path] + list(d.lights) if path_first else list(d.lights) + [path]
    d = D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, d.cam, list(d.mats), [], [], lights,
                    d.opts)
    img, st, _ = gpu_render(d, torch)
    ref, rs = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat).render(d)
    assert np.all(np.isfinite(img))
    assert rel_rmse(img, ref) <= 1e-3
    assert st.rays_camera == rs[0] and st.rays_extension == rs[1] and st.rays_shadow == rs[2]
    assert st.lcg_draws == rs[4]


def test_two_area_lights_around_the_pathlight(torch_cuda):
    """Area lights on both sides of the pathlight in light order (area,
    pathlight, area, sun): the split of each primary hit's draws around the
    pathlight's path draws (k_prep's `pre`, k_direct's jump) and the vertex
    pass's order.  GPU = oracle in image, rays and draws."""
    from test_oracle_fp64 import room_two_area_lights
    d = room_two_area_lights(res=(64, 48))
    img, st, _ = gpu_render(d, torch_cuda)
    ref, rs = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat).render(d)
    assert np.all(np.isfinite(img))
    assert rel_rmse(img, ref) <= 1e-3
    assert (st.rays_camera, st.rays_extension, st.rays_shadow) == tuple(int(x) for x in rs[:3])
    assert st.lcg_draws == rs[4]


def test_area_light_draws_beyond_32_bits_are_refused(torch_cuda):
    """The primary hits' ourRandom() stream offsets are 32-bit scans: a scene
    whose draws in one pass reach 2^32 (4096 pixels x 2 x 2^20 samples) is
    refused before any launch instead of drawing from wrong stream positions."""
    d = S.direct_room(64, 64, subdiv=1, samples=1 << 20)
    with pytest.raises(RuntimeError, match="too many area-light draws"):
        gpu_render(d, torch_cuda)


def test_area_light_inside_gi_paths_refuses_antialiasing(torch_cuda):
    """Antialiasing passes continue each area's ourRandom() stream pass by
    pass; with area lights at path vertices the launch is refused, not
    rendered differently from the reference."""
    d = S.direct_room(32, 24, subdiv=1, samples=2)
    o = D.render_opts(bias=1e-3, raydepth=3, aa_passes=1, aa_minsamples=2)
    d = D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, d.cam, list(d.mats), [], [],
                    list(d.lights) + [D.pathlight(samples=2, depth=2, use_qmc=True, threshold=1e3)], o)
    with pytest.raises(RuntimeError, match="antialiasing with area lights inside GI paths"):
        gpu_render(d, torch_cuda)


def test_batched_area_light_inside_gi_paths(torch_cuda):
    """A batch mixing rooms whose GI vertices draw from the area light (both
    light orders, two resolutions) with a terrain without area lights: each
    image equals its own yg_render bit for bit, and the draw counts add up."""
    torch = torch_cuda
    Y = _Y()
    descs = []
    for k, (rx, ry, path_first) in enumerate([(48, 36, False), (40, 30, True)]):
        d = S.direct_room(rx, ry, subdiv=2, samples=3 + k)
        path = D.pathlight(samples=3, depth=2, use_qmc=True, threshold=1e3)
        lights = [path] + list(d.lights) if path_first else list(d.lights) + [path]
        descs.append(D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, d.cam, list(d.mats), [], [],
                                 lights, d.opts))
    descs.insert(1, S.terrain_gi(po.heightfield, nx=40, ny=20, resx=32, resy=24, samples=3, depth=2))
    scenes, outs, singles, draws = [], [], [], 0
    for d in descs:
        sc = Y.Scene.from_desc(d)
        out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
        draws += sc.render(out).lcg_draws
        singles.append(out.clone())
        outs.append(torch.full_like(out, -1.0))
        scenes.append(sc)
    st = Y.render_batch(scenes, outs)
    torch.cuda.synchronize()
    for a, b in zip(outs, singles):
        assert torch.equal(a, b)
    assert st.lcg_draws == draws
