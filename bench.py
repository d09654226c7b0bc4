"""Benchmark: path-traced GI on the 1,045,506-triangle fBm terrain.

One step = one full yg_render of the terrain at 1280x720, 4 samples per
pixel, 3 bounces (pathlight use_QMC on, sun + constant sky), with the scene,
BVH and buffers already resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def gpu_leg(args):
    import numpy as np
    import torch

    from yafray_amd import descriptors as D
    from yafray_amd import lib as Y
    from yafray_amd import scenes as S

    torch.cuda.set_device(0)
    t0 = time.time()
    d = S.terrain_gi(Y.device_heights)
    sc = Y.Scene.from_desc(d)
    setup_s = time.time() - t0
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    # the autotuner picks BVH width, rays per lane and compaction threshold on
    # this frame (before the timed region) and leaves the scene built with them
    t0 = time.time()
    widths = (2, 4, 8) if not args.width else (args.width,)
    rays_c, comp_c = ((1, 2, 4), (1.0, 0.5, -1)) if not args.width else ((1,), (1.0,))
    best, best_ms, cands = sc.autotune(out, widths=widths, leaves=(4,), repeats=2, rays=rays_c, compacts=comp_c)
    tune = {"candidates": len(cands), "seconds": round(time.time() - t0, 1), "best_ms": round(best_ms, 3),
            "best": {"width": int(best.width), "max_leaf": int(best.max_leaf), "sah_bins": int(best.sah_bins),
                     "stack_lds": int(best.stack_lds) or 16, "rays_per_lane": int(best.rays_per_lane),
                     "compact_below": float(best.compact_below)},
            "slowest_ms": round(max(c[3] for c in cands), 3)}
    info = sc.info()
    opts = D.yg_render_opts.from_buffer_copy(d.opts)
    for _ in range(args.warmup):
        sc.render(out, opts)
    torch.cuda.synchronize()
    rays = 0
    modes = set()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        st = sc.render(out, opts)
        rays += st.rays_camera + st.rays_extension + st.rays_shadow
        modes.add(int(st.graph))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t_start
    # kernel-level timing of the GI bounce launches (HIP events on the render stream)
    kt = D.yg_render_opts.from_buffer_copy(d.opts)
    kt.flags = 3
    ks = sc.render(out, kt)
    torch.cuda.synchronize()
    img = out.cpu().numpy().reshape(d.cam.resy, d.cam.resx, 3)
    builders = build_compare(sc, out, torch, dict(width=int(best.width), max_leaf=int(best.max_leaf),
                                                  bins=int(best.sah_bins), stack_lds=int(best.stack_lds),
                                                  rays_per_lane=int(best.rays_per_lane),
                                                  compact_below=float(best.compact_below)))
    return {
        "desc": d, "img": img, "tune": tune, "builders": builders,
        "dt": dt, "rays": rays, "stats": ks, "info": info, "setup_s": setup_s,
        "graph_modes": sorted(modes),
        "finite": bool(np.all(np.isfinite(img))), "mean": float(img.mean()),
    }


def parity_leg(d, img, cpu_threads):
    """Bounded CPU sample: the same full 1280x720 frame rendered by the oracle
    on the host cores (timed), compared pixel by pixel with the GPU image.
    The timed CPU baseline is the restatement built with the reference's
    release flags (-O3 -ffast-math, liboracle_fast.so); the image compared
    with the GPU's comes from the strict build (also timed, reported beside)."""
    import numpy as np

    from oracle import pyoracle as po

    of = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat, fast=True)
    t0 = time.perf_counter()
    _, rf = of.render(d, threads=cpu_threads)
    fast_dt = time.perf_counter() - t0
    fast_rays = int(rf[0] + rf[1] + rf[2])
    of.close()
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    t0 = time.perf_counter()
    ref, rs = oc.render(d, threads=cpu_threads)
    strict_dt = time.perf_counter() - t0
    strict_rays = int(rs[0] + rs[1] + rs[2])
    cpu_dt, cpu_rays = fast_dt, fast_rays
    img = img.reshape(ref.shape).astype(np.float64)
    ref64 = ref.astype(np.float64)
    rel = float(np.sqrt(np.mean((img - ref64) ** 2)) / np.sqrt(np.mean(ref64 ** 2)))
    exact = float(np.mean(np.all(img == ref64, axis=2)))
    return {"cpu_dt": cpu_dt, "cpu_rays": cpu_rays, "strict_mrays": strict_rays / strict_dt / 1e6,
            "rel_rmse": rel, "exact_px": exact,
            "oracle": oc, "ref": ref}


FP64_SAMPLES = 4096


def path_spp(d):
    """Pathlight samples per pixel of a scene description (1 without one)."""
    from yafray_amd import descriptors as D
    return next((int(L.samples) for L in d.lights if L.type == D.YG_LIGHT_PATH), 1)


def cpu_cores():
    return min(os.cpu_count() or 1, 16)


def fp64_leg(d, images, n_pix, oc=None, seed=7):
    """Error against the converged FP64 render.  A bounded sample: `n_pix`
    pixels drawn with a fixed seed, each rendered through its centre with 4096
    samples of every sampled light (pathlight paths, area-light points) in
    double precision by the oracle (yo_render_fp64); each image in `images`
    ([resy, resx, 3] or flat) is compared with it on those pixels."""
    import numpy as np

    from oracle import pyoracle as po

    if oc is None:
        oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    rx, ry = d.cam.resx, d.cam.resy
    q = np.random.default_rng(seed).choice(rx * ry, size=min(n_pix, rx * ry), replace=False)
    px, py = (q % rx).astype(np.int32), (q // rx).astype(np.int32)
    threads = cpu_cores()
    t0 = time.perf_counter()
    conv = oc.render_fp64(d, samples=FP64_SAMPLES, px=px, py=py, seed=seed, threads=threads)
    dt = time.perf_counter() - t0
    out = {"samples": FP64_SAMPLES, "pixels": int(len(q)), "cpu_s": round(dt, 1), "threads": threads}
    den = np.sqrt(np.mean(conv ** 2))
    for name, img in images.items():
        v = np.asarray(img, np.float64).reshape(ry, rx, 3)[py, px]
        out["rel_rmse_" + name] = float(np.sqrt(np.mean((v - conv) ** 2)) / den)
    return out


def copy_probe():
    """A plain 1 GiB device-to-device copy (torch's copy kernel), run after the
    timed region so that a counter run of this script can scale the bounce
    kernel's FETCH_SIZE / WRITE_SIZE by what the counters read for a known
    number of bytes."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 2}


TRAFFIC_PROFILE = "profiles/gi_terrain_1280x720_hbm_traffic.md"


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE / WRITE_SIZE passes scaled by a 1 GiB copy), or None."""
    try:
        txt = open(os.path.join(ROOT, TRAFFIC_PROFILE)).read()
        block = txt.split("```json", 1)[1].split("```", 1)[0]
        return int(json.loads(block)["kernels"][kernel]["hbm_bytes_per_launch"])
    except (OSError, IndexError, KeyError, ValueError):
        return None


def build_compare(sc, out, torch, kw, reps=2, frames=5):
    """The scene's BVH built by the host builder (binned SAH on host threads)
    and by the device builders (the same binned SAH on the GPU; clustering
    on the GPU) at the same settings:
    wall time of yg_build (fastest of `reps`, host data already in memory)
    and the frame time each tree gives.  Leaves the scene built by the
    host builder."""
    from yafray_amd import descriptors as D
    res = {}
    for name, b in (("device_ploc", D.YG_BUILDER_DEVICE_PLOC), ("device_sah", D.YG_BUILDER_DEVICE),
                    ("host", D.YG_BUILDER_HOST)):
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sc.build(builder=b, **kw)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        ms, _ = _timed(lambda: sc.render(out), frames, 1)
        info = sc.info()
        res[name] = {"build_s": round(min(ts), 4), "frame_ms": round(ms * 1e3, 3), "nodes": int(info["nodes"]),
                     "max_depth": int(info["max_depth"])}
    return res


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def bench_batched(args):
    """64 heightfield scenes (1e3..2e6 triangles, four resolutions), 2 spp,
    3 bounces: one batched launch vs 64 per-scene renders on the same GPU."""
    import torch

    from yafray_amd import lib as Y
    from yafray_amd import scenes as S

    specs = S.batch_specs(args.batch_n)
    scenes, outs, tris, rays_px = [], [], 0, 0
    for k, (nx, ny, rx, ry) in enumerate(specs):
        d = S.terrain_gi(Y.device_heights, nx=nx, ny=ny, resx=rx, resy=ry, samples=2, depth=3,
                         seed_key=100 + k)
        scenes.append(Y.Scene.from_desc(d))
        outs.append(torch.zeros(rx * ry * 3, dtype=torch.float32, device="cuda"))
        tris += d.n_tris
        rays_px += rx * ry
    mode = args.batch_mode
    if mode == "sequential":
        # per-scene renders only (a kernel profile of that mode alone)
        seq_dt, _ = _timed(lambda: [s.render(o) for s, o in zip(scenes, outs)], args.steps, args.warmup)
        return {"metric": "sequential GI ms/iter", "value": round(seq_dt * 1e3, 3), "unit": "ms/iter",
                "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(seq_dt * 1e3, 3),
                "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
                "data": "synthetic", "config": {"workload": "%d heightfield scenes rendered one after another"
                                                            % len(specs), "triangles": int(tris)}}
    seq_dt = None
    if mode == "both":
        seq_dt, _ = _timed(lambda: [s.render(o) for s, o in zip(scenes, outs)], args.steps, args.warmup)
    ref = [o.clone() for o in outs]
    bat_dt, st = _timed(lambda: Y.render_batch(scenes, outs), args.steps, args.warmup)
    same = all(torch.equal(a, b) for a, b in zip(ref, outs)) if seq_dt is not None else None
    rays = st.rays_camera + st.rays_extension + st.rays_shadow
    fp64 = None
    if not args.no_cpu:
        # the first scene of each resolution, 256 pixels each
        fp64, seen = [], set()
        for s, o in zip(scenes, outs):
            key = (s.desc.cam.resx, s.desc.cam.resy)
            if key in seen:
                continue
            seen.add(key)
            e = fp64_leg(s.desc, {"gpu": o.cpu().numpy()}, 256)
            e["scene"] = "%dx%d, %d tris" % (key[0], key[1], s.desc.n_tris)
            fp64.append(e)
    return {
        "metric": "batched GI ms/iter", "value": round(bat_dt * 1e3, 3), "unit": "ms/iter",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(bat_dt * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "%d heightfield scenes, triangles log-uniform in [1e3, 2e6], "
                               "resolutions from 256^2..1920x1080, 2 spp, 3 bounces" % len(specs),
                   "triangles": int(tris), "pixels": int(rays_px)},
        "sequential_ms_per_iter": None if seq_dt is None else round(seq_dt * 1e3, 3),
        "speedup_vs_sequential": None if seq_dt is None else round(seq_dt / bat_dt, 3),
        "batched_mrays_per_s": round(rays / bat_dt / 1e6, 2),
        "batched_msamples_per_s": round(rays_px * path_spp(scenes[0].desc) / bat_dt / 1e6, 2),
        "images_identical_to_sequential": same,
        "error_vs_fp64": fp64,
    }


def bench_adjoint(args):
    """Differentiable GI: terrain at 512x512, 16 spp, 3 bounces, 256 materials,
    Musgrave-modulated albedo; forward vs forward+adjoint (replay) time."""
    import torch

    from yafray_amd import lib as Y
    from yafray_amd import scenes as S
    from yafray_amd.autograd import render_gi

    d = S.terrain_diff(Y.device_heights)
    sc = Y.Scene.from_desc(d)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    fwd_dt, _ = _timed(lambda: sc.render(out), args.steps, args.warmup)
    import numpy as np
    alb0 = np.array([list(m.color) for m in d.mats])
    target = render_gi(sc, torch.tensor(alb0 * 0.9 + 0.04), torch.tensor(0.4), torch.tensor(2.7),
                       torch.tensor(2.5)).detach().clone()
    albedo = torch.tensor(alb0, requires_grad=True)
    amp = torch.tensor(0.5, requires_grad=True)
    size = torch.tensor(3.0, requires_grad=True)
    power = torch.tensor(3.0, requires_grad=True)

    def step():
        for t in (albedo, amp, size, power):
            t.grad = None
        im = render_gi(sc, albedo, amp, size, power)
        L = ((im - target) ** 2).sum()
        L.backward()
        return float(L)

    both_dt, loss = _timed(step, args.steps, args.warmup)
    # device time of the adjoint replay alone (REUSE of the forward just made)
    from yafray_amd import descriptors as D
    o = D.yg_render_opts.from_buffer_copy(sc.desc.opts)
    sc.render(out, o)
    o.flags |= 1 | 4
    gt = torch.ones(out.numel(), dtype=torch.float32, device="cuda")
    gm = torch.zeros(len(sc.desc.mats) * 3, dtype=torch.float64, device="cuda")
    go = torch.zeros(max(len(sc.desc.mods), 1) * 2, dtype=torch.float64, device="cuda")
    gl = torch.zeros(len(sc.desc.lights), dtype=torch.float64, device="cuda")
    adj_ms = []
    for _ in range(3):
        st = sc.render_adjoint(gt, out, gm, go, gl, opts=o)
        adj_ms.append(st.ms_trace)
    fst = sc.render(out, timing=True)
    fp64 = grad = None
    if not args.no_cpu:
        # forward image at the starting parameters against the converged FP64 render
        sc.render(out)
        torch.cuda.synchronize()
        fp64 = fp64_leg(d, {"gpu": out.cpu().numpy()}, 512)
        grad = adjoint_gradient_leg(d, sc, torch)
    return {
        "metric": "fwd+adjoint ms/iter", "value": round(both_dt * 1e3, 3), "unit": "ms/iter",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(both_dt * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "differentiable GI: 1,045,506-tri terrain, 512x512, 16 spp, 3 bounces, "
                               "256 materials, Musgrave albedo texture (mul)",
                   "params": 256 * 3 + 3},
        "forward_ms": round(fwd_dt * 1e3, 3),
        "forward_msamples_per_s": round(d.cam.resx * d.cam.resy * path_spp(d) / fwd_dt / 1e6, 2),
        "ratio_fwd_adj_over_fwd": round(both_dt / fwd_dt, 3),
        "adjoint_replay_device_ms": round(min(adj_ms), 3),
        "forward_device_ms": round(fst.ms, 3),
        "loss": loss,
        "error_vs_fp64": fp64,
        "gradient_vs_oracle": grad,
    }


def adjoint_gradient_leg(d, sc, torch, n_sample=64):
    """The GPU adjoint's gradient against the oracle's double-precision
    gradient of the same L2 loss (target rendered from other parameters), at
    the workload's full size: relative L2 error of the whole gradient
    (768 albedo components, modulator amplitude and frequency, sun power)
    and the largest relative error over `n_sample` sampled albedo components
    among those at least 5% of the largest."""
    import numpy as np

    from oracle import pyoracle as po
    from yafray_amd import scenes as S

    def render(desc):
        sc.set_desc(desc)
        o = torch.zeros(desc.cam.resx * desc.cam.resy * 3, dtype=torch.float32, device="cuda")
        sc.render(o)
        torch.cuda.synchronize()
        return o.cpu().numpy().astype(np.float64)

    alb = np.array([list(m.color) for m in d.mats], np.float64)
    target = render(S.with_params(d, albedo=alb * 0.9 + 0.04, amp=0.4, size=2.7, sun_power=2.5))
    g = (2 * (render(d) - target)).astype(np.float32)
    gt = torch.from_numpy(g).cuda()
    out = torch.zeros_like(gt)
    gm = torch.zeros(len(d.mats) * 3, dtype=torch.float64, device="cuda")
    go = torch.zeros(2, dtype=torch.float64, device="cuda")
    gl = torch.zeros(len(d.lights), dtype=torch.float64, device="cuda")
    sc.render_adjoint(gt, out, gm, go, gl)
    torch.cuda.synchronize()
    gpu = np.concatenate([gm.cpu().numpy(), go.cpu().numpy(), gl.cpu().numpy()])
    threads = cpu_cores()
    t0 = time.perf_counter()
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    _, rm, ro, rl = oc.render_adjoint(d, g, threads=threads)
    dt = time.perf_counter() - t0
    ref = np.concatenate([rm.reshape(-1), ro.reshape(-1), rl])
    nm = 3 * len(d.mats)
    big = np.flatnonzero(np.abs(ref[:nm]) >= 0.05 * np.abs(ref[:nm]).max())
    pick = np.random.default_rng(7).choice(big, size=min(n_sample, len(big)), replace=False)
    rel = np.abs(gpu[pick] - ref[pick]) / np.abs(ref[pick])
    return {"rel_l2_error": float(np.linalg.norm(gpu - ref) / np.linalg.norm(ref)),
            "sampled_albedo_max_rel_error": float(rel.max()), "sampled": int(len(pick)),
            "modulator_amp_rel_error": float(abs(gpu[nm] - ref[nm]) / abs(ref[nm])),
            "modulator_size_rel_error": float(abs(gpu[nm + 1] - ref[nm + 1]) / abs(ref[nm + 1])),
            "sun_power_rel_error": float(abs(gpu[-1] - ref[-1]) / abs(ref[-1])),
            "oracle": "oracle/yo_render.c yo_render_adjoint (double-precision sums), %d threads, %.1f s"
                      % (threads, dt)}


def bench_photon(args):
    """Photon mapping: 2M photons from a 45-degree spot into 4,096 icospheres
    (20,971,522 triangles), K = 50 gather at 1920x1080.  One step = photon
    build (emission, tracing, clustering, gather index) + render.

    Also: the K-nearest gather kernel alone (yg_photon_gather over the
    frame's primary hit points, HIP events on its stream) against the HBM
    roofline; a sampled-pixel parity check of the GPU frame against the
    oracle (same photon map built on the host); and the oracle's photon build
    plus the sampled pixels as the CPU baseline, scaled to a full frame."""
    import numpy as np
    import torch

    from yafray_amd import descriptors as D
    from yafray_amd import lib as Y
    from yafray_amd import scenes as S

    t0 = time.time()
    d = S.photon_field()
    sc = Y.Scene.from_desc(d, builder=D.YG_BUILDER_DEVICE)
    setup_s = time.time() - t0
    pst = sc.photon_build(0)
    build_dt, _ = _timed(lambda: sc.photon_build(0), max(1, args.steps // 2), 0)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    rdt, st = _timed(lambda: sc.render(out), args.steps, args.warmup)
    img = out.cpu().numpy().reshape(d.cam.resy, d.cam.resx, 3)
    finite = bool(np.all(np.isfinite(img)))
    L = d.lights[0]
    K = int(L.search)
    # -- the gather kernel on the frame's primary hit points (pixel centres)
    rx, ry = d.cam.resx, d.cam.resy
    o, dirs = _camera_rays(d.cam, np.arange(rx * ry))
    frm = torch.from_numpy(o.astype(np.float32)).cuda()
    dd = torch.from_numpy(dirs.astype(np.float32)).cuda()
    hits = torch.empty((rx * ry, 4), dtype=torch.float32, device="cuda")
    sc.first_hit(frm, dd, hits)
    h = hits.cpu().numpy()
    prim = h[:, 0].view(np.int32)
    ok = prim >= 0
    P = (o + dirs * (D.MIN_RAYDIST + h[:, 1:2]))[ok].astype(np.float32)
    tri = d.verts[d.idx[prim[ok]]].astype(np.float64)
    ng = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    ng = (ng / np.linalg.norm(ng, axis=1, keepdims=True)).astype(np.float32)
    eye = (o[ok] - P).astype(np.float32)
    mat = d.tri_mat[prim[ok]].astype(np.int32)
    nq = int(ok.sum())
    tP, tN, tE = (torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (P, ng, eye))
    tM = torch.from_numpy(mat).cuda()
    rgb = torch.empty((nq, 3), dtype=torch.float32, device="cuda")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(2):
        sc.photon_gather(tP, tN, tN, tE, tM, rgb)
    e0.record()
    for _ in range(args.steps):
        sc.photon_gather(tP, tN, tN, tE, tM, rgb)
    e1.record()
    torch.cuda.synchronize()
    g_ms = e0.elapsed_time(e1) / args.steps
    # algorithmic bytes per query: 48 B query (P, N, Ng, eye) + 12 B estimate
    # + K x 36 B photons + 27 cells x 8 B (cell ranges of the 3x3x3 block)
    bq = 48 + 12 + K * 36 + 27 * 8
    achieved = nq * bq / (g_ms * 1e-3) / 1e9
    line = {
        "metric": "photon map ms/frame", "value": round((build_dt + rdt) * 1e3, 3), "unit": "ms/frame",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round((build_dt + rdt) * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "4096 icospheres (20,971,522 tris), 1,999,396 photons, K=50, 1920x1080",
                   "triangles": int(d.n_tris)},
        "photon_build_ms": round(build_dt * 1e3, 3), "render_ms": round(rdt * 1e3, 3),
        "emitted": int(pst.emitted), "stored": int(pst.stored), "clustered": int(pst.clustered),
        "scene_setup_s": round(setup_s, 1), "image_finite": finite,
        "bvh": "device binned-SAH builder, width 4",
        "roofline": {"bound": "hbm", "kernel": "k_gather_points", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "queries": nq, "ms_per_launch": round(g_ms, 4),
                     "alg_bytes_per_query": bq,
                     "note": "the clustered photon map (%d marks, %d KB) stays in L2: the gather is bound by "
                             "its LDS heap work, not by HBM" % (int(pst.clustered), int(pst.clustered) * 36 // 1024)},
    }
    # the two builders on this scene (the photon map is rebuilt for each tree)
    bvh = {}
    for name, b in (("host", D.YG_BUILDER_HOST), ("device_ploc", D.YG_BUILDER_DEVICE_PLOC),
                    ("device_sah", D.YG_BUILDER_DEVICE)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sc.build(builder=b)
        torch.cuda.synchronize()
        bs = time.perf_counter() - t0
        fdt, _ = _timed(lambda: (sc.photon_build(0), sc.render(out)), 2, 1)
        bvh[name] = {"build_s": round(bs, 3), "ms_per_frame": round(fdt * 1e3, 3), "nodes": int(sc.info()["nodes"])}
    line["bvh_builders"] = bvh
    if not args.no_cpu:
        line.update(_photon_cpu_legs(d, img, pst, args, sc=sc))
    return line


def _camera_rays(cam, idx):
    """Camera rays through pixel centres (camera.cc:27-87, scene.cc:736-756)
    in float64: (origins, unit directions) for flat pixel indices idx."""
    import numpy as np
    pos = np.array(cam.pos[:], np.float64)
    look = np.array(cam.look[:], np.float64)
    up = np.array(cam.up[:], np.float64)
    vup, vto = up - pos, look - pos
    vright = np.cross(vup, vto)
    vup = np.cross(vright, vto)
    vup /= np.linalg.norm(vup)
    vright /= -np.linalg.norm(vright)
    vto /= np.linalg.norm(vto)
    vup *= cam.aspect * cam.resy / cam.resx
    vto = vto * cam.focal - 0.5 * (vup + vright)
    vup /= cam.resy
    vright /= cam.resx
    i, j = idx // cam.resx, idx % cam.resx
    dv = vright[None] * (j + 0.5)[:, None] + vup[None] * (i + 0.5)[:, None] + vto[None]
    dv /= np.linalg.norm(dv, axis=1, keepdims=True)
    return np.repeat(pos[None], len(idx), 0), dv


def _photon_cpu_legs(d, img, pst, args, sc=None, n_pix=4096, seed=11):
    """Oracle legs of the photon workload, on the host.

    * map: the oracle builds the photon map itself (timed: with the sampled
      pixels below, scaled to the full frame, the CPU baseline); marks are
      compared cell by cell with the GPU's;
    * estimate: the oracle renders a fixed-seed sample of pixels with the
      GPU's map loaded (yo_photon_load), so the K-nearest estimate at full
      size is compared on equal photons;
    * camera rays of the sample whose first hit differs are classified by an
      FP64 edge test of the oracle's triangle.

    The reference draws bounce directions from one Halton sequence in photon
    order, so a single first hit decided differently (the reference's float
    edge test is tree-dependent at far silhouettes) shifts the bounce of
    every later photon: maps built independently differ in a few percent of
    their cells, and the K-nearest estimate passes that on to the image."""
    import numpy as np
    import torch

    from oracle import pyoracle as po

    print("photon: oracle tree over %d triangles" % d.n_tris, file=sys.stderr, flush=True)
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    print("photon: oracle photon build", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    os_ = oc.photon_build(d, 0)
    cpu_build = time.perf_counter() - t0
    omap = oc.photon_read()
    gbuf = torch.zeros(int(pst.clustered) * 9, dtype=torch.float32, device="cuda")
    sc.photon_read(gbuf)
    gmap = gbuf.cpu().numpy().reshape(-1, 9)
    cell = float(d.lights[0].cluster)
    gk = {tuple(np.floor(r[3:6] / cell).astype(np.int64)): r for r in gmap}
    ok_ = {tuple(np.floor(r[3:6] / cell).astype(np.int64)): r for r in omap}
    common = set(gk) & set(ok_)
    same = sum(1 for k in common if np.allclose(gk[k], ok_[k], rtol=1e-5, atol=1e-5))
    rx, ry = d.cam.resx, d.cam.resy
    q = np.random.default_rng(seed).choice(rx * ry, size=n_pix, replace=False)
    px, py = (q % rx).astype(np.float32) + 0.5, (q // rx).astype(np.float32) + 0.5
    print("photon: oracle sample of %d pixels" % n_pix, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    oc.trace_camera(d, px, py)
    cpu_px = time.perf_counter() - t0
    oc.photon_load(d, 0, gmap)
    ref = oc.trace_camera(d, px, py).astype(np.float64)
    g = img.reshape(ry, rx, 3)[q // rx, q % rx].astype(np.float64)
    # first hits of the sample's camera rays on both sides, FP64 check of mismatches
    o, dirs = _camera_rays(d.cam, q)
    o32, d32 = o.astype(np.float32), dirs.astype(np.float32)
    hits = torch.empty((n_pix, 4), dtype=torch.float32, device="cuda")
    sc.first_hit(torch.from_numpy(o32).cuda(), torch.from_numpy(d32).cuda(), hits)
    pg = hits.cpu().numpy()[:, 0].view(np.int32)
    pc, _ = oc.first_hit(o32, d32)
    mism = np.nonzero(pg != pc)[0]
    false_pos = 0
    for k in mism:
        p = pc[k] if pc[k] >= 0 else pg[k]
        A, B, C = d.verts[d.idx[p]].astype(np.float64)
        n = np.cross(B - A, C - A)
        f = o32[k].astype(np.float64) + d32[k].astype(np.float64) * 5e-5
        r = d32[k].astype(np.float64)
        r = -r if np.dot(r, n) < 0 else r
        va, vb, vc = A - f, B - f, C - f
        tp = [np.dot(np.cross(va, vb), r), np.dot(np.cross(vb, vc), r), np.dot(np.cross(vc, va), r)]
        false_pos += int(min(tp) < 0 < max(tp))
    agree = np.ones(n_pix, bool)
    agree[mism] = False

    def rrmse(x, y):
        return float(np.sqrt(np.mean((x - y) ** 2)) / max(np.sqrt(np.mean(y ** 2)), 1e-30))

    frame_ms = (cpu_build + cpu_px / n_pix * rx * ry) * 1e3
    return {
        "cpu_baseline": {"value": round(frame_ms, 1), "unit": "ms/frame", "cores": 1, "kind": "port",
                         "sample": "oracle photon build (%.1f s) + %d sampled pixels (%.2f s) scaled to the "
                                   "%dx%d frame, one thread" % (cpu_build, n_pix, cpu_px, rx, ry)},
        "rmse_vs_oracle": {
            "rel_rmse": rrmse(g, ref), "rel_rmse_same_first_hit": rrmse(g[agree], ref[agree]),
            "pixels": n_pix, "camera_first_hit_differs": int(len(mism)),
            "of_which_float_false_positive": int(false_pos),
            "sample": "fixed-seed pixel sample of the full frame; the oracle renders it with the GPU's photon "
                      "map; a differing first hit whose triangle the ray misses in FP64 is a false positive "
                      "of the reference's float edge test, which trees decide differently"},
        "photon_map_vs_oracle": {
            "emitted": [int(pst.emitted), int(os_[0])], "stored": [int(pst.stored), int(os_[1])],
            "clustered": [int(pst.clustered), int(os_[2])], "cells_equal": int(same),
            "cells_common": int(len(common))},
    }


def bench_gphoton(args):
    """Global photon map (globalPhotonLight_t) read at pathlight vertices:
    the room (icosphere 20,480 + 2 boxes + 5 walls) with its area light a
    dummy emitter of 1,000,000 photons, depth 2, search 100, radius 1, a sun,
    and a 16-sample one-bounce pathlight whose path vertices read the
    irradiance map, at 640x480.  One step = map build (tracing, box merging,
    pre-gathering) + the frame.

    Also: device time of the photon tracing and pre-gathering launches and
    of the frame's extension / shadow launches (HIP events on the library's
    stream); the frame against the oracle rendering it with the GPU's map
    (same arithmetic: equal within rounding) and against the oracle's own
    serial map (a different photon stream: statistical agreement); the
    oracle's serial map build and its frame as the CPU baseline."""
    import numpy as np
    import torch

    from yafray_amd import descriptors as D
    from yafray_amd import lib as Y
    from yafray_amd import scenes as S

    d = S.gphoton_room(640, 480, subdiv=5, photons=1000000, depth=2, search=100, radius=1.0, samples=16,
                       path_depth=1, sun=(0.3, -0.2, 1.0))
    sc = Y.Scene.from_desc(d)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")

    def step():
        sc.gphoton_build(1)
        return sc.render(out)

    dt, st = _timed(step, args.steps, args.warmup)
    kt = D.yg_render_opts.from_buffer_copy(d.opts)
    kt.flags = 3
    gst = sc.gphoton_build(1, kt)
    bdt, _ = _timed(lambda: sc.gphoton_build(1), args.steps, 1)
    rdt, _ = _timed(lambda: sc.render(out), args.steps, 1)
    rst = sc.render(out, kt)
    torch.cuda.synchronize()
    img = out.cpu().numpy().reshape(d.cam.resy, d.cam.resx, 3)
    finite = bool(np.all(np.isfinite(img)))
    if not finite:
        raise RuntimeError("gphoton: non-finite pixels in the GPU image")
    rays = int(st.rays_camera + st.rays_extension + st.rays_shadow)
    nbox = int(gst.clustered)
    K = int(d.lights[1].search)
    # pre-gathering: per box 24 B query (position, normal) + 12 B irradiance out
    # + K x 20 B (position, direction code, RGBE colour of each box photon used)
    # + 27 x 4 B cell starts
    irr_b = nbox * (24 + 12 + K * 20 + 27 * 4)
    kern = {"k_gp_trace": {"ms": float(gst.ms_trace), "segments": int(gst.rays)},
            "k_gp_irr": {"ms": float(gst.ms_irr), "boxes": nbox, "alg_bytes": irr_b},
            "k_ext": {"ms": float(rst.ms_ext), "rays": int(rst.rays_extension)},
            "k_shadow": {"ms": float(rst.ms_shadow), "rays": int(rst.rays_bounce - rst.rays_extension)}}
    achieved = irr_b / (gst.ms_irr * 1e-3) / 1e9 if gst.ms_irr > 0 else 0.0
    line = {
        "metric": "global photon map + GI ms/frame", "value": round(dt * 1e3, 3), "unit": "ms/frame",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "room + icosphere(20480) + 2 boxes, 640x480; global photon map of 1,000,000 "
                               "photons from the area light, depth 2, search 100, radius 1; sun; pathlight "
                               "16 spp, 1 bounce, reading the map at its vertices",
                   "triangles": int(d.n_tris)},
        "map_build_ms": round(bdt * 1e3, 3), "render_ms": round(rdt * 1e3, 3),
        "rays_per_frame": rays, "mrays_per_s_frame": round(rays / rdt / 1e6, 3),
        "emitted": int(gst.emitted), "stored": int(gst.stored), "boxes": nbox,
        "kernels_ms": kern, "image_finite": finite,
        "roofline": {"bound": "hbm", "kernel": "k_gp_irr", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "ms_per_launch": round(float(gst.ms_irr), 4), "alg_bytes_per_launch": irr_b},
    }
    if not args.no_cpu:
        line.update(_gphoton_cpu_legs(d, img, rst, sc, torch))
    return line


def bench_psamp(args):
    """photonSampler_t: the gphoton room (1,000,000 photons, depth 2,
    search 100, radius 1, sun) with a 128-sample one-bounce pathlight, which
    samples by photonSampler_t (pathlight samples > 96 with the map
    published) at 640x480.  One step = the frame (the map is built once
    before the timed region).  Parity leg: the same scene at 96x72 rendered
    by the oracle from the GPU's map (boxes and raw photons); its time is
    the CPU baseline (one thread: the reference's Halton starts are drawn in
    area order)."""
    import numpy as np
    import torch

    from oracle import pyoracle as po
    from yafray_amd import lib as Y
    from yafray_amd import scenes as S

    def scene(rx, ry):
        return S.gphoton_room(rx, ry, subdiv=5, photons=1000000, depth=2, search=100, radius=1.0, samples=128,
                              path_depth=1, sun=(0.3, -0.2, 1.0))

    d = scene(640, 480)
    sc = Y.Scene.from_desc(d)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    sc.gphoton_build(1)
    dt, st = _timed(lambda: sc.render(out), args.steps, args.warmup)
    torch.cuda.synchronize()
    img = out.cpu().numpy()
    if not np.all(np.isfinite(img)):
        raise RuntimeError("psamp: non-finite pixels in the GPU image")
    rays = int(st.rays_camera + st.rays_extension + st.rays_shadow)
    line = {
        "metric": "photonSampler_t GI ms/frame", "value": round(dt * 1e3, 3), "unit": "ms/frame",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "room + icosphere(20480) + 2 boxes, 640x480; global photon map of 1,000,000 "
                               "photons, depth 2, search 100, radius 1; sun; pathlight 128 spp (photonSampler_t: "
                               "96-photon gather and 4 x 8 sector plan per pixel), 1 bounce",
                   "triangles": int(d.n_tris)},
        "rays_per_frame": rays, "mrays_per_s": round(rays / dt / 1e6, 3), "image_finite": True,
    }
    if not args.no_cpu:
        ds = scene(96, 72)
        scs = Y.Scene.from_desc(ds)
        outs = torch.zeros(96 * 72 * 3, dtype=torch.float32, device="cuda")
        scs.gphoton_build(1)
        sst = scs.render(outs)
        n = scs.gphoton_read()
        boxes = torch.zeros(max(n, 1) * 15, dtype=torch.float32, device="cuda")
        scs.gphoton_read(boxes)
        nr = scs.gphoton_read_photons()
        recs = torch.zeros(max(nr, 1) * 5, dtype=torch.int32, device="cuda")
        scs.gphoton_read_photons(recs)
        torch.cuda.synchronize()
        gimg = outs.cpu().numpy().reshape(72, 96, 3)
        print("psamp: oracle frame on the GPU's map (96x72)", file=sys.stderr, flush=True)
        oc = po.OracleScene.from_desc(ds, fast=True)
        oc.gphoton_load(ds, 1, boxes.cpu().numpy()[: n * 15].reshape(n, 15), stateless=True)
        oc.gphoton_load_photons(recs.cpu().numpy().view(np.uint32)[: nr * 5].reshape(nr, 5))
        t0 = time.perf_counter()
        ref, rs = oc.render(ds)
        cpu_s = time.perf_counter() - t0
        oc2 = po.OracleScene.from_desc(ds)
        oc2.gphoton_load(ds, 1, boxes.cpu().numpy()[: n * 15].reshape(n, 15), stateless=True)
        oc2.gphoton_load_photons(recs.cpu().numpy().view(np.uint32)[: nr * 5].reshape(nr, 5))
        ref2, rs2 = oc2.render(ds)
        rel = float(np.sqrt(np.mean((gimg.astype(np.float64) - ref2) ** 2)) / np.sqrt(np.mean(ref2.astype(np.float64) ** 2)))
        equal = (int(sst.rays_camera), int(sst.rays_extension), int(sst.rays_shadow)) == tuple(int(x) for x in rs2[:3])
        if rel > 1e-4 or not equal:
            raise RuntimeError("psamp parity failed: rel rmse %.3g, ray counts equal %s" % (rel, equal))
        crays = int(rs[0] + rs[1] + rs[2])
        line["rmse_vs_oracle"] = {"rel_rmse": rel, "bit_exact_pixels": float(np.mean(gimg == ref2)),
                                  "ray_counts_equal": equal, "sample": "96x72 frame of the same scene, oracle "
                                  "rendering from the GPU's map"}
        line["cpu_baseline"] = {"value": round(crays / cpu_s / 1e6, 4), "unit": "Mrays/s", "cores": 1,
                                "kind": "port", "sample": "oracle (-O3 -ffast-math build) rendering the 96x72 frame "
                                "(%d rays in %.1f s, one thread)" % (crays, cpu_s)}
    return line


def _gphoton_cpu_legs(d, img, rst, sc, torch):
    """Oracle legs of the global photon workload, on the host: the frame
    rendered with the GPU's boxes loaded (equal arithmetic), the oracle's own
    serial map and its frame (statistical agreement), and the CPU baseline:
    the serial map build timed on a bounded photon count and scaled, plus
    the frame on the host cores."""
    import numpy as np

    from oracle import pyoracle as po

    def rrmse(x, y):
        return float(np.sqrt(np.mean((x - y) ** 2)) / max(np.sqrt(np.mean(y ** 2)), 1e-30))

    threads = cpu_cores()
    n = sc.gphoton_read()
    boxes = torch.zeros(max(n, 1) * 15, dtype=torch.float32, device="cuda")
    sc.gphoton_read(boxes)
    torch.cuda.synchronize()
    boxes = boxes.cpu().numpy()[: n * 15].reshape(n, 15)
    print("gphoton: oracle frame on the GPU's map", file=sys.stderr, flush=True)
    oc = po.OracleScene.from_desc(d)
    oc.gphoton_load(d, 1, boxes, stateless=True)
    t0 = time.perf_counter()
    ref, rs = oc.render(d, threads=threads)
    frame_s = time.perf_counter() - t0
    g = img.astype(np.float64)
    same = rrmse(g, ref.astype(np.float64))
    counts_equal = (int(rst.rays_camera), int(rst.rays_extension), int(rst.rays_shadow)) == tuple(int(x) for x in rs[:3])
    if not (same <= 1e-3 and counts_equal):
        raise RuntimeError("gphoton parity failed: rel rmse %.3g, ray counts equal %s" % (same, counts_equal))
    print("gphoton: oracle serial map", file=sys.stderr, flush=True)
    o2 = po.OracleScene.from_desc(d)
    t0 = time.perf_counter()
    oc_counts = o2.gphoton_build(d, 1)
    map_s = time.perf_counter() - t0
    own, _ = o2.render(d, threads=threads)
    stat = rrmse(g, own.astype(np.float64))
    full = int(d.lights[1].photons)
    return {
        "cpu_baseline": {"value": round((map_s + frame_s) * 1e3, 1), "unit": "ms/frame", "cores": threads,
                         "kind": "port",
                         "sample": "oracle serial map build of %d photons (%.1f s, one thread, as the reference "
                                   "shoots them one after another) + the 640x480 frame on %d threads (%.1f s)"
                                   % (full, map_s, threads, frame_s)},
        "rmse_vs_oracle": {"rel_rmse_same_map": same, "ray_counts_equal": counts_equal,
                           "rel_rmse_oracle_own_map": stat,
                           "oracle_emitted_stored_boxes": [int(x) for x in oc_counts[:3]],
                           "sample": "full frame; same_map: the oracle renders with the GPU's boxes; own_map: "
                                     "the oracle's serial photon stream (statistical)"},
    }


def bench_cache(args):
    """Pathlight with the irradiance cache (cache = on) on the 1,045,506-
    triangle fBm terrain at 1280x720: 64 samples per cache sample, 3
    bounces, cache_size 0.01, search 50, shadow_threshold 0.3 (the
    factory's defaults), threshold 1e3 (no refinement passes).  One step =
    the whole frame: camera rays, the fake pass that places the samples,
    their takeSample paths, the cache's tree order and the render pass's
    interpolation plus the sun.

    Also: device time of the GI levels' extension and shadow launches (HIP
    events), the frame against the oracle's restatement of the reference
    (image and every sample), and the oracle's frame on the host cores as
    the CPU baseline (its fake pass is sequential, as the reference's)."""
    import numpy as np
    import torch

    from yafray_amd import descriptors as D
    from yafray_amd import lib as Y
    from yafray_amd import scenes as S

    d0 = S.terrain_gi(Y.device_heights)
    cached = D.pathlight(samples=64, depth=3, use_qmc=True, threshold=1e3, cache=True, cache_size=0.01, search=50,
                         shadow_threshold=0.3)
    lights = [cached if L.type == D.YG_LIGHT_PATH else L for L in d0.lights]
    d = D.SceneDesc(d0.verts, d0.idx, d0.tri_obj, d0.tri_mat, d0.obj_flags, d0.cam, list(d0.mats), list(d0.mods),
                    list(d0.texs), lights, d0.opts)
    sc = Y.Scene.from_desc(d)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    dt, st = _timed(lambda: sc.render(out), args.steps, args.warmup)
    kt = D.yg_render_opts.from_buffer_copy(d.opts)
    kt.flags = 3
    ks = sc.render(out, kt)
    torch.cuda.synchronize()
    img = out.cpu().numpy().reshape(d.cam.resy, d.cam.resx, 3)
    if not np.all(np.isfinite(img)):
        raise RuntimeError("cache: non-finite pixels in the GPU image")
    gs = sc.cache_read()
    rays = int(st.rays_camera + st.rays_extension + st.rays_shadow)
    bm = json.load(open(os.path.join(ROOT, "tests", "golden", "byte_model.json")))
    ext_b, sh_b = bm["extension"]["B_ray"], bm["shadow"]["B_ray"]
    bshadow = int(ks.rays_bounce - ks.rays_extension)
    kern = {"k_ext": {"ms": float(ks.ms_ext), "bytes": int(ks.rays_extension) * ext_b, "rays": int(ks.rays_extension)},
            "k_shadow": {"ms": float(ks.ms_shadow), "bytes": bshadow * sh_b, "rays": bshadow}}
    dom = max(kern, key=lambda k: kern[k]["ms"])
    launches = max(int(ks.bounce_launches), 1)
    achieved = kern[dom]["bytes"] / (kern[dom]["ms"] * 1e-3) / 1e9 if kern[dom]["ms"] > 0 else 0.0
    line = {
        "metric": "cached GI ms/frame", "value": round(dt * 1e3, 3), "unit": "ms/frame",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "pathlight irradiance cache on the 1,045,506-triangle fBm terrain, 1280x720, "
                               "64 samples per cache sample, 3 bounces, cache_size 0.01, search 50, "
                               "shadow_threshold 0.3, sun + sky", "triangles": int(d.n_tris)},
        "cache_samples": int(len(gs)), "rays_per_frame": rays, "mrays_per_s": round(rays / dt / 1e6, 3),
        "kernels_ms": {k: round(v["ms"], 3) for k, v in kern.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "ms_per_launch": round(kern[dom]["ms"] / launches, 4),
                     "alg_bytes_per_launch": int(kern[dom]["bytes"] / launches),
                     "model": "tests/golden/byte_model.json (terrain GI rays, canonical BVH2)"},
    }
    if not args.no_cpu:
        line.update(_cache_cpu_legs(d, img, gs))
    return line


def _cache_cpu_legs(d, img, gs, n_fp64=512):
    """Oracle legs of the cache workload, on the host.

    * the oracle renders the same frame with the cache (timed: the CPU
      baseline); its samples are compared with the GPU's one by one, up to
      the first that differs.  The cache is order-dependent (a visit takes a
      sample only if no earlier one covers it), so one camera or first-hit
      ray decided differently changes every later decision of its render
      area: the first differing sample's camera ray is classified with an
      FP64 edge test of the oracle's triangle (a false positive of the
      reference's float edge test, which trees decide differently);
    * both images against the converged FP64 render (4096 samples of the
      uncached pathlight, fixed-seed pixel sample): the statistical bar is
      that the GPU's cached image is as close to it as the oracle's."""
    import numpy as np

    from oracle import pyoracle as po
    from yafray_amd import descriptors as D

    threads = cpu_cores()
    print("cache: oracle frame", file=sys.stderr, flush=True)
    oc = po.OracleScene.from_desc(d)
    t0 = time.perf_counter()
    ref, _ = oc.render(d, threads=threads)
    cpu_dt = time.perf_counter() - t0
    os_, nnew = oc.cache_read()
    g64, r64 = img.astype(np.float64), ref.astype(np.float64)
    rel = float(np.sqrt(np.mean((g64 - r64) ** 2)) / np.sqrt(np.mean(r64 ** 2)))
    n = min(len(gs), len(os_))
    diff = np.nonzero(np.any(gs[:n, 0:15] != os_[:n, 0:15], axis=1) | (gs[:n, 16] != os_[:n, 16]))[0]
    first = int(diff[0]) if len(diff) else None
    cls = None
    if first is not None and gs[first, 16] == os_[first, 16] and np.any(gs[first, 0:3] != os_[first, 0:3]):
        # the camera ray of that pixel hit differently: FP64 edge test of the oracle's triangle
        pix = int(os_[first, 16])
        o, dirs = _camera_rays(d.cam, np.array([pix]))
        o32, d32 = o.astype(np.float32), dirs.astype(np.float32)
        pc, _ = oc.first_hit(o32, d32)
        A, B, C = d.verts[d.idx[int(pc[0])]].astype(np.float64)
        nrm = np.cross(B - A, C - A)
        f = o32[0].astype(np.float64) + d32[0].astype(np.float64) * D.MIN_RAYDIST
        r = d32[0].astype(np.float64)
        r = -r if np.dot(r, nrm) < 0 else r
        va, vb, vc = A - f, B - f, C - f
        tp = [np.dot(np.cross(va, vb), r), np.dot(np.cross(vb, vc), r), np.dot(np.cross(vc, va), r)]
        cls = {"pixel": pix, "oracle_triangle": int(pc[0]), "fp64_edge_products": [float(x) for x in tp],
               "oracle_hit_is_float_false_positive": bool(min(tp) < 0)}
    # the converged render of the same scene without the cache
    lights = [D.yg_light.from_buffer_copy(L) for L in d.lights]
    for L in lights:
        L.cache = 0
    dn = D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, d.cam, list(d.mats), list(d.mods),
                     list(d.texs), lights, d.opts)
    print("cache: FP64 converged pixels", file=sys.stderr, flush=True)
    fp = fp64_leg(dn, {"gpu_cache": img, "oracle_cache": ref}, n_fp64, oc=oc)
    ratio = fp["rel_rmse_gpu_cache"] / fp["rel_rmse_oracle_cache"]
    if not (ratio <= 1.05 and nnew == 0):
        raise RuntimeError("cache: GPU error vs converged %.4g is not within 5%% of the oracle's %.4g"
                           % (fp["rel_rmse_gpu_cache"], fp["rel_rmse_oracle_cache"]))
    return {
        "cpu_baseline": {"value": round(cpu_dt * 1e3, 1), "unit": "ms/frame", "cores": threads, "kind": "port",
                         "sample": "the whole frame by the oracle: fake pass and its samples on one thread "
                                   "(sequential in the reference), render pass on %d threads" % threads},
        "rmse_vs_oracle": {"rel_rmse": rel, "exact_px": float(np.mean(np.all(img == ref, axis=2))),
                           "samples": [int(len(gs)), int(len(os_))],
                           "samples_equal_before_first_difference": first if first is not None else int(n),
                           "first_difference": cls, "render_pass_new_samples": int(nnew)},
        "error_vs_fp64": dict(fp, gpu_over_oracle=round(ratio, 4),
                              bar="GPU cached image within 5% of the oracle's cached image in error against "
                                  "the converged render"),
    }


VALU_FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak (packed FMA); unpacked code reaches half
FBM_FLOP_PER_POINT = 384       # 6 octaves x ~64 FLOP (improved Perlin: 8 gradients, 7 lerps, 3 fades)


def bench_noise(args):
    """Improved-Perlin fBm on the device, the terrain's parameters (H 1,
    lacunarity 2, 6 octaves), over 64M points resident in HBM."""
    import numpy as np
    import torch

    from yafray_amd import descriptors as D
    from yafray_amd import lib as Y

    n = 1 << 26
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    pts = (torch.rand((n, 3), device="cuda", generator=g) * 64.0 - 32.0).contiguous()
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    tex = D.musgrave(H=1.0, lacunarity=2.0, octaves=6.0)
    for _ in range(args.warmup):
        Y.noise_eval(D.YG_EVAL_FBM, tex, pts, out)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.steps):
        Y.noise_eval(D.YG_EVAL_FBM, tex, pts, out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    tflops = n * FBM_FLOP_PER_POINT / (ms * 1e-3) / 1e12
    line = {
        "metric": "fBm Gpoints/s", "value": round(n / (ms * 1e-3) / 1e9, 3), "unit": "Gpoints/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "improved-Perlin fBm, H 1, lacunarity 2, 6 octaves, 67,108,864 points "
                               "uniform in [-32, 32]^3", "points": n},
        "roofline": {"bound": "valu", "achieved": round(tflops, 2), "peak": VALU_FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(tflops / VALU_FP32_PEAK_TFLOPS, 4),
                     "flop_per_point": FBM_FLOP_PER_POINT,
                     "hbm_gbs": round(n * 16 / (ms * 1e-3) / 1e9, 1)},
    }
    if not args.no_cpu:
        from oracle import pyoracle as po
        q = np.random.default_rng(3).choice(n, 65536, replace=False)
        host = pts.cpu().numpy()[q]
        t0 = time.perf_counter()
        ref = po.noise_eval(1, tex, host)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(len(q) / dt / 1e9, 6), "unit": "Gpoints/s", "cores": 1,
                                "kind": "port", "sample": "oracle fBm of 65,536 of the points, one thread"}
        line["exact_vs_oracle"] = float(np.mean(out.cpu().numpy()[q] == ref))
    return line


def bench_direct(args):
    """Plain direct lighting: box room + icosphere + two boxes, 1x1 area light
    with 16 samples, 640x480."""
    import torch

    from yafray_amd import lib as Y
    from yafray_amd import scenes as S

    d = S.direct_room()
    sc = Y.Scene.from_desc(d)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    dt, st = _timed(lambda: sc.render(out), args.steps, args.warmup)
    rays = st.rays_camera + st.rays_extension + st.rays_shadow
    fp64 = None
    if not args.no_cpu:
        fp64 = fp64_leg(d, {"gpu": out.cpu().numpy()}, 1024)
    return {
        "metric": "direct-lighting Mrays/s", "value": round(rays / dt / 1e6, 3), "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "room + icosphere(20480) + 2 boxes, area light 16 samples, 640x480",
                   "triangles": int(d.n_tris)},
        "rays_per_frame": int(rays), "lcg_draws": int(st.lcg_draws),
        "error_vs_fp64": fp64,
    }


def bench_room_gi(args):
    """Pathlight GI whose path vertices light an area light (use_in_indirect):
    the direct-lighting room, 640x480, area light 16 samples, pathlight 4
    samples and 2 bounces.  Also times the same frame with the area light kept
    out of the paths, the cost of the deferred vertex pass beside it."""
    import torch

    from yafray_amd import descriptors as D
    from yafray_amd import lib as Y
    from yafray_amd import scenes as S

    base = S.direct_room()
    out = torch.zeros(base.cam.resx * base.cam.resy * 3, dtype=torch.float32, device="cuda")
    res = {}
    for in_paths in (1, 0):
        lights = list(base.lights)
        lights[0].use_in_indirect = in_paths
        lights.append(D.pathlight(samples=4, depth=2, use_qmc=True, threshold=1e3))
        d = D.SceneDesc(base.verts, base.idx, base.tri_obj, base.tri_mat, base.obj_flags, base.cam,
                        list(base.mats), [], [], lights, base.opts)
        sc = Y.Scene.from_desc(d)
        dt, st = _timed(lambda: sc.render(out), args.steps, args.warmup)
        res[in_paths] = (dt, st.rays_camera + st.rays_extension + st.rays_shadow, st.lcg_draws)
        if in_paths:
            finite = bool(torch.isfinite(out).all().item())
            if not finite:
                raise RuntimeError("room_gi: non-finite pixels in the GPU image")
            parity = None if args.no_cpu else _room_gi_parity(d)
    dt, rays, draws = res[1]
    return {
        "metric": "GI Mrays/s", "value": round(rays / dt / 1e6, 3), "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "room + icosphere(20480) + 2 boxes, 640x480, area light 16 samples lit at "
                               "the primary hit and at every GI vertex, pathlight 4 spp, 2 bounces",
                   "triangles": int(base.n_tris)},
        "rays_per_frame": int(rays), "lcg_draws": int(draws),
        "area_light_outside_paths": {"ms_per_step": round(res[0][0] * 1e3, 3),
                                     "rays_per_frame": int(res[0][1]),
                                     "mrays_per_s": round(res[0][1] / res[0][0] / 1e6, 3)},
        "image_finite": finite,
        "rmse_vs_oracle": parity,
        "error_vs_fp64": None,
    }


def _room_gi_parity(d, res=(160, 120)):
    """The room_gi frame at reduced resolution (same camera, scene and
    lights) on the GPU and on the oracle: image relative RMSE, and the
    ourRandom() draws and ray counts of both, which must be equal."""
    import numpy as np
    import torch

    from oracle import pyoracle as po
    from yafray_amd import descriptors as D
    from yafray_amd import lib as Y

    cam = D.yg_camera.from_buffer_copy(d.cam)
    cam.resx, cam.resy = res
    ds = D.SceneDesc(d.verts, d.idx, d.tri_obj, d.tri_mat, d.obj_flags, cam, list(d.mats), [], [],
                     list(d.lights), d.opts)
    sc = Y.Scene.from_desc(ds)
    out = torch.zeros(res[0] * res[1] * 3, dtype=torch.float32, device="cuda")
    st = sc.render(out)
    torch.cuda.synchronize()
    img = out.cpu().numpy().reshape(res[1], res[0], 3).astype(np.float64)
    ref, rs = po.OracleScene(ds.verts, ds.idx, ds.tri_obj, ds.tri_mat).render(ds, threads=cpu_cores())
    ref = ref.astype(np.float64)
    rel = float(np.sqrt(np.mean((img - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    same = (int(st.lcg_draws) == int(rs[4]) and
            (st.rays_camera, st.rays_extension, st.rays_shadow) == tuple(int(x) for x in rs[:3]))
    if not (rel <= 1e-3 and same):
        raise RuntimeError("room_gi parity failed: rel rmse %.3g, draws %d vs %d" % (rel, st.lcg_draws, rs[4]))
    return {"rel_rmse": rel, "lcg_draws_equal": same, "sample": "same scene at %dx%d, GPU vs oracle" % res}


def bench_autotune(args):
    """Autotuner on the GI terrain frame (1280x720, 4 spp, 3 bounces): every
    (BVH width, leaf size) candidate built and timed on the frame."""
    import torch

    from yafray_amd import lib as Y
    from yafray_amd import scenes as S

    d = S.terrain_gi(Y.device_heights)
    sc = Y.Scene.from_desc(d)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    default_dt, _ = _timed(lambda: sc.render(out), args.steps, args.warmup)
    t0 = time.perf_counter()
    best, best_ms, cands = sc.autotune(out, widths=(2, 4), leaves=(2, 4, 8), stacks=(-1, 8, 16, 32), repeats=2)
    tune_s = time.perf_counter() - t0
    tuned_dt, _ = _timed(lambda: sc.render(out), args.steps, args.warmup)
    return {
        "metric": "autotuned GI ms/frame", "value": round(tuned_dt * 1e3, 3), "unit": "ms/frame",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(tuned_dt * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "path-traced GI on the 1,045,506-triangle fBm terrain, 1280x720, 4 spp, "
                               "3 bounces; BVH width 2/4 x max_leaf 2/4/8 x LDS stack entries all/8/16/32, "
                               "32 SAH bins"},
        "default_build_ms_per_frame": round(default_dt * 1e3, 3),
        "best": {"width": best.width, "max_leaf": best.max_leaf, "sah_bins": best.sah_bins,
                 "stack_lds": best.stack_lds, "frame_device_ms": round(best_ms, 3)},
        "candidates": [{"width": w, "max_leaf": lf, "sah_bins": b, "stack_lds": sl, "frame_device_ms": round(ms, 3),
                        "build_ms": round(bms, 1)} for (w, lf, b, ms, bms, sl) in cands],
        "tune_s": round(tune_s, 2),
    }


def bench_aa(args):
    """Adaptive antialiasing on the GI terrain frame: 1280x720, 4 spp, 3 bounces,
    AA_passes 2, AA_minsamples 4, AA_threshold 0.03, jittered first pass."""
    import torch

    from yafray_amd import descriptors as D
    from yafray_amd import lib as Y
    from yafray_amd import scenes as S

    d = S.terrain_gi(Y.device_heights)
    d.opts = D.render_opts(bias=2e-3, raydepth=3, background=(0.4, 0.5, 0.7), aa_passes=2, aa_minsamples=4,
                           aa_threshold=0.03, aa_jitterfirst=True)
    sc = Y.Scene.from_desc(d)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    dt, st = _timed(lambda: sc.render(out), args.steps, args.warmup)
    rays = st.rays_camera + st.rays_extension + st.rays_shadow
    plain = D.yg_render_opts.from_buffer_copy(d.opts)
    plain.aa_passes = 0
    n_first = sc.render(out, plain).rays_camera  # first-pass camera rays (area overlaps included)
    return {
        "metric": "antialiased GI Mrays/s", "value": round(rays / dt / 1e6, 2), "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "path-traced GI on the 1,045,506-triangle fBm terrain, 1280x720, 4 spp, "
                               "3 bounces, AA_passes 2, AA_minsamples 4, AA_threshold 0.03, AA_jitterfirst"},
        "rays_per_frame": int(rays), "camera_rays": int(st.rays_camera),
        "aa_samples_per_first_pass_ray": round((st.rays_camera - n_first) / n_first, 4),
    }


def bench_rays(args):
    """Ray-query entry points on the 1,045,506-triangle terrain: 4M incoherent
    rays per batch (origins above the terrain, random downward-biased
    directions), closest hit (kdTree_t::Intersect), any hit (IntersectS) and
    the scene_t::firstHit / isShadowed forms, device arrays in HBM."""
    import numpy as np
    import torch

    from yafray_amd import lib as Y
    from yafray_amd import scenes as S

    verts, idx, band = S.terrain_mesh(1024, 512, Y.device_heights)
    sc = Y.Scene()
    sc.add_mesh(verts, idx, band)
    sc.build()
    n = 1 << 22
    rng = np.random.default_rng(7)
    o = np.stack([rng.uniform(0, 2, n), rng.uniform(0, 1, n), rng.uniform(0.3, 0.8, n)], 1).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d[:, 2] = -np.abs(d[:, 2]) - 0.2
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 4:7], rays[:, 7] = o, d, np.inf
    dr = torch.from_numpy(rays).cuda()
    hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    occ = torch.empty(n, dtype=torch.uint8, device="cuda")
    frm, dd = torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()
    sun = torch.from_numpy(np.tile(np.float32([0.53, 0.27, 0.80]), (n, 1))).cuda()
    self_obj = torch.zeros(n, dtype=torch.int32, device="cuda")
    res = {}
    for name, fn in (("closest", lambda: sc.trace_closest(dr, hits)),
                     ("any", lambda: sc.trace_any(dr, occ)),
                     ("first_hit", lambda: sc.first_hit(frm, dd, hits)),
                     ("is_shadowed_dir", lambda: sc.is_shadowed(frm, sun, self_obj, occ, False))):
        dt, _ = _timed(fn, args.steps, args.warmup)
        res[name] = {"ms": round(dt * 1e3, 3), "mrays_per_s": round(n / dt / 1e6, 1)}
    return {
        "metric": "ray queries Mrays/s (closest hit)", "value": res["closest"]["mrays_per_s"], "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["closest"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "4,194,304 incoherent rays per batch on the 1,045,506-triangle fBm terrain "
                               "(BVH width 4)"},
        "entry_points": res,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="gi",
                    choices=["gi", "batched", "adjoint", "photon", "direct", "autotune", "aa", "rays",
                             "noise", "room_gi", "gphoton", "cache", "psamp"])
    ap.add_argument("--batch-n", type=int, default=64)
    ap.add_argument("--batch-mode", default="both", choices=["both", "sequential", "batched"],
                    help="batched workload: time both modes (default) or one of them alone")
    ap.add_argument("--width", type=int, default=0, choices=[0, 2, 4, 8],
                    help="GI workload: BVH width to measure (0 = the autotuner's choice)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB device copy after the timed region (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(rendering would shard by image tiles over GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.workload != "gi":
        import torch
        torch.cuda.set_device(0)
        fn = {"batched": bench_batched, "adjoint": bench_adjoint, "photon": bench_photon,
              "direct": bench_direct, "autotune": bench_autotune, "aa": bench_aa,
              "rays": bench_rays, "noise": bench_noise, "room_gi": bench_room_gi,
              "gphoton": bench_gphoton, "cache": bench_cache, "psamp": bench_psamp}[args.workload]
        print(json.dumps(fn(args)))
        return

    g = gpu_leg(args)
    ms = 1000.0 * g["dt"] / args.steps
    value = g["rays"] / g["dt"] / 1e6
    st = g["stats"]
    bm = json.load(open(os.path.join(ROOT, "tests", "golden", "byte_model.json")))
    # algorithmic bytes of the GI levels' traversal launches: extension rays
    # (k_ext, closest hit) and shadow rays (k_shadow, any hit) at the frozen
    # canonical-BVH2 visit counts; the dominant one is reported
    ext_b = bm["extension"]["B_ray"]
    sh_b = bm["shadow"]["B_ray"]
    bshadow = st.rays_bounce - st.rays_extension
    levels = max(int(st.bounce_launches), 1)
    kern = {
        "k_ext": {"ms": st.ms_ext, "bytes": st.rays_extension * ext_b, "rays": int(st.rays_extension)},
        "k_shadow": {"ms": st.ms_shadow, "bytes": bshadow * sh_b, "rays": int(bshadow)},
    }
    for k in kern.values():
        k["gbs"] = k["bytes"] / (k["ms"] * 1e-3) / 1e9 if k["ms"] > 0 else 0.0
    dom = max(kern, key=lambda k: kern[k]["ms"])
    achieved = kern[dom]["gbs"]
    traffic = measured_traffic(dom)
    line = {
        "metric": "GI Mrays/s",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        # pixel samples: camera rays x pathlight samples per frame
        "msamples_per_s": round(g["desc"].cam.resx * g["desc"].cam.resy * path_spp(g["desc"])
                                / (ms * 1e-3) / 1e6, 3),
        "config": {
            "workload": "path-traced GI: 1024x512-vertex fBm terrain (1,045,506 tris), 1280x720, "
                        "4 spp, 3 bounces, sun + constant sky, 64 height-band materials",
            "triangles": int(g["info"]["tris"]),
            "resolution": "1280x720",
            "spp": 4,
            "depth": 3,
            "rays_per_step": int(g["rays"] / args.steps),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": TRAFFIC_PROFILE if traffic is not None else None,
            "launches": levels,
            "ms_per_launch": round(kern[dom]["ms"] / levels, 4),
            "alg_bytes_per_launch": round(kern[dom]["bytes"] / levels),
            "kernels": {k: {"ms_per_launch": round(v["ms"] / levels, 4), "rays": v["rays"],
                            "achieved_gbs": round(v["gbs"], 1)} for k, v in kern.items()},
            "gi_levels_ms": round(st.ms_bounce, 3),
        },
        "bvh": {"nodes": int(g["info"]["nodes"]), "max_depth": int(g["info"]["max_depth"]),
                "dev_bytes": int(g["info"]["dev_bytes"])},
        "autotune": g["tune"],
        # BVH built on the host and on the device at the tuned settings (after the timed region)
        "bvh_builders": g["builders"],
        "image_finite": g["finite"],
        # 1 = frame replayed from its captured hipGraph, 2 = captured in that call, 0 = kernel by kernel
        "frame_graph_modes": g["graph_modes"],
    }
    if args.copy_probe:
        line["copy_probe"] = copy_probe()
    if not args.no_cpu:
        cores = cpu_cores()
        p = parity_leg(g["desc"], g["img"], cores)
        line["cpu_baseline"] = {
            "value": round(p["cpu_rays"] / p["cpu_dt"] / 1e6, 4),
            "unit": "Mrays/s",
            "cores": cores,
            "kind": "port",
            "sample": "oracle (C restatement, built -O3 -ffast-math like the reference) rendering the full "
                      "1280x720 frame (4 spp, 3 bounces): %d rays in %.1f s on %d threads (%.0f core-seconds)"
                      % (p["cpu_rays"], p["cpu_dt"], cores, p["cpu_dt"] * cores),
            "strict_build_mrays_per_s": round(p["strict_mrays"], 4),
        }
        line["rmse_vs_oracle"] = {"rel_rmse": p["rel_rmse"], "bit_exact_pixels": p["exact_px"],
                                  "sample": "full 1280x720 frame, GPU image vs oracle image"}
        # both 4-spp images (GPU and the FP32 oracle) against the converged FP64 render
        line["error_vs_fp64"] = fp64_leg(g["desc"], {"gpu": g["img"], "oracle_fp32": p["ref"]}, 1024,
                                         oc=p["oracle"])
    print(json.dumps(line))


if __name__ == "__main__":
    main()
