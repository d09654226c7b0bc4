"""Frozen algorithmic byte model for the roofline of the GI bounce kernel.

Traces a fixed 65,536-ray sample of the path-traced GI workload (terrain of
1024x512 vertices, camera of the 1280x720 render) through the oracle's
canonical BVH2 (binned SAH, 32 bins, leaf <= 4, FP64 boxes) and records the
mean node visits and triangle tests per ray:

  B_ray = 32 B ray + 16 B hit (4 B for a shadow query) + n_node * 32 B + n_tri * 36 B

Extension rays: cosine-weighted directions (exact base-2/3 radical inverses)
from the camera rays' first hits.  Shadow rays: from the same points toward
the sun.  Output: tests/golden/byte_model.json.

Run: python tests/golden/make_byte_model.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from yafray_amd import scenes as S  # noqa: E402

N = 65536


def radinv(b, i):
    r, f = 0.0, 1.0 / b
    while i:
        r += (i % b) * f
        i //= b
        f /= b
    return r


def camera_rays(cam, resx, resy, idx):
    pos = np.array(cam.pos, np.float64)
    look = np.array(cam.look, np.float64)
    up = np.array(cam.up, np.float64)
    vup, vto = up - pos, look - pos
    vright = np.cross(vup, vto)
    vup = np.cross(vright, vto)
    vup /= np.linalg.norm(vup)
    vright /= -np.linalg.norm(vright)
    vto /= np.linalg.norm(vto)
    vup *= cam.aspect * resy / resx
    vto = vto * cam.focal - 0.5 * (vup + vright)
    vup /= resy
    vright /= resx
    i, j = idx // resx, idx % resx
    d = vright[None] * (j + 0.5)[:, None] + vup[None] * (i + 0.5)[:, None] + vto[None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.repeat(pos[None], len(idx), 0), d


def main():
    d = S.terrain_gi(po.heightfield)
    sc = po.OracleScene(d.verts, d.idx, None, d.tri_mat)
    resx, resy = d.cam.resx, d.cam.resy
    idx = (np.arange(N, dtype=np.int64) * (resx * resy // N)) % (resx * resy)
    o, dr = camera_rays(d.cam, resx, resy, idx)
    prim, t = sc.first_hit(o.astype(np.float32), dr.astype(np.float32))
    ok = prim >= 0
    p = (o + dr * (t[:, None] + 5e-5))[ok]
    n = d.verts[d.idx[prim[ok]]].astype(np.float64)
    nrm = np.cross(n[:, 1] - n[:, 0], n[:, 2] - n[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm *= np.sign(np.sum(nrm * -dr[ok], axis=1))[:, None]
    m = len(p)
    z1 = np.array([radinv(2, k + 1) for k in range(m)])
    z2 = 2 * np.pi * np.array([radinv(3, k + 1) for k in range(m)])
    u = np.cross(nrm, [0.0, 0.0, 1.0] * np.ones((m, 1)))
    bad = np.linalg.norm(u, axis=1) < 1e-6
    u[bad] = [1.0, 0.0, 0.0]
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    v = np.cross(nrm, u)
    w = (u * np.cos(z2)[:, None] + v * np.sin(z2)[:, None]) * np.sqrt(1 - z1)[:, None] + nrm * np.sqrt(z1)[:, None]
    rays = np.zeros((m, 8), np.float32)
    rays[:, :3] = p + w * 5e-5
    rays[:, 4:7] = w
    rays[:, 7] = np.inf
    sc.trace(rays, mode=0, threads=8)
    ext_nodes, ext_tris = sc.last_counts
    sun = np.array([1.0, 0.5, 1.5])
    sun /= np.linalg.norm(sun)
    srays = np.zeros((m, 8), np.float32)
    srays[:, :3] = p + sun * 2e-3
    srays[:, 4:7] = sun
    srays[:, 7] = np.inf
    sc.trace(srays, mode=1, threads=8)
    sh_nodes, sh_tris = sc.last_counts
    out = {
        "workload": "path-traced GI, 1,045,506-triangle fBm terrain, 1280x720 camera",
        "rays": int(m),
        "canonical_bvh": "binned SAH, 32 bins, leaf <= 4, FP64 boxes (oracle/yo_core.c)",
        "extension": {"n_node": ext_nodes / m, "n_tri": ext_tris / m},
        "shadow": {"n_node": sh_nodes / m, "n_tri": sh_tris / m},
        "bytes": {"ray": 32, "hit": 16, "occluded": 4, "node": 32, "tri": 36},
    }
    out["extension"]["B_ray"] = 32 + 16 + 32 * out["extension"]["n_node"] + 36 * out["extension"]["n_tri"]
    out["shadow"]["B_ray"] = 32 + 4 + 32 * out["shadow"]["n_node"] + 36 * out["shadow"]["n_tri"]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "byte_model.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
