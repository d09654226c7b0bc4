"""Frozen algorithmic byte model for the roofline of the GI bounce kernels.

Renders the path-traced GI workload (1,045,506-triangle fBm terrain,
1280x720, 4 samples, 3 bounces) with the oracle while logging every ray
query, then prices a fixed-seed 65,536-query sample of each kind on the
canonical tight BVH2 of oracle/yo_bytemodel.c (binned SAH, 32 bins, leaf
<= 4, exact FP64 boxes, near child first, closest-hit shrinking):

  B_ray = 32 B ray + 16 B hit (4 B for an any-hit query)
          + n_node * 32 B (nodes visited) + n_tri * 36 B (triangles tested)

(n_box_tested, the child boxes the visits test to order their children, is
recorded beside it; a layout holding both child boxes in the parent fetches
those, so the figure above is the smaller of the two accountings.)

* extension: the pathlight's firstHit queries of every level (the k_ext
  launches), i.e. all logged first-hit queries except the camera rays;
* shadow: the sun's isShadowed queries at path vertices (the k_shadow
  launches), i.e. all logged shadow queries except those at the camera rays'
  hits (the primary hits' sun rays run in k_sun_vis).

The parity tree of oracle/yo_core.c is not used: its boxes are grown (5% of
their extent plus slack) so that the oracle agrees bit for bit with the
reference's float edge test, which would overstate the work.

Output: tests/golden/byte_model.json.  Run: python tests/golden/make_byte_model.py
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
SEED = 20250607


def _rows(a):
    """float32 [n, 3] -> hashable byte rows"""
    a = np.ascontiguousarray(a, np.float32)
    return a.view(np.dtype((np.void, 12))).reshape(-1)


def main():
    d = S.terrain_gi(po.heightfield)
    oc = po.OracleScene(d.verts, d.idx, None, d.tri_mat)
    img, st, ext, sh = oc.render_logged(d, threads=8, ext_cap=10_000_000, sh_cap=7_000_000)
    mind = np.float32(d.opts.min_raydis)
    cam = np.array(d.cam.pos[:], np.float32)
    frm, ray = ext[:, 0:3], ext[:, 3:6]
    is_cam = np.all(frm == cam[None, :], axis=1)
    # primary hit points, as the oracle forms them: (from + ray*min_raydis) + ray*t
    f = (frm[is_cam] + ray[is_cam] * mind).astype(np.float32)
    t = ext[is_cam, 7:8]
    prim_ok = ext[is_cam, 6].view(np.int32) >= 0
    P = (f + ray[is_cam] * t).astype(np.float32)[prim_ok]
    at_primary = np.isin(_rows(sh[:, 0:3]), _rows(P))
    ext_q = ext[~is_cam]
    sh_q = sh[~at_primary]
    rng = np.random.default_rng(SEED)
    pe = ext_q[rng.choice(len(ext_q), N, replace=False)]
    ps = sh_q[rng.choice(len(sh_q), N, replace=False)]
    # extension queries: origin from + ray*min_raydis, closest hit over (0, inf)
    er = np.zeros((N, 8), np.float32)
    er[:, 0:3] = pe[:, 0:3] + pe[:, 3:6] * mind
    er[:, 4:7] = pe[:, 3:6]
    er[:, 7] = np.inf
    # sun shadow queries (direction form): the terrain is the shading point's
    # own and only object, so every triangle is tested from p + dir*self_bias
    assert np.all((ps[:, 7].view(np.int32) >> 1) == 0)
    dn = ps[:, 3:6] / np.linalg.norm(ps[:, 3:6], axis=1, keepdims=True)
    sr = np.zeros((N, 8), np.float32)
    sr[:, 0:3] = ps[:, 0:3] + dn * np.float32(d.opts.self_bias)
    sr[:, 4:7] = dn
    sr[:, 7] = np.inf
    en, et, eb, nodes = po.canonical_counts(d.verts, d.idx, er, 0)
    sn, stt, sb, _ = po.canonical_counts(d.verts, d.idx, sr, 1)
    out = {
        "workload": "path-traced GI, 1,045,506-triangle fBm terrain, 1280x720, 4 spp, 3 bounces",
        "canonical_bvh": "tight binned-SAH BVH2, 32 bins, leaf <= 4, exact FP64 boxes, near child first "
                         "(oracle/yo_bytemodel.c), %d nodes" % nodes,
        "queries": {"extension_all_levels": int(len(ext_q)), "camera_excluded": int(is_cam.sum()),
                    "shadow_at_path_vertices": int(len(sh_q)), "shadow_at_primary_excluded": int(at_primary.sum()),
                    "sample_per_kind": N, "seed": SEED},
        "extension": {"n_node": en, "n_tri": et, "n_box_tested": eb},
        "shadow": {"n_node": sn, "n_tri": stt, "n_box_tested": sb},
        "bytes": {"ray": 32, "hit": 16, "occluded": 4, "node": 32, "tri": 36},
    }
    out["extension"]["B_ray"] = 32 + 16 + 32 * en + 36 * et
    out["shadow"]["B_ray"] = 32 + 4 + 32 * sn + 36 * stt
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "byte_model.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
