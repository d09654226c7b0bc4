"""Full-size parity on the 1,045,506-triangle fBm terrain at 1280x720, 4 spp,
3 bounces: the GPU image against the oracle image, and every first-hit and
shadow query the oracle's render made, replayed through the C ABI.

Bar: image relative RMSE <= 1e-3; first-hit primitive ids identical on all
but a 1e-6 fraction of queries, and each remaining difference a tie between
two triangles that share a vertex, both accepted by the reference's float edge
test (triangle.h:48-58) -- near-edge-on triangles seen from afar pass that
test by rounding, and which of two such neighbours is reached first depends
on the acceleration structure, in the reference's kd-tree as here.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from yafray_amd import lib as Y
from yafray_amd import scenes as S


def _shares_vertex(idx, a, b):
    return len(set(idx[a].tolist()) & set(idx[b].tolist())) > 0


@pytest.mark.gpu
def test_terrain_full_frame_parity(torch_cuda):
    torch = torch_cuda
    d = S.terrain_gi(Y.device_heights)
    print("full frame: terrain 1280x720 4spp depth 3", flush=True)
    oc = po.OracleScene(d.verts, d.idx, d.tri_obj, d.tri_mat)
    ref, rs, ext, sh = oc.render_logged(d, threads=min(os.cpu_count() or 1, 16),
                                        ext_cap=9_500_000, sh_cap=7_500_000)
    sc = Y.Scene.from_desc(d)
    out = torch.zeros(d.cam.resx * d.cam.resy * 3, dtype=torch.float32, device="cuda")
    st = sc.render(out)
    img = out.cpu().numpy().reshape(ref.shape).astype(np.float64)
    ref64 = ref.astype(np.float64)
    rel = np.sqrt(np.mean((img - ref64) ** 2)) / np.sqrt(np.mean(ref64 ** 2))
    exact = np.mean(np.all(img == ref64, axis=2))
    print("rel_rmse %.3g exact_px %.7f" % (rel, exact), flush=True)
    assert rel <= 1e-3
    assert exact >= 0.9999
    assert abs(int(st.rays_extension) - int(rs[1])) <= 1e-5 * rs[1]

    # first-hit queries (camera rays and path extensions)
    mr, sb = d.opts.min_raydis, d.opts.self_bias
    frm = torch.from_numpy(np.ascontiguousarray(ext[:, 0:3])).cuda()
    dr = torch.from_numpy(np.ascontiguousarray(ext[:, 3:6])).cuda()
    hits = torch.empty((len(ext), 4), dtype=torch.float32, device="cuda")
    sc.first_hit(frm, dr, hits, mr)
    h = hits.cpu().numpy()
    gp, gt = h[:, 0].view(np.int32), h[:, 1]
    op, ot = ext[:, 6].view(np.int32), ext[:, 7]
    bad = np.nonzero((gp != op) | ((op >= 0) & (gt != ot)))[0]
    print("first-hit queries %d, differing %d" % (len(ext), len(bad)), flush=True)
    assert len(bad) <= 1e-6 * len(ext)
    idx = d.idx.reshape(-1, 3)
    for i in bad:
        assert gp[i] >= 0 and op[i] >= 0, "hit/miss difference on query %d" % i
        assert _shares_vertex(idx, gp[i], op[i]), "non-adjacent difference on query %d" % i
        f = ext[i, 0:3] + ext[i, 3:6] * np.float32(mr)
        ray = np.array([[*f, 0, *ext[i, 3:6], np.inf]], np.float32)
        # both answers are triangles the float edge test accepts along this ray
        for tri in (gp[i], op[i]):
            one = po.OracleScene(d.verts, idx[tri:tri + 1])
            p1, _, _, _ = one.trace(ray, brute=True)
            assert p1[0] == 0

    # shadow queries (sun, direction form)
    p = torch.from_numpy(np.ascontiguousarray(sh[:, 0:3])).cuda()
    l = torch.from_numpy(np.ascontiguousarray(sh[:, 3:6])).cuda()
    so = torch.from_numpy(np.ascontiguousarray(sh[:, 6].view(np.int32))).cuda()
    code = sh[:, 7].view(np.int32)
    assert np.all((code >> 1) == 0)
    occ = torch.empty(len(sh), dtype=torch.uint8, device="cuda")
    sc.is_shadowed(p, l, so, occ, False, mr, sb)
    diff = int(np.sum(occ.cpu().numpy() != (code & 1)))
    print("shadow queries %d, differing %d" % (len(sh), diff), flush=True)
    assert diff <= 1e-6 * len(sh)
