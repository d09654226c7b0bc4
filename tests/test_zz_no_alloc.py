"""yg_trace_any_noalloc (the allocation-free any-hit entry point) against
yg_trace_any through the C ABI: identical occlusion bits on two calls in a
row, the second with fewer rays, on trees whose traversal stacks overflow
the LDS part into the scratch array."""
import numpy as np
import pytest

from yafray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _rays(n, seed, tmax):
    rng = np.random.default_rng(seed)
    o = np.stack([rng.uniform(0, 2, n), rng.uniform(0, 1, n), rng.uniform(0.3, 0.8, n)], 1)
    d = rng.normal(size=(n, 3))
    d[:, 2] = -np.abs(d[:, 2]) - 0.2
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3], r[:, 4:7], r[:, 7] = o, d, tmax
    return r


@pytest.mark.parametrize("width", [2, 4, 8])
def test_noalloc_any_hit_equals_yg_trace_any(torch_cuda, width):
    torch = torch_cuda
    from yafray_amd import lib as Y
    verts, idx, band = S.terrain_mesh(512, 256, Y.device_heights)
    sc = Y.Scene()
    sc.add_mesh(verts, idx, band)
    sc.build(width=width)
    assert sc.info()["max_depth"] > 16   # stacks deeper than the 16 LDS entries
    for call, (n, seed) in enumerate([(1 << 21, 3), (100_003, 4)]):
        print("noalloc call %d: width %d, %d rays, seed %d" % (call, width, n, seed), flush=True)
        rays = torch.from_numpy(_rays(n, seed, 0.8)).cuda()
        ref = torch.full((n,), 7, dtype=torch.uint8, device="cuda")
        got = torch.full((n,), 9, dtype=torch.uint8, device="cuda")
        sc.trace_any(rays, ref)
        sc.trace_any_noalloc(rays, got)
        torch.cuda.synchronize()
        assert torch.equal(ref, got)
        frac = float(got.float().mean())
        assert 0.05 < frac < 0.95
