"""ff_trace_closest_noalloc, the closest-hit entry point that allocates no
device memory (its work counters and traversal stack overflow live in a
fixed per-device scratch), called through the C ABI and checked against
ff_trace_closest: triangle id, t, u, v and the traversal counters
bit-identical, on two calls in a row (the second smaller, so the shared
scratch's leftovers from the first must not matter), for two-child, leaf
batching and 4-wide traversal variants; the device's free memory is the same
before and after; a BVH deeper than the scratch holds is refused."""
import sys

import numpy as np
import pytest

from firefly_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _variant(width, ww):
    for v, _, _, _, w, wd in abi.traversal_variants():
        if (wd, w) == (width, ww):
            return v
    raise AssertionError("no such variant")


def _rays(sd, n, seed):
    r = scenes.random_rays(n, seed, lo=-1.3, hi=1.3, y_lo=0.0, y_hi=2.4)
    return r


@pytest.fixture(scope="module")
def mesh(gpu):
    sd = scenes.icosphere_scene(64, 64, 1, 4, subdiv=6)
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    yield sd, sc
    sc.close()


@pytest.mark.parametrize("shape", [(2, 0), (2, 5), (4, 0)], ids=["binary", "leaf_batching", "wide"])
def test_noalloc_equals_trace_closest(gpu, mesh, shape):
    import torch
    sd, sc = mesh
    v = _variant(*shape)
    abi.set_traversal_variant(gpu, v)
    try:
        for call, (n, seed, two_sided) in enumerate(((300000, 11, False), (1537, 12, True))):
            print(f"noalloc: variant {v}, call {call}, {n} rays", file=sys.stderr, flush=True)
            rays = torch.from_numpy(_rays(sd, n, seed)).cuda()
            h_ref = torch.empty((n, 4), dtype=torch.float32, device="cuda")
            h_na = torch.full((n, 4), 7.0, dtype=torch.float32, device="cuda")
            c_ref = torch.zeros(2, dtype=torch.int64, device="cuda")
            c_na = torch.zeros(2, dtype=torch.int64, device="cuda")
            sc.trace_closest(rays, h_ref, two_sided, c_ref)
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            sc.trace_closest_noalloc(rays, h_na, two_sided, c_na)
            torch.cuda.synchronize()
            free1 = torch.cuda.mem_get_info()[0]
            a, b = h_ref.cpu().numpy(), h_na.cpu().numpy()
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"call {call}: hits differ"
            assert torch.equal(c_ref, c_na), f"call {call}: traversal counters differ"
            assert np.count_nonzero(a[:, 3].view(np.int32) >= 0) > n // 10
            assert free1 == free0, f"call {call}: free device memory {free0} -> {free1}"
    finally:
        abi.set_traversal_variant(gpu, 0)


def test_noalloc_refuses_a_bvh_deeper_than_its_scratch(gpu):
    """60 triangles along a line, each 4x farther than the last: an SAH tree
    of depth 43, beyond the 39 levels the fixed scratch holds for two-child
    records.  ff_trace_closest (which sizes its own overflow area) still
    traces it."""
    import torch
    n = 60
    c = 2.0 ** (np.arange(n) * 2.0)
    tri = np.zeros((n, 3, 3), np.float32)
    tri[:, 0, 0] = c
    tri[:, 1, 0] = c * 1.01
    tri[:, 2, 0] = c
    tri[:, 2, 1] = c * 0.01
    nrm = np.zeros_like(tri)
    nrm[..., 2] = 1.0
    sc = abi.Scene(gpu, tri, nrm, np.zeros((n, 3, 2), np.float32), np.zeros(n, np.int32),
                   scenes.material(abi.FF_MAT_DIFFUSE)[None])
    assert sc.max_depth > 39
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0] = c * 1.002
    rays[:, 1] = c * 0.002
    rays[:, 2] = -1.0
    rays[:, 6] = 1.0
    r = torch.from_numpy(rays).cuda()
    h = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(abi.FFError) as e:
        sc.trace_closest_noalloc(r, h)
    assert e.value.status == 5 and b"depth" in abi._lib().ff_last_error()
    sc.trace_closest(r, h, True)
    torch.cuda.synchronize()
    _, ordered = sc.bvh()
    slot = h.cpu().numpy()[:, 3].view(np.int32)
    # (the edge products of the far triangles overflow float; the near ones are exact)
    assert [int(ordered[s]) if s >= 0 else -1 for s in slot[:30]] == list(range(30))
    sc.close()
