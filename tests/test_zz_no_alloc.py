"""pm_gather_radiance_noalloc: the allocation-free radiance gather equals
pm_gather_radiance bit for bit.

Called through the C ABI twice in a row on one map, the second call with a
smaller query set, so the second call runs over scratch contents the first
one left behind.  The first call meets a map whose radius bounds for its k are
not cached (they are computed into the fixed scratch array); the second one
reads the bounds an allocating gather cached.  Also: the query limit, an
empty map, and k >= N.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(pmap, q, mats, k, n_emit):
    import torch

    from photon_mapping_amd import _lib

    m = int(q.shape[0])
    rgb = torch.full((m, 3), float("nan"), dtype=torch.float32, device=q.device)
    r2 = torch.full((m,), float("nan"), dtype=torch.float32, device=q.device)
    key = torch.full((m,), -1, dtype=torch.int64, device=q.device)
    stream = _lib.C.c_void_p(torch.cuda.current_stream(q.device).cuda_stream)
    _lib.call("pm_gather_radiance_noalloc", pmap.handle, m, _lib.ptr(q), _lib.ptr(mats),
              int(mats.shape[0]), int(k), float(n_emit), _lib.ptr(rgb), _lib.ptr(r2),
              _lib.ptr(key), stream)
    return rgb, r2, key


def _same(a, b):
    import torch

    for x, y in zip(a, b):
        assert torch.equal(x.view(torch.int32) if x.dtype == torch.float32 else x,
                           y.view(torch.int32) if y.dtype == torch.float32 else y)


@pytest.fixture(scope="module")
def scene_data(gpu):
    import torch

    from photon_mapping_amd import scene as sc

    s = sc.box_scene()
    with sc.NativeScene(s) as ns:
        pos, wi, flux, n_emit = sc.trace_photons(ns, 300_000, seed=3)
        q = sc.camera_queries(ns, 256, 256, 8, depth=1, seed=3)
    mats = torch.from_numpy(s.materials()).cuda()
    torch.cuda.synchronize()
    return pos, wi, flux, float(n_emit), q, mats


def test_noalloc_equals_gather_radiance(gpu, scene_data):
    import torch

    pos, wi, flux, n_emit, q, mats = scene_data
    k = 100
    pmap = gpu.PhotonMap()
    pmap.set_photons_soa(pos, wi, flux)
    pmap.build()
    print("noalloc: first call, bounds not cached", flush=True)
    first = _noalloc(pmap, q, mats, k, n_emit)            # 524,288 queries
    torch.cuda.synchronize()
    expect = pmap.gather_radiance(q, mats, k, n_emit)     # allocating; caches the bounds
    torch.cuda.synchronize()
    _same(first, expect)
    print("noalloc: second call, smaller, bounds cached", flush=True)
    q2 = q[100_000:170_000].contiguous()
    second = _noalloc(pmap, q2, mats, k, n_emit)
    torch.cuda.synchronize()
    _same(second, pmap.gather_radiance(q2, mats, k, n_emit))
    # another k: bounds not cached again, smaller still
    q3 = q[:5000].contiguous()
    third = _noalloc(pmap, q3, mats, 32, n_emit)
    torch.cuda.synchronize()
    _same(third, pmap.gather_radiance(q3, mats, 32, n_emit))


def test_noalloc_limits_and_edges(gpu, scene_data):
    import torch

    from photon_mapping_amd import _lib

    pos, wi, flux, n_emit, q, mats = scene_data
    small = gpu.PhotonMap()
    small.set_photons_soa(pos[:60], wi[:60], flux[:60])
    small.build()
    print("noalloc: k >= N", flush=True)
    got = _noalloc(small, q[:3000].contiguous(), mats, 100, n_emit)
    torch.cuda.synchronize()
    _same(got, small.gather_radiance(q[:3000].contiguous(), mats, 100, n_emit))
    empty = gpu.PhotonMap()
    z = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
    empty.set_photons_soa(z, z, z)
    empty.build()
    print("noalloc: empty map", flush=True)
    rgb, r2, key = _noalloc(empty, q[:500].contiguous(), mats, 10, 1.0)
    torch.cuda.synchronize()
    assert not rgb.any() and not r2.any() and not key.any()
    big = torch.zeros((_lib.PM_NOALLOC_MAX_QUERIES + 1, 10), dtype=torch.float32,
                      device="cuda")
    with pytest.raises(gpu.PmError) as e:
        _noalloc(small, big, mats, 8, 1.0)
    assert e.value.code == _lib.PM_ERR_UNSUPPORTED
