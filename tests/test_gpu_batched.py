"""pm_gather_batched: many instances (different maps, sizes and k, a shared
map with two k, an empty map, an empty query list) in one persistent launch
equal the per-instance pm_gather_radiance results bit for bit."""

import numpy as np
import pytest

from util import materials, photons_clustered, photons_surfaces, queries_for, query_records

pytestmark = pytest.mark.gpu


def _instance_data(rng, n, m, gen):
    if gen == "surfaces":
        pos, wi, flux, nrm = photons_surfaces(rng, n)
        pick = rng.integers(0, n, m)
        qpos = (pos[pick] + nrm[pick] * 1e-3).astype(np.float32)
        q = query_records(rng, qpos, nrm[pick])
    else:
        pos, wi, flux = photons_clustered(rng, n)
        q = query_records(rng, queries_for(rng, pos, m, jitter=0.005))
    return pos, wi, flux, q


def test_batched_equals_single_calls(gpu, rng):
    import torch

    dev = torch.device("cuda:0")
    mats = torch.from_numpy(materials(3, with_specular=True)).to(dev)
    specs = [(20000, 3000, 32, "surfaces"), (5000, 700, 100, "clustered"),
             (60000, 5000, 200, "surfaces"), (150, 200, 64, "clustered"),
             (40, 100, 100, "surfaces")]
    maps, inst = [], []
    for n, m, k, gen in specs:
        pos, wi, flux, q = _instance_data(rng, n, m, gen)
        pmap = gpu.PhotonMap()
        pmap.set_photons_soa(torch.from_numpy(pos).to(dev), torch.from_numpy(wi).to(dev),
                             torch.from_numpy(flux).to(dev))
        pmap.build()
        maps.append(pmap)
        inst.append((pmap, torch.from_numpy(q).to(dev), mats, k, 1000.0 + n))
    # the first map again with another k, and an empty query list
    q0 = inst[0][1][:900].contiguous()
    inst.append((maps[0], q0, mats, 100, 5000.0))
    inst.append((maps[1], inst[1][1][:0], mats, 10, 1.0))
    empty = gpu.PhotonMap()
    empty.set_photons_soa(torch.zeros((0, 3), device=dev), torch.zeros((0, 3), device=dev),
                          torch.zeros((0, 3), device=dev))
    empty.build()
    inst.append((empty, inst[3][1], mats, 8, 1.0))
    print(f"batched: {len(inst)} instances", flush=True)
    got = gpu.gather_batched(inst)
    torch.cuda.synchronize()
    for (pmap, q, mt, k, n_emit), (rgb, r2, key) in zip(inst, got):
        if q.shape[0] == 0:
            assert rgb.shape[0] == 0
            continue
        e_rgb, e_r2, e_key = pmap.gather_radiance(q, mt, k, n_emit)
        torch.cuda.synchronize()
        assert torch.equal(rgb.view(torch.int32), e_rgb.view(torch.int32))
        assert torch.equal(r2.view(torch.int32), e_r2.view(torch.int32))
        assert torch.equal(key, e_key)


def test_batched_rejects_foreign_context(gpu, rng):
    import torch

    dev = torch.device("cuda:0")
    pos, wi, flux, q = _instance_data(rng, 500, 64, "surfaces")
    other = gpu.Context(0)
    pmap = gpu.PhotonMap(other)
    pmap.set_photons_soa(pos, wi, flux)
    pmap.build()
    mats = torch.from_numpy(materials(3)).to(dev)
    with pytest.raises(gpu.PmError):
        gpu.gather_batched([(pmap, torch.from_numpy(q).to(dev), mats, 8, 1.0)],
                           ctx=gpu.Context.default())


def _plan_instances(gpu, rng, dev, mats):
    import torch

    inst = []
    for n, m, k, gen in [(30000, 4000, 64, "surfaces"), (8000, 1500, 100, "clustered"),
                         (3000, 900, 200, "surfaces")]:
        pos, wi, flux, q = _instance_data(rng, n, m, gen)
        pmap = gpu.PhotonMap()
        pmap.set_photons_soa(torch.from_numpy(pos).to(dev), torch.from_numpy(wi).to(dev),
                             torch.from_numpy(flux).to(dev))
        pmap.build()
        inst.append((pmap, torch.from_numpy(q).to(dev), mats, k, 500.0 + n))
    return inst


def test_plan_launch_equals_batched(gpu, rng):
    import torch

    dev = torch.device("cuda:0")
    mats = torch.from_numpy(materials(3, with_specular=True)).to(dev)
    inst = _plan_instances(gpu, rng, dev, mats)
    print("plan launch vs batched", flush=True)
    plan = gpu.BatchedPlan(inst)
    expect = gpu.gather_batched(inst)
    for _ in range(2):
        plan.launch()
        torch.cuda.synchronize()
        for (rgb, r2, key), (e_rgb, e_r2, e_key) in zip(plan.outputs, expect):
            assert torch.equal(rgb.view(torch.int32), e_rgb.view(torch.int32))
            assert torch.equal(r2.view(torch.int32), e_r2.view(torch.int32))
            assert torch.equal(key, e_key)
    plan.close()


def test_plan_graph_replays(gpu, rng):
    """One frame captured in a HIP graph and replayed four times, the photon
    flux changed between replays (eager work in between); every replay equals
    the eager batched gather of that frame bit for bit."""
    import torch

    dev = torch.device("cuda:0")
    mats = torch.from_numpy(materials(3)).to(dev)
    inst = _plan_instances(gpu, rng, dev, mats)
    base_flux = [torch.rand((p.getNPhotons(), 3), device=dev) + 0.1 for p, *_ in inst]
    plan = gpu.BatchedPlan(inst)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.launch()  # warm-up outside the capture
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        plan.launch()
    torch.cuda.synchronize()
    for frame in range(4):
        print(f"graph replay {frame}", flush=True)
        for (pmap, *_), f in zip(inst, base_flux):
            pmap.update_flux(f * (1.0 + frame))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        expect = gpu.gather_batched(inst)
        torch.cuda.synchronize()
        for (rgb, r2, key), (e_rgb, e_r2, e_key) in zip(plan.outputs, expect):
            assert torch.equal(rgb.view(torch.int32), e_rgb.view(torch.int32))
            assert torch.equal(r2.view(torch.int32), e_r2.view(torch.int32))
            assert torch.equal(key, e_key)
    del graph
    plan.close()


def test_batched_many_k_on_one_map(gpu, rng):
    """Five distinct k on one map in one batch (more than the map's four
    cached radius-bound sets): no set an earlier instance uses is replaced
    before the launch, so no group falls back to an unbounded search, and
    every instance equals its single call."""
    import torch

    from photon_mapping_amd import _lib

    dev = torch.device("cuda:0")
    mats = torch.from_numpy(materials(3)).to(dev)
    pos, wi, flux, q = _instance_data(rng, 40000, 3000, "surfaces")
    pmap = gpu.PhotonMap()
    pmap.set_photons_soa(torch.from_numpy(pos).to(dev), torch.from_numpy(wi).to(dev),
                         torch.from_numpy(flux).to(dev))
    pmap.build()
    qd = torch.from_numpy(q).to(dev)
    # descending: a later, smaller k evicting an earlier set would leave that
    # instance with bounds too tight for its k
    inst = [(pmap, qd, mats, k, 1e4) for k in (150, 100, 64, 32, 16, 8)]
    ctx = pmap.ctx
    print("batched: six k on one map", flush=True)
    ctx.set_option(_lib.PM_OPT_COUNT_STATS, 1)
    try:
        got = gpu.gather_batched(inst)
        torch.cuda.synchronize()
        st = ctx.stats()
    finally:
        ctx.set_option(_lib.PM_OPT_COUNT_STATS, 0)
    assert st["fallbacks"] == 0, st
    for (pm_, q_, mt, k, ne), (rgb, r2, key) in zip(inst, got):
        e_rgb, e_r2, e_key = pm_.gather_radiance(q_, mt, k, ne)
        torch.cuda.synchronize()
        assert torch.equal(rgb.view(torch.int32), e_rgb.view(torch.int32)), k
        assert torch.equal(key, e_key), k


def test_batched_sort_tiles_at_segment_edges(gpu, rng):
    """Instances whose query counts sit on and around the sort's 4,096-key
    tiles (1, 4,095, 4,096, 4,097, 8,193 and 12,288 queries) and an empty one:
    each instance is sorted inside its own range, so the batched results still
    equal the single calls bit for bit."""
    import torch

    dev = torch.device("cuda:0")
    mats = torch.from_numpy(materials(3)).to(dev)
    pos, wi, flux, _ = _instance_data(rng, 30000, 1, "surfaces")
    pmap = gpu.PhotonMap()
    pmap.set_photons_soa(pos, wi, flux)
    pmap.build()
    inst = []
    for m in (4095, 1, 4096, 0, 4097, 8193, 12288):
        pick = rng.integers(0, len(pos), m)
        q = query_records(rng, (pos[pick] + np.float32(1e-3)).astype(np.float32))
        inst.append((pmap, torch.from_numpy(q).to(dev), mats, 16 + (m % 7), 1000.0))
    got = gpu.gather_batched(inst)
    torch.cuda.synchronize()
    for (pm_, q, mt, k, n_emit), (rgb, r2, key) in zip(inst, got):
        if q.shape[0] == 0:
            assert rgb.shape[0] == 0
            continue
        e_rgb, e_r2, e_key = pm_.gather_radiance(q, mt, k, n_emit)
        torch.cuda.synchronize()
        assert torch.equal(rgb.view(torch.int32), e_rgb.view(torch.int32))
        assert torch.equal(r2.view(torch.int32), e_r2.view(torch.int32))
        assert torch.equal(key, e_key)
