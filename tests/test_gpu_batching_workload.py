"""The batching configuration at full size, checked against the oracle.

64 box-scene instances of uneven size drawn with seed 0 exactly as the
benchmark draws them (photons traced from the light, log-uniform in
[1e4, 4e6] stored photons, k from {32, 64, 100, 200}, square resolutions from
{128, 256, 512, 1024}, 16 spp final-gathering queries, 0-3 mirror or glass
spheres): 48.5M photons and 438M queries in one pm_gather_batched launch
(work counters on, so unbounded fallbacks are reported).  Every instance is
checked on a seeded sample of 256 of its queries against the oracle's exact
brute-force kNN under the reference's (FP32 d^2, index) order
(include/photon_map.h:107-183) and the FP64 estimate
(include/integrator.h:133-152): r^2 bit for bit, the key's index word equal
to the k-th neighbour, radiance within 1e-5 relative.  The same launch
prepared as a plan and replayed from a HIP graph gives the same bits.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-5
PER_INSTANCE = 256


def test_batching_workload_matches_oracle(gpu, oracle):
    import torch

    from photon_mapping_amd import _lib
    from photon_mapping_amd import scene as sc

    print("batching workload: generating 64 instances (seed 0)", flush=True)
    work = sc.batching_workload(64, 0, keep_photons=True)
    inst = [(w["map"], w["queries"], w["materials"], w["k"], w["n_emit"]) for w in work]
    assert sorted({w["k"] for w in work}) == [32, 64, 100, 200]
    total_q = sum(int(w["queries"].shape[0]) for w in work)
    print(f"batched launch: {len(inst)} instances, {total_q} queries, "
          f"{sum(w['n'] for w in work)} photons", flush=True)
    ctx = gpu.Context.default()
    ctx.set_option(_lib.PM_OPT_COUNT_STATS, 1)
    try:
        outs = gpu.gather_batched(inst, ctx=ctx)
        torch.cuda.synchronize()
        st = ctx.stats()
    finally:
        ctx.set_option(_lib.PM_OPT_COUNT_STATS, 0)
    assert st["fallbacks"] == 0, st
    rng = np.random.default_rng(0)
    for i, (w, (rgb, r2, key)) in enumerate(zip(work, outs)):
        m = int(w["queries"].shape[0])
        sel = np.sort(rng.choice(m, min(PER_INSTANCE, m), replace=False))
        st_ = torch.from_numpy(sel).cuda()
        q = w["queries"][st_].cpu().numpy()
        g_rgb = rgb[st_].cpu().numpy()
        g_r2 = r2[st_].cpu().numpy()
        g_key = key[st_].cpu().numpy().view(np.uint64)
        pos, wi, flux = w["photons"]
        mats = w["materials"].cpu().numpy()
        bidx, br2 = oracle.brute_knn(pos, q[:, :3], w["k"])
        label = f"instance {i}: n={w['n']} k={w['k']} res={w['res']}"
        np.testing.assert_array_equal(g_r2.view(np.uint32), br2.view(np.uint32), label)
        np.testing.assert_array_equal((g_key & np.uint64(0xffffffff)).astype(np.int64),
                                      bidx[:, 0], label)
        expect = oracle.radiance_fp64(wi, flux, q, mats, bidx, br2, w["n_emit"])
        zero = expect == 0
        assert (g_rgb[zero] == 0).all(), label
        np.testing.assert_allclose(g_rgb[~zero].astype(np.float64), expect[~zero], rtol=RTOL,
                                   atol=0, err_msg=label)
    print("all 64 instances match the oracle on their samples", flush=True)

    # the same batch prepared once and replayed from a HIP graph: same bits
    plan = gpu.BatchedPlan(inst, ctx=ctx)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.launch()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        plan.launch()
    for rep in range(2):
        graph.replay()
        torch.cuda.synchronize()
        for (a_rgb, a_r2, a_key), (b_rgb, b_r2, b_key) in zip(plan.outputs, outs):
            assert torch.equal(a_rgb.view(torch.int32), b_rgb.view(torch.int32)), rep
            assert torch.equal(a_r2.view(torch.int32), b_r2.view(torch.int32)), rep
            assert torch.equal(a_key, b_key), rep
    del graph
    plan.close()
