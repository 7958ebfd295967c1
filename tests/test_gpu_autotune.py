"""pm_autotune: picks a setting from its grid, keeps it in force, and the
gather at the tuned setting still matches the FP64 oracle."""

import numpy as np
import pytest

from util import materials, photons_surfaces, query_records

pytestmark = pytest.mark.gpu


def test_autotune_then_parity(gpu, oracle, rng):
    import torch

    pos, wi, flux, nrm = photons_surfaces(rng, 40000)
    sel = rng.integers(0, len(pos), 20000)
    q = query_records(rng, pos[sel] + nrm[sel] * 1e-3, nrm[sel])
    mats = materials(3)
    dev = torch.device("cuda:0")
    ctx = gpu.Context(0)  # a context of its own: tuning changes its options
    pmap = gpu.PhotonMap(ctx)
    pmap.set_photons_soa(torch.from_numpy(pos).to(dev), torch.from_numpy(wi).to(dev),
                         torch.from_numpy(flux).to(dev))
    pmap.build()
    qt = torch.from_numpy(q).to(dev)
    print("autotune", flush=True)
    rep = gpu.autotune(pmap, qt, torch.from_numpy(mats).to(dev), 50, 1e4)
    assert rep["leaf_size"] in (8, 16, 32)
    assert rep["waves_per_cu"] in (8, 12, 16, 20)
    assert 0.5 <= rep["run_split"] <= 1.25
    assert rep["n_tried"] >= 10
    assert 0 < rep["best_ms"] <= rep["initial_ms"]
    rgb, r2, _ = pmap.gather_radiance(qt, torch.from_numpy(mats).to(dev), 50, 1e4)
    bidx, br2 = oracle.brute_knn(pos, q[:, :3], 50)
    np.testing.assert_array_equal(r2.cpu().numpy().view(np.uint32), br2.view(np.uint32))
    ref = oracle.radiance_fp64(wi, flux, q, mats, bidx, br2, 1e4)
    np.testing.assert_allclose(rgb.cpu().numpy(), ref, rtol=1e-5, atol=1e-7 * np.abs(ref).max())
