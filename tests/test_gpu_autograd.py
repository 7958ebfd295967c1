"""torch.autograd wrapper (photon_mapping_amd.autograd): gradients with
respect to per-photon flux and per-material albedo equal the FP64 oracle
gradient, in both modes (reused neighbour keys, recomputed kNN), and a
final-gathering albedo factor composes through ordinary torch products."""

import numpy as np
import pytest

from util import materials, photons_surfaces, query_records

pytestmark = pytest.mark.gpu


def _setup(gpu, rng, n=20000, m=3000):
    import torch

    pos, wi, flux, nrm = photons_surfaces(rng, n)
    sel = rng.integers(0, n, m)
    qpos = pos[sel] + rng.normal(scale=0.002, size=(m, 3)).astype(np.float32)
    q = query_records(rng, qpos, nrm[sel])
    mats = materials(3)
    dev = torch.device("cuda:0")
    pmap = gpu.PhotonMap()
    pmap.set_photons_soa(torch.from_numpy(pos).to(dev), torch.from_numpy(wi).to(dev),
                         torch.from_numpy(flux).to(dev))
    pmap.build()
    return pos, wi, flux, q, mats, pmap, dev


@pytest.mark.parametrize("recompute", [False, True])
def test_autograd_matches_fp64_oracle(gpu, oracle, rng, recompute):
    import torch

    from photon_mapping_amd.autograd import gather_radiance

    pos, wi, flux, q, mats, pmap, dev = _setup(gpu, rng)
    k, n_emit = 64, 2e4
    g = rng.normal(size=(len(q), 3)).astype(np.float32)
    print(f"autograd recompute={recompute}", flush=True)
    flux_t = torch.from_numpy(flux).to(dev).requires_grad_(True)
    alb_t = torch.from_numpy(mats[:, :3].copy()).to(dev).requires_grad_(True)
    types = torch.from_numpy(mats.view(np.int32)[:, 3].copy())
    qt = torch.from_numpy(q).to(dev)
    rgb = gather_radiance(pmap, qt, flux_t, alb_t, types, k, n_emit, recompute=recompute)
    (rgb * torch.from_numpy(g).to(dev)).sum().backward()
    bidx, br2 = oracle.brute_knn(pos, q[:, :3], k)
    ef, er = oracle.adjoint_fp64(wi, flux, q, mats, bidx, br2, n_emit, g)
    np.testing.assert_allclose(flux_t.grad.cpu().numpy(), ef.T, rtol=1e-4,
                               atol=1e-5 * np.abs(ef).max())
    np.testing.assert_allclose(alb_t.grad.cpu().numpy(), er, rtol=1e-4,
                               atol=1e-6 * np.abs(er).max())
    ref = oracle.radiance_fp64(wi, flux, q, mats, bidx, br2, n_emit)
    np.testing.assert_allclose(rgb.detach().cpu().numpy(), ref, rtol=1e-5,
                               atol=1e-7 * np.abs(ref).max())


def test_autograd_final_gathering_factor(gpu, oracle, rng):
    """L_fg = rho[fg_mat] * L: the albedo gradient collects both the estimate
    term and the final-gathering factor."""
    import torch

    from photon_mapping_amd.autograd import gather_radiance

    pos, wi, flux, q, mats, pmap, dev = _setup(gpu, rng, n=15000, m=2000)
    k, n_emit = 32, 1e4
    fg = rng.integers(0, 3, len(q)).astype(np.int64)
    g = rng.normal(size=(len(q), 3)).astype(np.float32)
    flux_t = torch.from_numpy(flux).to(dev).requires_grad_(True)
    alb_t = torch.from_numpy(mats[:, :3].copy()).to(dev).requires_grad_(True)
    types = torch.from_numpy(mats.view(np.int32)[:, 3].copy())
    rgb = gather_radiance(pmap, torch.from_numpy(q).to(dev), flux_t, alb_t, types, k, n_emit)
    out = alb_t[torch.from_numpy(fg).to(dev)] * rgb
    (out * torch.from_numpy(g).to(dev)).sum().backward()
    bidx, br2 = oracle.brute_knn(pos, q[:, :3], k)
    rho_fg = mats[fg, :3]
    ef, er = oracle.adjoint_fp64(wi, flux, q, mats, bidx, br2, n_emit, g * rho_fg)
    L = oracle.radiance_fp64(wi, flux, q, mats, bidx, br2, n_emit)
    for mi in range(3):
        er[mi] += np.sum((g * L)[fg == mi], axis=0)
    np.testing.assert_allclose(flux_t.grad.cpu().numpy(), ef.T, rtol=1e-4,
                               atol=1e-5 * np.abs(ef).max())
    np.testing.assert_allclose(alb_t.grad.cpu().numpy(), er, rtol=1e-4,
                               atol=1e-6 * np.abs(er).max())


def test_autograd_fits_albedo(gpu, rng):
    """A few gradient steps recover a perturbed albedo from a target image."""
    import torch

    from photon_mapping_amd.autograd import gather_radiance

    pos, wi, flux, q, mats, pmap, dev = _setup(gpu, rng, n=20000, m=4000)
    k, n_emit = 50, 2e4
    types = torch.from_numpy(mats.view(np.int32)[:, 3].copy())
    qt = torch.from_numpy(q).to(dev)
    flux_t = torch.from_numpy(flux).to(dev)
    true_alb = torch.from_numpy(mats[:, :3].copy()).to(dev)
    target = gather_radiance(pmap, qt, flux_t, true_alb, types, k, n_emit).detach()
    alb = (true_alb * 0.6).clone().requires_grad_(True)
    opt = torch.optim.Adam([alb], lr=0.05)
    first = None
    for _ in range(60):
        opt.zero_grad()
        loss = ((gather_radiance(pmap, qt, flux_t, alb, types, k, n_emit) - target) ** 2).mean()
        loss.backward()
        opt.step()
        first = first if first is not None else float(loss.detach())
    assert float(loss.detach()) < 1e-3 * first
    assert torch.allclose(alb.detach(), true_alb, rtol=2e-2, atol=2e-3)


def test_autograd_two_forwards_before_backward(gpu, oracle, rng):
    """Two forward passes on one map with different flux, then both backward
    passes: each albedo gradient uses its own forward flux, and the map
    keeps the newer flux afterwards."""
    import torch

    from photon_mapping_amd.autograd import gather_radiance

    pos, wi, flux, q, mats, pmap, dev = _setup(gpu, rng, n=8000, m=1000)
    k, n_emit = 32, 1e4
    g = rng.normal(size=(len(q), 3)).astype(np.float32)
    types = torch.from_numpy(mats.view(np.int32)[:, 3].copy())
    qt = torch.from_numpy(q).to(dev)
    fluxes = [flux, (flux * np.float32(2.5) + np.float32(0.1)).astype(np.float32)]
    albs, outs = [], []
    for f in fluxes:
        alb = torch.from_numpy(mats[:, :3].copy()).to(dev).requires_grad_(True)
        out = gather_radiance(pmap, qt, torch.from_numpy(f).to(dev), alb, types, k, n_emit)
        albs.append(alb)
        outs.append(out)
    for out in outs:
        (out * torch.from_numpy(g).to(dev)).sum().backward()
    bidx, br2 = oracle.brute_knn(pos, q[:, :3], k)
    for f, alb in zip(fluxes, albs):
        _, er = oracle.adjoint_fp64(wi, f, q, mats, bidx, br2, n_emit, g)
        np.testing.assert_allclose(alb.grad.cpu().numpy(), er, rtol=1e-4,
                                   atol=1e-6 * np.abs(er).max())
    # the map still holds the newer flux
    rgb, _, _ = pmap.gather_radiance(qt, torch.from_numpy(mats).to(dev), k, n_emit)
    np.testing.assert_array_equal(rgb.cpu().numpy(), outs[1].detach().cpu().numpy())
