"""torch.autograd wrapper of the photon-density radiance estimate.

``gather_radiance(pmap, queries, flux, albedo, types, k, n_emit)`` returns
the estimate of ``include/integrator.h:133-152`` at every query as an (m, 3)
tensor that is differentiable with respect to the per-photon flux (n, 3) and
the per-material albedo (n_mats, 3).  The forward pass is
``pm_gather_radiance`` on the map (after ``pm_photonmap_update_flux`` with the
given flux), the backward pass ``pm_gather_adjoint``, which either reuses
the forward pass's neighbour-set keys or recomputes the kNN (checkpointing).
Positions, normals and directions are not differentiated (the neighbour sets
are piecewise constant in them).  A final-gathering bounce is an ordinary
torch product outside: ``albedo[fg_mat] * gather_radiance(...)``, whose
albedo gradient torch forms itself.

The map holds the flux of the last forward pass.  When it has been replaced
since (another forward pass on the same map, an update_flux call), the
backward pass uploads the flux its forward pass used, runs the adjoint, and
puts the newer flux back, so every backward pass sees its own forward flux.
"""

from __future__ import annotations

import torch

from .photon_map import PhotonMap


def _materials(albedo: torch.Tensor, types: torch.Tensor) -> torch.Tensor:
    """(n_mats, 4) float32 records: rho rgb and the type as int32 bits."""
    rec = torch.empty((albedo.shape[0], 4), dtype=torch.float32, device=albedo.device)
    rec[:, :3] = albedo.detach().to(torch.float32)
    rec.view(torch.int32)[:, 3] = types.to(device=albedo.device, dtype=torch.int32)
    return rec


class _GatherRadiance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, flux, albedo, pmap: PhotonMap, queries, types, k: int, n_emit: float,
                recompute: bool):
        fl = flux.detach().to(torch.float32).contiguous()
        pmap.update_flux(fl)
        mats = _materials(albedo, types)
        rgb, _, key = pmap.gather_radiance(queries, mats, k, n_emit, want_key=not recompute)
        ctx.pmap = pmap
        ctx.k = k
        ctx.n_emit = n_emit
        ctx.flux_version = pmap.flux_version
        ctx.save_for_backward(queries, mats, key if key is not None else torch.empty(0), fl)
        ctx.recompute = recompute
        return rgb

    @staticmethod
    def backward(ctx, g_rgb):
        queries, mats, key, fl = ctx.saved_tensors
        pmap = ctx.pmap
        newer = None
        if pmap.flux_version != ctx.flux_version:
            newer = pmap._flux_ref
            pmap.update_flux(fl)
        dflux, dalb = pmap.gather_adjoint(
            queries, mats, ctx.k, ctx.n_emit, g_rgb.to(torch.float32).contiguous(),
            key=None if ctx.recompute else key)
        if newer is not None:
            pmap.update_flux(newer)
        return dflux.t(), dalb, None, None, None, None, None, None


def gather_radiance(pmap: PhotonMap, queries: torch.Tensor, flux: torch.Tensor,
                    albedo: torch.Tensor, types: torch.Tensor, k: int, n_emit: float,
                    recompute: bool = False) -> torch.Tensor:
    """Differentiable radiance estimate (m, 3) at the query records (m, 10)."""
    return _GatherRadiance.apply(flux, albedo, pmap, queries, types, int(k), float(n_emit),
                                 bool(recompute))
