"""Caustic photon tracing and the reference's count semantics
(pmx_trace_photons kind CAUSTICS, pmx_trace_photon_map;
include/integrator.h:393-563).

Pins: with a flat mirror floor of albedo 1 in the closed box, a caustic path
(light -> mirror -> first Lambert hit) keeps its initial flux exactly, since
f cos / pdf = rho |cos_o| |cos_i| / (|cos_o| |cos_i|) = 1, so every caustic
photon carries T0 = Le * A_light * pi and none lies on the mirror; a path
whose first hit is Lambert stores nothing.  The map variant emits exactly
n_paths paths and must hold the photons the fixed-count variant writes for
the same paths, in the same order."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mirror_floor_box():
    from photon_mapping_amd import _lib
    from photon_mapping_amd import scene as sc

    s = sc.box_scene()
    front = sc._quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, -1), 0)
    mirror = len(s.albedo)
    s.tris["material"][:2] = mirror                        # the floor quad
    s.tris = np.concatenate([s.tris, np.array(front, dtype=s.tris.dtype)])
    s.albedo = np.concatenate([s.albedo, np.ones((1, 3), np.float32)])
    s.mat_types = np.concatenate([s.mat_types, np.array([_lib.PM_MATERIAL_MIRROR], np.int32)])
    return s


def test_caustics_through_a_flat_mirror_keep_their_flux(gpu):
    from photon_mapping_amd import scene as sc

    s = _mirror_floor_box()
    print("caustics mirror floor", flush=True)
    pos, wi, flux, n_emit = sc.trace_photons(s, 100_000, seed=2, kind="caustics")
    t0 = np.asarray(s.light_le, np.float64) * s.light_area * math.pi
    f = flux.double().cpu().numpy()
    np.testing.assert_allclose(f, np.broadcast_to(t0, f.shape), rtol=2e-5)
    # every caustic photon left the mirror floor upwards: wi = -d points down
    # (a photon stored on the floor itself would have arrived from above)
    w = wi.cpu().numpy()
    assert (w[:, 1] < 0).all()
    # at most the small light's view factor to the 1 x 1 floor one unit below
    # (0.2394, the inner boxes shadow part of it)
    x = 0.5
    corner = (x / math.sqrt(1 + x * x) * math.atan(x / math.sqrt(1 + x * x))) / math.pi
    assert 0.05 < 100_000 / n_emit < 4 * corner * 1.02


def test_global_photons_skip_specular_surfaces(gpu):
    from photon_mapping_amd import scene as sc

    s = sc.caustics_scene()
    pos, wi, flux, n_emit = sc.trace_photons(s, 200_000, seed=1)
    p = pos.double().cpu().numpy()
    for c in ((0.3, 0.18, 0.4), (0.7, 0.18, 0.6)):
        # the sphere's facets lie within 2e-4 of radius 0.18; it touches the floor
        r = np.linalg.norm(p - np.asarray(c), axis=1)
        on_sphere = (np.abs(r - 0.18) < 2e-3) & (p[:, 1] > 1e-3)
        assert not on_sphere.any()
    assert (flux.cpu().numpy() >= 0).all()


def test_caustics_scene_and_water_store_caustics(gpu):
    from photon_mapping_amd import scene as sc

    for s in (sc.caustics_scene(), sc.water_scene(0)):
        print(f"caustics in a scene of {len(s.tris)} triangles", flush=True)
        pos, wi, flux, n_emit = sc.trace_photons(s, 50_000, seed=3, kind="caustics")
        assert n_emit >= 50_000
        f = flux.cpu().numpy()
        assert np.isfinite(f).all() and (f >= 0).all() and (f.sum(1) > 0).any()


def test_photon_map_emits_exactly_n_paths(gpu):
    from photon_mapping_amd import scene as sc

    s = sc.box_scene()
    n_paths = 5000
    with sc.NativeScene(s) as ns:
        pmap, n_stored = sc.trace_photon_map(ns, n_paths, seed=9)
        assert n_stored == pmap.getNPhotons() > n_paths
        pos, wi, flux, n_emit = sc.trace_photons(ns, n_stored, seed=9)
        assert n_emit <= n_paths
        for i in (0, 1, n_stored // 2, n_stored - 1):
            ph = pmap.getIthPhoton(i)
            assert np.allclose(ph.position, pos[i].cpu().numpy(), rtol=0, atol=0)
            assert np.allclose(ph.wi, wi[i].cpu().numpy(), rtol=0, atol=0)
            assert np.allclose(ph.throughput, flux[i].cpu().numpy(), rtol=0, atol=0)
        pmap.build()
        idx, r2 = pmap.knn(np.array([[0.5, 0.0, 0.5]], np.float32), 10)
        assert (idx >= 0).all() and r2[0] > 0
        empty, n0 = sc.trace_photon_map(ns, 0, seed=9)
        assert n0 == 0 and empty.getNPhotons() == 0
        cm, nc = sc.trace_photon_map(ns, 2000, seed=9, kind="caustics")
        assert nc == 0                                     # no specular surface in the box
