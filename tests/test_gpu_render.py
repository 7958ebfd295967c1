"""Device renders (pmx_render_photon_mapping, pmx_render_path_tracing)
against the CPU restatement of the reference's integrators and render loop
(oracle/render_oracle.cpp), and against closed forms.

Both sides estimate the same pixel expectations with independent random
numbers; the photon-mapping renders use the same photon arrays (traced once
on the device and handed to both), so their estimates differ only by eye-path
noise.  Pixel means are compared within their Monte Carlo confidence
intervals: |m_gpu - m_cpu| / sqrt(var_gpu + var_cpu) exceeds 2.576 (the 99%
two-sided quantile) for about 1% of pixel channels when both estimate the
same thing, so the test allows 3%, no channel beyond 6, and the image means
within 4 standard errors."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cam_array(cam):
    return np.concatenate([np.asarray(cam[k], np.float32).ravel()
                           for k in ("position", "forward", "right", "up")]
                          + [np.array([cam["focal"]], np.float32)])


def _ci_check(mg, vg, mc, vc, label):
    mg, vg, mc, vc = (np.asarray(a, np.float64) for a in (mg, vg, mc, vc))
    tot = vg + vc
    both_flat = tot <= 0
    assert np.all(np.abs(mg - mc)[both_flat] <= 1e-6 * (1 + np.abs(mc[both_flat]))), label
    z = np.abs(mg - mc)[~both_flat] / np.sqrt(tot[~both_flat])
    frac = float((z > 2.576).mean()) if z.size else 0.0
    print(f"{label}: {z.size} channels, {frac:.4f} beyond the 99% interval, max |z| "
          f"{z.max() if z.size else 0:.2f}", flush=True)
    assert frac <= 0.03, label
    assert z.max() < 6.0, label
    P = mg.shape[0] * mg.shape[1]
    zi = np.abs(mg.mean((0, 1)) - mc.mean((0, 1))) / np.sqrt(tot.sum((0, 1)) / P ** 2)
    assert (zi < 4.0).all(), (label, zi)


def _specular_box():
    """The box scene with the tall box a mirror and the short box glass
    (IOR 1.5), both of albedo 1: specular splits, caustics, light leaks."""
    from photon_mapping_amd import _lib
    from photon_mapping_amd import scene as sc

    s = sc.box_scene()
    mirror, glass = len(s.albedo), len(s.albedo) + 1
    s.tris["material"][12:24] = mirror
    s.tris["material"][24:36] = glass
    s.albedo = np.concatenate([s.albedo, np.ones((2, 3), np.float32)])
    s.mat_types = np.concatenate([s.mat_types, np.array([_lib.PM_MATERIAL_MIRROR,
                                                         _lib.PM_MATERIAL_GLASS], np.int32)])
    s.ior = np.full(len(s.albedo), 1.5, np.float32)
    return s


def _oracle_render(s, mode, w, h, spp, **kw):
    import pm_oracle

    return pm_oracle.render(mode, s.tris, s.materials(), s.ior, s.light_le,
                            _cam_array(s.camera), w, h, spp, mat_le=s.mat_le, **kw)


def _maps(s, n_global, n_caustics):
    """Traced photon arrays (host copies) and built device maps."""
    import photon_mapping_amd as pm
    from photon_mapping_amd import scene as sc

    import torch

    out = {}
    for kind, n in (("global", n_global), ("caustics", n_caustics)):
        if n == 0:                                       # an empty map
            pos = wi = flux = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
            n_emit = 1
        else:
            pos, wi, flux, n_emit = sc.trace_photons(s, n, seed=5, kind=kind)
        pmap = pm.PhotonMap()
        pmap.set_photons_soa(pos, wi, flux)
        pmap.build()
        host = tuple(a.cpu().numpy() for a in (pos, wi, flux))
        out[kind] = (pmap, host, float(n_emit))
    return out


@pytest.mark.parametrize("name", ["box", "specular"])
def test_path_tracing_matches_oracle(gpu, name):
    from photon_mapping_amd import scene as sc

    s = sc.box_scene() if name == "box" else _specular_box()
    w, h, spp = 24, 24, 256
    print(f"path tracing {name}", flush=True)
    mg, vg, st = sc.render_path_tracing(s, w, h, spp, seed=1)
    mc, vc, dropped = _oracle_render(s, "pt", w, h, spp, seed=2)
    assert st["samples"] == w * h * spp and st["dropped_samples"] == 0 and dropped == 0
    _ci_check(mg.cpu().numpy(), vg.cpu().numpy(), mc, vc, f"pt {name}")


def test_path_tracing_furnace(gpu):
    """Closed box, every face emitting Le = 1 with albedo rho: each bounce
    adds rho^k, so the expected radiance is (1 - rho^D) / (1 - rho) for a
    depth limit D (Russian roulette is unbiased)."""
    from photon_mapping_amd import scene as sc

    s = sc.box_scene()
    s.tris = s.tris[:10].copy()                          # walls only
    front = sc._quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, -1), 0)
    s.tris = np.concatenate([s.tris, np.array(front, dtype=s.tris.dtype)])
    s.tris["is_light"] = 1
    s.tris["material"] = 0
    rho, D = 0.5, 100
    s.albedo[:] = rho
    s.light_le = np.array([1.0, 1.0, 1.0])
    s.camera = dict(position=np.array([0.5, 0.5, 0.9]), forward=np.array([0.0, 0.0, -1.0]),
                    right=np.array([0.3, 0.0, 0.0]), up=np.array([0.0, 0.3, 0.0]), focal=1.0)
    mg, vg, st = sc.render_path_tracing(s, 16, 16, 512, max_depth=D, seed=3)
    mg, vg = mg.cpu().numpy().astype(np.float64), vg.cpu().numpy().astype(np.float64)
    expect = (1 - rho ** D) / (1 - rho)
    z = np.abs(mg - expect) / np.sqrt(vg)
    assert float((z > 2.576).mean()) <= 0.03 and z.max() < 6.0
    se = math.sqrt(vg.sum() / vg.size ** 2 * 3)
    assert abs(mg.mean() - expect) < 4 * se


@pytest.mark.parametrize("name,fg", [("box", 0), ("box", 1), ("specular", 1)])
def test_photon_mapping_matches_oracle(gpu, name, fg):
    from photon_mapping_amd import scene as sc

    s = sc.box_scene() if name == "box" else _specular_box()
    maps = _maps(s, 50_000, 0 if name == "box" else 20_000)
    gmap, ghost, gemit = maps["global"]
    cmap, chost, cemit = maps["caustics"]
    w, h, spp, k = 24, 24, 32, 50
    print(f"photon mapping {name} fg={fg}", flush=True)
    mg, vg, st = sc.render_photon_mapping(s, gmap, gemit, w, h, spp, k=k, caustics_map=cmap,
                                          n_emit_caustics=cemit, final_gathering_depth=fg,
                                          seed=1)
    mc, vc, dropped = _oracle_render(s, "pm", w, h, spp, fg_depth=fg, k=k, global_map=ghost,
                                     n_emit_global=gemit, caustics_map=chost,
                                     n_emit_caustics=cemit, seed=2)
    assert st["samples"] == w * h * spp
    assert st["dropped_samples"] == dropped == 0
    assert st["gathers_global"] > 0
    if name == "box":  # an empty caustics map estimates 0: no lookups are made
        assert st["gathers_caustics"] == 0
    _ci_check(mg.cpu().numpy(), vg.cpu().numpy(), mc, vc, f"pm {name} fg={fg}")


@pytest.mark.parametrize("name", ["caustics", "water"])
def test_photon_mapping_config_scenes_match_oracle(gpu, name):
    """The caustics configuration (two 5120-triangle icospheres, a mirror and
    IOR-1.5 glass) and the differentiation configuration (the 131,072-triangle
    water height field, IOR 1.33, 64 wall patches) rendered with final
    gathering, a global and a caustics map and k = 100, at a reduced size:
    the device frame against the CPU restatement within the 99% intervals."""
    from photon_mapping_amd import scene as sc

    s = sc.caustics_scene() if name == "caustics" else sc.water_scene(0)
    maps = _maps(s, 60_000, 30_000)
    gmap, ghost, gemit = maps["global"]
    cmap, chost, cemit = maps["caustics"]
    w, h, spp, k = 24, 24, 48, 100
    print(f"photon mapping {name} fg=1 k={k}", flush=True)
    mg, vg, st = sc.render_photon_mapping(s, gmap, gemit, w, h, spp, k=k, caustics_map=cmap,
                                          n_emit_caustics=cemit, final_gathering_depth=1,
                                          seed=1)
    mc, vc, dropped = _oracle_render(s, "pm", w, h, spp, fg_depth=1, k=k, global_map=ghost,
                                     n_emit_global=gemit, caustics_map=chost,
                                     n_emit_caustics=cemit, seed=2)
    assert st["samples"] == w * h * spp
    assert st["gathers_caustics"] > 0 and st["gathers_global"] > 0
    assert st["dropped_samples"] == dropped == 0
    _ci_check(mg.cpu().numpy(), vg.cpu().numpy(), mc, vc, f"pm {name} fg=1")


def test_plain_configuration_full_size_matches_oracle(gpu):
    """The CPU-runnable plain configuration at its full size: the box scene,
    100,000 emitted photon paths (every stored photon kept, the reference's
    count semantics), k = 100, 256 x 256 x 16 spp, no final gathering."""
    from photon_mapping_amd import scene as sc

    import torch

    s = sc.box_scene()
    with sc.NativeScene(s) as ns:
        gmap, n_stored = sc.trace_photon_map(ns, 100_000, seed=0)
        gmap.build()
        pos, wi, flux, _ = sc.trace_photons(ns, n_stored, seed=0)
        host = tuple(a.cpu().numpy() for a in (pos, wi, flux))
        print(f"plain configuration: {n_stored} photons from 100000 paths", flush=True)
        mg, vg, st = sc.render_photon_mapping(ns, gmap, 100_000, 256, 256, 16, k=100, seed=1)
        torch.cuda.synchronize()
    mc, vc, dropped = _oracle_render(s, "pm", 256, 256, 16, fg_depth=0, k=100,
                                     global_map=host, n_emit_global=100_000, seed=2)
    assert st["samples"] == 256 * 256 * 16 and st["dropped_samples"] == dropped == 0
    _ci_check(mg.cpu().numpy(), vg.cpu().numpy(), mc, vc, "plain configuration 256x256x16")


def test_render_is_deterministic(gpu):
    from photon_mapping_amd import scene as sc

    s = _specular_box()
    maps = _maps(s, 30_000, 10_000)
    gmap, _, gemit = maps["global"]
    cmap, _, cemit = maps["caustics"]
    kw = dict(k=32, caustics_map=cmap, n_emit_caustics=cemit, final_gathering_depth=1, seed=4)
    a, va, _ = sc.render_photon_mapping(s, gmap, gemit, 20, 18, 8, **kw)
    b, vb, _ = sc.render_photon_mapping(s, gmap, gemit, 20, 18, 8, **kw)
    c, vcc, st = sc.render_photon_mapping(s, gmap, gemit, 20, 18, 8, batch_pixels=37, **kw)
    assert st["batches"] == math.ceil(20 * 18 / 37)
    # the same call twice: bit for bit (queries gathered in (sample, lineage)
    # order, sample sums in fixed point)
    assert bool((b == a).all()) and bool((vb == va).all())
    # another batching: the same kNN sets and paths; the estimates' FP32 sums
    # may round differently because a query's neighbours are summed in an order
    # that depends on which queries share its wavefront group
    np.testing.assert_allclose(c.cpu().numpy(), a.cpu().numpy(), rtol=2e-5, atol=1e-7)
    p, _, _ = sc.render_path_tracing(s, 20, 18, 8, seed=4)
    q, _, _ = sc.render_path_tracing(s, 20, 18, 8, seed=4, batch_pixels=7)
    assert bool((p == q).all())


def test_photon_mapping_converges_to_path_tracing(gpu):
    """Final gathering over a dense global map: the image mean comes within a
    few percent of the unbiased path tracer's (the kNN density estimate is
    biased, but only at the second bounce)."""
    from photon_mapping_amd import scene as sc

    s = sc.box_scene()
    maps = _maps(s, 400_000, 0)
    gmap, _, gemit = maps["global"]
    cmap, _, cemit = maps["caustics"]
    pm_img, _, _ = sc.render_photon_mapping(s, gmap, gemit, 48, 48, 64, k=100, caustics_map=cmap,
                                            n_emit_caustics=cemit, final_gathering_depth=1,
                                            seed=7)
    pt_img, pt_var, _ = sc.render_path_tracing(s, 48, 48, 1024, seed=8)
    a = pm_img.double().mean((0, 1)).cpu().numpy()
    b = pt_img.double().mean((0, 1)).cpu().numpy()
    print("pm vs pt image means", a, b, flush=True)
    np.testing.assert_allclose(a, b, rtol=0.05)
