"""Closed-form pins of the oracle's render restatement (oracle/render_oracle.cpp).

The device renders are compared with this restatement within Monte Carlo
intervals, so the restatement itself is pinned here, on the CPU, by
expectations the mathematics fixes:

* furnace: a closed box whose every face emits Le = 1 with albedo rho; a
  path of depth limit D collects sum_{k<D} rho^k = (1 - rho^D) / (1 - rho)
  (Russian roulette is unbiased; include/integrator.h:60-101);
* direct light under a rectangular emitter: a Lambert floor of albedo rho
  below a parallel rectangle of radiance Le at height h reflects
  L = rho Le F, with F the point-to-rectangle form factor in closed form.
  The path tracer reaches it with one bounce (depth limit 2), photon mapping
  with final gathering over empty maps through next-event estimation alone
  (include/integrator.h:178-213), so a wrong light pdf, a dropped cosine or
  a wrong Lambert factor fails it;
* the restated query producer puts its queries on Lambert faces inside the
  box, with unit normals and wo on the front side (depth 0 always; depth 1
  but for rays that meet a face from behind), and never a primary query on
  the light.
"""

import math

import numpy as np
import pytest


def _cam(position, forward, right, up, focal):
    return np.concatenate([np.asarray(v, np.float32).ravel() for v in
                           (position, forward, right, up)] + [np.float32([focal])])


def _ci(mean, var, expect, label):
    mean, var, expect = (np.asarray(a, np.float64) for a in (mean, var, expect))
    z = np.abs(mean - expect) / np.sqrt(np.maximum(var, 1e-30))
    frac = float((z > 2.576).mean())
    print(f"{label}: {z.size} channels, {frac:.4f} beyond the 99% interval, max |z| "
          f"{z.max():.2f}")
    assert frac <= 0.03, label
    assert z.max() < 6.0, label
    se = np.sqrt(var.sum((0, 1))) / (var.shape[0] * var.shape[1])
    zi = np.abs(mean.mean((0, 1)) - expect.mean((0, 1))) / se
    assert (zi < 4.0).all(), (label, zi)


def _furnace(rho):
    from photon_mapping_amd import scene as sc

    s = sc.box_scene()
    tris = list(s.tris[:10])
    tris += sc._quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, -1), 0)
    t = np.array(tris, dtype=sc.TRI_DTYPE)
    t["is_light"] = 1
    t["material"] = 0
    from photon_mapping_amd.photon_map import make_materials

    mats = make_materials(np.full((1, 3), rho, np.float32))
    cam = _cam((0.5, 0.5, 0.9), (0, 0, -1), (0.3, 0, 0), (0, 0.3, 0), 1.0)
    return t, mats, cam


@pytest.mark.parametrize("rho,depth", [(0.5, 100), (0.8, 5)])
def test_render_oracle_path_tracing_furnace(oracle, rho, depth):
    t, mats, cam = _furnace(rho)
    mean, var, dropped = oracle.render("pt", t, mats, None, np.ones(3, np.float32), cam, 12,
                                       12, 512, max_depth=depth, seed=11)
    assert dropped == 0
    expect = (1 - rho ** depth) / (1 - rho)
    _ci(mean, var, np.full_like(mean, expect), f"furnace rho={rho} D={depth}")


def test_render_oracle_furnace_light_hit_is_le(oracle):
    """Photon mapping returns Le at a light hit before anything else
    (include/integrator.h:297-299): every camera ray of the furnace sees
    exactly 1."""
    t, mats, cam = _furnace(0.5)
    z = np.zeros((0, 3), np.float32)
    mean, var, _ = oracle.render("pm", t, mats, None, np.ones(3, np.float32), cam, 6, 6, 4,
                                 fg_depth=1, k=10, global_map=(z, z, z),
                                 caustics_map=(z, z, z))
    assert (mean == 1.0).all() and (var == 0.0).all()


def rect_form_factor(px, pz, h, x0, x1, z0, z1):
    """Form factor from a differential area (normal +y, at (px, 0, pz)) to a
    parallel rectangle [x0, x1] x [z0, z1] at height h facing it: the corner
    formula G(a, b) of the rectangle [0, a] x [0, b], summed with signs."""
    def g(a, b):
        ra = np.sqrt(a * a + h * h)
        rb = np.sqrt(b * b + h * h)
        return (a / ra * np.arctan(b / ra) + b / rb * np.arctan(a / rb)) / (2 * math.pi)

    return (g(x1 - px, z1 - pz) - g(x0 - px, z1 - pz) - g(x1 - px, z0 - pz)
            + g(x0 - px, z0 - pz))


def test_rect_form_factor_against_quadrature():
    """The closed form against midpoint quadrature of cos cos / (pi r^2)."""
    h, x0, x1, z0, z1 = 1.0, -0.2, 0.3, -0.25, 0.15
    n = 400
    xs = x0 + (np.arange(n) + 0.5) * (x1 - x0) / n
    zs = z0 + (np.arange(n) + 0.5) * (z1 - z0) / n
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    dA = (x1 - x0) * (z1 - z0) / n ** 2
    for px, pz in [(0.0, 0.0), (0.4, -0.3), (-0.7, 0.5), (0.05, 0.1)]:
        r2 = (X - px) ** 2 + (Z - pz) ** 2 + h * h
        quad = float(np.sum(h * h / (math.pi * r2 * r2)) * dA)
        assert abs(rect_form_factor(px, pz, h, x0, x1, z0, z1) - quad) < 1e-6 * max(quad, 1e-3)


def _direct_light_scene():
    from photon_mapping_amd import scene as sc
    from photon_mapping_amd.photon_map import make_materials

    x0, x1, z0, z1, h = -0.2, 0.3, -0.25, 0.15, 1.0
    tris = sc._quad((-20, 0, -20), (20, 0, -20), (20, 0, 20), (-20, 0, 20), (0, 1, 0), 0)
    tris += sc._quad((x0, h, z0), (x1, h, z0), (x1, h, z1), (x0, h, z1), (0, -1, 0), 1, 1)
    t = np.array(tris, dtype=sc.TRI_DTYPE)
    rho = np.float32([0.5, 0.3, 0.7])
    mats = make_materials(np.stack([rho, np.zeros(3, np.float32)]))
    le = np.float32([1.0, 2.0, 0.5])
    pos, fwd = np.array([0.05, 0.9, -0.03]), np.array([0.0, -1.0, 0.0])
    right, up, focal = np.array([0.5, 0.0, 0.0]), np.array([0.0, 0.0, 0.5]), 0.5
    cam = _cam(pos, fwd, right, up, focal)
    return t, mats, le, cam, (pos, fwd, right, up, focal), rho, (x0, x1, z0, z1, h)


def _expected_direct(w, h_img, camera, rho, rect, le, sub=8):
    """Pixel means of rho Le F over the floor points the pixel's camera rays
    reach (sub x sub midpoints per pixel)."""
    pos, fwd, right, up, focal = camera
    x0, x1, z0, z1, h = rect
    out = np.zeros((h_img, w, 3))
    off = (np.arange(sub) + 0.5) / sub
    for i in range(h_img):
        for j in range(w):
            a, b = np.meshgrid(off, off, indexing="ij")
            u = (2.0 * (j + a) - w) / h_img
            v = (2.0 * (i + b) - h_img) / h_img
            sensor = pos + u[..., None] * right + v[..., None] * up
            d = pos + focal * fwd - sensor
            d /= np.linalg.norm(d, axis=-1, keepdims=True)
            t = -sensor[..., 1] / d[..., 1]
            p = sensor + t[..., None] * d
            F = rect_form_factor(p[..., 0], p[..., 2], h, x0, x1, z0, z1)
            out[i, j] = F.mean() * rho * le
    return out


@pytest.mark.parametrize("mode", ["pt", "pm"])
def test_render_oracle_direct_light_closed_form(oracle, mode):
    t, mats, le, cam, camera, rho, rect = _direct_light_scene()
    w = h = 12
    expect = _expected_direct(w, h, camera, rho, rect, le)
    z = np.zeros((0, 3), np.float32)
    if mode == "pt":
        mean, var, dropped = oracle.render("pt", t, mats, None, le, cam, w, h, 600, max_depth=2,
                                           seed=5)
    else:
        mean, var, dropped = oracle.render("pm", t, mats, None, le, cam, w, h, 300, fg_depth=1,
                                           k=10, global_map=(z, z, z), caustics_map=(z, z, z),
                                           seed=5)
    assert dropped == 0
    assert expect.min() > 0.01  # every pixel sees the light
    _ci(mean, var, expect, f"direct light {mode}")


def test_oracle_fg_queries_lie_on_lambert_faces(oracle):
    from photon_mapping_amd import scene as sc

    s = sc.box_scene()
    cam = _cam(s.camera["position"], s.camera["forward"], s.camera["right"], s.camera["up"],
               s.camera["focal"])
    for depth in (0, 1):
        q, fg = oracle.fg_queries(s.tris, s.materials(), s.ior, s.light_le, cam, 16, 16, 4, depth,
                                  seed=1)
        mat = q.view(np.int32)[:, 9]
        assert (mat >= 0).all()
        if depth == 0:
            assert (mat != 3).all()   # a primary hit on the light is redrawn
        else:
            # a final-gathering ray may end on the light's Lambert face, where
            # the reference gathers too (include/integrator.h:239-244)
            assert 0 < (mat == 3).mean() < 0.05
        p = q[:, :3]
        assert (p >= -1e-4).all() and (p <= 1 + 1e-4).all()  # inside the box
        # unit shading normals of flat faces, wo on their front side
        np.testing.assert_allclose(np.linalg.norm(q[:, 3:6], axis=1), 1.0, rtol=1e-5)
        # wo on the front side, but for a few final-gathering rays that meet a
        # face from behind (from the ceiling onto the light quad hanging just
        # below it, or from a point on an inner box's edge); the one-sided
        # Lambert makes those estimates 0
        front = np.sum(q[:, 3:6] * q[:, 6:9], axis=1) > 0
        if depth == 0:
            assert front.all()
        else:
            assert front[mat != 3].mean() > 0.99
        if depth == 1:
            assert (fg >= 0).all()
        else:
            assert (fg == -1).all()


@pytest.mark.parametrize("name", ["box", "caustics", "water"])
def test_oracle_bvh_equals_brute_force(oracle, name):
    """The restatement's BVH returns exactly the brute-force loop's hits
    (triangle and t bits): random rays from inside the box, axis-aligned
    rays, and rays aimed at triangle vertices (ties between neighbours)."""
    from photon_mapping_amd import scene as sc

    s = {"box": sc.box_scene, "caustics": sc.caustics_scene,
         "water": lambda: sc.water_scene(0)}[name]()
    rng = np.random.default_rng(4)
    n = 4000
    org = rng.random((n, 3), dtype=np.float32)
    d = rng.normal(size=(n, 3))
    d[: n // 8] = np.eye(3)[rng.integers(0, 3, n // 8)] * rng.choice([-1.0, 1.0], (n // 8, 1))
    vtx = s.tris["v"][rng.integers(0, len(s.tris), n // 8), rng.integers(0, 3, n // 8)]
    d[n // 8: n // 4] = vtx - org[n // 8: n // 4]
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    a_tri, a_t = oracle.closest_hits(s.tris, org, d)
    b_tri, b_t = oracle.closest_hits(s.tris, org, d, brute_force=True)
    np.testing.assert_array_equal(a_tri, b_tri)
    np.testing.assert_array_equal(a_t.view(np.uint32), b_t.view(np.uint32))
    assert (a_tri >= 0).mean() > 0.5
