"""Procedural scenes of the benchmark configurations and their workloads.

The box scene ("Cornell box" of the configs): the closed box [0,1]^3 without
its z = 1 face, five Lambertian walls as 10 triangles (left wall albedo
(0.63, 0.065, 0.05), right (0.14, 0.45, 0.09), floor/ceiling/back 0.75 grey),
a 0.25 x 0.25 area light centred under the ceiling at y = 0.999 facing down
(2 triangles, radiance (34, 24, 8), Lambert 0.75 like a light face with an MTL
Kd, include/scene.h:235-265), and two grey boxes (0.3 x 0.6 x 0.3 rotated +15
degrees about y at (0.35, 0.3, 0.35); 0.3 x 0.3 x 0.3 rotated -15 degrees at
(0.65, 0.15, 0.62)), 12 triangles each.  Walls face inward, boxes outward; the
winding of every triangle gives its face normal (include/scene.h:182-189).

Camera: the reference's pinhole (include/camera.h:21-47) with a 40 degree
vertical field of view, pinhole at (0.5, 0.5, 2.3) looking down -z, so every
camera ray enters the box through its open face.

Also the caustics scene (two spheres), the water scene (a 131,072-triangle
height field) and the batching instances with spheres.  Every scene is traced
through NativeScene, the device copy with its BVH.

Surrogate photons (a stand-in array for tests): uniform by area over all
triangles, wi cosine-distributed about the face normal,
flux = (Le * A_light * pi) * albedo * (0.5 + 0.5 u).
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import call, ptr

TRI_DTYPE = np.dtype([("v", np.float32, (3, 3)), ("n", np.float32, (3, 3)),
                      ("material", np.int32), ("is_light", np.int32)])
assert TRI_DTYPE.itemsize == 80


class pmx_camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("forward", C.c_float * 3),
                ("right", C.c_float * 3), ("up", C.c_float * 3), ("focal", C.c_float)]


class pmx_scene_info(C.Structure):
    _fields_ = [("n_tri", C.c_int32), ("n_lights", C.c_int32), ("bvh_nodes", C.c_int32),
                ("bvh_leaves", C.c_int32), ("bvh_depth", C.c_int32), ("reserved", C.c_int32),
                ("sah_cost", C.c_double), ("light_area", C.c_double)]


assert C.sizeof(pmx_scene_info) == 40


class pmx_render_params(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("max_depth", C.c_int32), ("final_gathering_depth", C.c_int32),
                ("k", C.c_int32), ("n_emit_global", C.c_double),
                ("n_emit_caustics", C.c_double), ("seed", C.c_uint64),
                ("batch_pixels", C.c_int32), ("out_mem", C.c_int32)]


class pmx_render_stats(C.Structure):
    _fields_ = [("samples", C.c_int64), ("dropped_samples", C.c_int64), ("rays", C.c_int64),
                ("gathers_global", C.c_int64), ("gathers_caustics", C.c_int64),
                ("batches", C.c_int32), ("iterations", C.c_int32), ("ms", C.c_double)]


assert C.sizeof(pmx_render_params) == 56
assert C.sizeof(pmx_render_stats) == 56


@dataclass
class Scene:
    tris: np.ndarray                   # TRI_DTYPE records
    albedo: np.ndarray                 # (n_mat, 3)
    mat_types: np.ndarray              # (n_mat,) int32
    light_le: np.ndarray               # (3,)
    light_area: float
    camera: dict = field(default_factory=dict)
    ior: np.ndarray | None = None     # (n_mat,) float32 for glass; None: 1.5
    mat_le: np.ndarray | None = None  # (n_mat, 3) emission per material (MTL Ke), or None

    @property
    def light_power(self) -> np.ndarray:
        """Le * A * pi: the flux of one emitted photon before division by
        the number of emitted paths (include/integrator.h:283-285)."""
        return (self.light_le * self.light_area * math.pi).astype(np.float32)

    def materials(self) -> np.ndarray:
        from .photon_map import make_materials

        return make_materials(self.albedo, self.mat_types)


def _quad(p0, p1, p2, p3, normal, mat, light=0):
    """Two triangles (p0, p1, p2), (p0, p2, p3) wound so that
    cross(v1 - v0, v2 - v0) points along `normal`."""
    p = [np.asarray(x, np.float64) for x in (p0, p1, p2, p3)]
    nrm = np.asarray(normal, np.float64)
    out = []
    for a, b, c in ((p[0], p[1], p[2]), (p[0], p[2], p[3])):
        if np.dot(np.cross(b - a, c - a), nrm) < 0:
            b, c = c, b
        rec = np.zeros((), TRI_DTYPE)
        rec["v"] = np.stack([a, b, c]).astype(np.float32)
        rec["n"] = np.tile(nrm.astype(np.float32), (3, 1))
        rec["material"] = mat
        rec["is_light"] = light
        out.append(rec)
    return out


def _box(center, size, angle_deg, mat):
    cx, cy, cz = center
    sx, sy, sz = size
    a = math.radians(angle_deg)
    ca, sa = math.cos(a), math.sin(a)

    def rot(x, z):
        return (cx + ca * x + sa * z, cz - sa * x + ca * z)

    hx, hy, hz = sx / 2, sy / 2, sz / 2
    corners = {}
    for ix, x in ((0, -hx), (1, hx)):
        for iy, y in ((0, -hy), (1, hy)):
            for iz, z in ((0, -hz), (1, hz)):
                wx, wz = rot(x, z)
                corners[(ix, iy, iz)] = (wx, cy + y, wz)

    def n_of(x, z):
        wx, wz = ca * x + sa * z, -sa * x + ca * z
        return (wx, 0.0, wz)

    c = corners
    tris = []
    tris += _quad(c[0, 1, 0], c[1, 1, 0], c[1, 1, 1], c[0, 1, 1], (0, 1, 0), mat)    # top
    tris += _quad(c[0, 0, 0], c[0, 0, 1], c[1, 0, 1], c[1, 0, 0], (0, -1, 0), mat)   # bottom
    tris += _quad(c[0, 0, 0], c[0, 1, 0], c[0, 1, 1], c[0, 0, 1], n_of(-1, 0), mat)  # -x
    tris += _quad(c[1, 0, 0], c[1, 0, 1], c[1, 1, 1], c[1, 1, 0], n_of(1, 0), mat)   # +x
    tris += _quad(c[0, 0, 0], c[1, 0, 0], c[1, 1, 0], c[0, 1, 0], n_of(0, -1), mat)  # -z
    tris += _quad(c[0, 0, 1], c[0, 1, 1], c[1, 1, 1], c[1, 0, 1], n_of(0, 1), mat)   # +z
    return tris


def box_scene() -> Scene:
    GREY, LEFT, RIGHT, LIGHT = 0, 1, 2, 3
    albedo = np.array([[0.75, 0.75, 0.75], [0.63, 0.065, 0.05], [0.14, 0.45, 0.09],
                       [0.75, 0.75, 0.75]], np.float32)
    tris = []
    tris += _quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), GREY)    # floor
    tris += _quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0), (0, -1, 0), GREY)   # ceiling
    tris += _quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 1), GREY)    # back
    tris += _quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 0, 0), LEFT)    # left
    tris += _quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), (-1, 0, 0), RIGHT)  # right
    lo, hi, y = 0.375, 0.625, 0.999
    tris += _quad((lo, y, lo), (hi, y, lo), (hi, y, hi), (lo, y, hi), (0, -1, 0), LIGHT, 1)
    tris += _box((0.35, 0.3, 0.35), (0.3, 0.6, 0.3), 15.0, GREY)
    tris += _box((0.65, 0.15, 0.62), (0.3, 0.3, 0.3), -15.0, GREY)
    arr = np.array(tris, dtype=TRI_DTYPE)
    fov = math.radians(40.0)
    focal = 1.0 / math.tan(0.5 * fov)
    pinhole = np.array([0.5, 0.5, 2.3])
    forward = np.array([0.0, 0.0, -1.0])
    cam = dict(position=pinhole - focal * forward, forward=forward, right=np.array([1.0, 0, 0]),
               up=np.array([0.0, 1.0, 0.0]), focal=focal)
    return Scene(arr, albedo, np.zeros(len(albedo), np.int32), np.array([34.0, 24.0, 8.0]),
                 0.25 * 0.25, cam)


def box_scene_instance(rng) -> Scene:
    """The box scene with the five walls' albedos drawn uniform in [0.1, 0.9]
    (the inner boxes share the floor/ceiling/back material)."""
    s = box_scene()
    s.albedo[:3] = rng.uniform(0.1, 0.9, (3, 3)).astype(np.float32)
    return s


def _box_shell(with_inner_boxes: bool = True):
    """Walls, light and (optionally) the two inner boxes of the box scene,
    with materials 0 grey, 1 left, 2 right, 3 light."""
    s = box_scene()
    if not with_inner_boxes:
        s.tris = s.tris[:12].copy()
    return s


def icosphere(center, radius, subdivisions, mat):
    """Triangles of an icosphere (20 * 4^subdivisions faces), outward
    winding, smooth per-vertex normals (the unit direction from the centre),
    so its shading and geometric normals differ (include/integrator.h:25-46)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.asarray(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]

        nxt = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nxt += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nxt
    v = np.asarray(verts)
    f = np.asarray(faces)
    out = np.zeros(len(f), TRI_DTYPE)
    pts = np.asarray(center, np.float64) + radius * v[f]           # (F, 3, 3)
    nrm = v[f]
    flip = np.einsum("fi,fi->f", np.cross(pts[:, 1] - pts[:, 0], pts[:, 2] - pts[:, 0]),
                     nrm.mean(1)) < 0
    pts[flip] = pts[flip][:, [0, 2, 1]]
    nrm[flip] = nrm[flip][:, [0, 2, 1]]
    out["v"] = pts.astype(np.float32)
    out["n"] = nrm.astype(np.float32)
    out["material"] = mat
    return out


def caustics_scene() -> Scene:
    """Caustics configuration: the box scene's walls and light (the inner
    boxes replaced by two spheres), a mirror sphere (rho 1) at (0.3, 0.18,
    0.4) and a glass sphere (rho 1, IOR 1.5) at (0.7, 0.18, 0.6), radius
    0.18, icospheres subdivided 4 times (5120 triangles each)."""
    s = _box_shell(False)
    mirror, glass = len(s.albedo), len(s.albedo) + 1
    sph = [icosphere((0.3, 0.18, 0.4), 0.18, 4, mirror),
           icosphere((0.7, 0.18, 0.6), 0.18, 4, glass)]
    s.tris = np.concatenate([s.tris] + sph)
    s.albedo = np.concatenate([s.albedo, np.ones((2, 3), np.float32)])
    s.mat_types = np.concatenate([s.mat_types, np.array([_lib.PM_MATERIAL_MIRROR,
                                                         _lib.PM_MATERIAL_GLASS], np.int32)])
    s.ior = np.concatenate([np.full(len(s.albedo) - 1, 1.5, np.float32),
                            np.array([1.5], np.float32)])
    return s


def _patch_walls():
    """The box scene's five walls cut into 64 Lambert patches, one material
    each: floor, ceiling and back wall 4 x 4 patches (materials 0-15, 16-31,
    32-47), left and right walls 2 (along z) x 4 (along y) patches (48-55,
    56-63).  Patches keep their wall's albedo (grey 0.75, left (0.63, 0.065,
    0.05), right (0.14, 0.45, 0.09)) and its inward winding.  Returns the
    triangles (256) and the (64, 3) albedos."""
    grey, left, right = (0.75, 0.75, 0.75), (0.63, 0.065, 0.05), (0.14, 0.45, 0.09)
    tris, albedo = [], []

    def wall(nu, nv, corner, du, dv, normal, rho):
        o, a, b = (np.asarray(x, np.float64) for x in (corner, du, dv))
        for i in range(nu):
            for j in range(nv):
                p0 = o + a * (i / nu) + b * (j / nv)
                p1 = o + a * ((i + 1) / nu) + b * (j / nv)
                p2 = o + a * ((i + 1) / nu) + b * ((j + 1) / nv)
                p3 = o + a * (i / nu) + b * ((j + 1) / nv)
                tris.extend(_quad(p0, p1, p2, p3, normal, len(albedo)))
                albedo.append(rho)

    wall(4, 4, (0, 0, 0), (1, 0, 0), (0, 0, 1), (0, 1, 0), grey)     # floor
    wall(4, 4, (0, 1, 0), (1, 0, 0), (0, 0, 1), (0, -1, 0), grey)    # ceiling
    wall(4, 4, (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), grey)     # back
    wall(2, 4, (0, 0, 0), (0, 0, 1), (0, 1, 0), (1, 0, 0), left)     # left
    wall(2, 4, (1, 0, 0), (0, 0, 1), (0, 1, 0), (-1, 0, 0), right)   # right
    return np.array(tris, dtype=TRI_DTYPE), np.asarray(albedo, np.float32)


def water_scene(seed: int = 0, grid: int = 256, patches: bool = True) -> Scene:
    """Differentiation configuration: the box scene's walls and light over a
    water surface, a height field on a grid x grid lattice over the floor
    (2 grid^2 triangles, 131,072 at 256) at y = 0.4 with heights the sum of 8
    sinusoids (wavevector direction uniform, magnitude uniform in [4, 40]
    rad/unit, amplitude uniform in [0, 0.02], phase uniform; numpy seed),
    glass of rho 1 and IOR 1.33, smooth normals from the analytic gradient.
    With patches (the default) the walls are the 64 one-material patches of
    _patch_walls, so the scene has 64 Lambert albedo parameters (materials
    0-63), the light's material (64) and the water (65)."""
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, 2 * math.pi, 8)
    mag = rng.uniform(4.0, 40.0, 8)
    amp = rng.uniform(0.0, 0.02, 8)
    phase = rng.uniform(0, 2 * math.pi, 8)
    kx, kz = mag * np.cos(ang), mag * np.sin(ang)
    g = np.linspace(0.0, 1.0, grid + 1)
    X, Z = np.meshgrid(g, g, indexing="ij")
    arg = X[..., None] * kx + Z[..., None] * kz + phase
    H = 0.4 + (amp * np.sin(arg)).sum(-1)
    dHx = (amp * kx * np.cos(arg)).sum(-1)
    dHz = (amp * kz * np.cos(arg)).sum(-1)
    P = np.stack([X, H, Z], -1)
    N = np.stack([-dHx, np.ones_like(H), -dHz], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    i, j = np.meshgrid(np.arange(grid), np.arange(grid), indexing="ij")
    i, j = i.ravel(), j.ravel()
    # two triangles per cell, wound so that their normals point up (+y)
    a, b, c, d = (i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)
    tri_idx = [(a, d, c), (a, c, b)]
    s = _box_shell(False)
    if patches:
        walls, alb = _patch_walls()
        light = s.tris[10:12].copy()
        light["material"] = len(alb)
        s.tris = np.concatenate([walls, light])
        s.albedo = np.concatenate([alb, s.albedo[3:4]])
        s.mat_types = np.zeros(len(s.albedo), np.int32)
    water = len(s.albedo)
    recs = []
    for p0, p1, p2 in tri_idx:
        r = np.zeros(len(i), TRI_DTYPE)
        r["v"] = np.stack([P[p0], P[p1], P[p2]], 1).astype(np.float32)
        r["n"] = np.stack([N[p0], N[p1], N[p2]], 1).astype(np.float32)
        r["material"] = water
        recs.append(r)
    hf = np.stack(recs, 1).reshape(-1)
    s.tris = np.concatenate([s.tris, hf])
    s.albedo = np.concatenate([s.albedo, np.ones((1, 3), np.float32)])
    s.mat_types = np.concatenate([s.mat_types, np.array([_lib.PM_MATERIAL_GLASS], np.int32)])
    s.ior = np.concatenate([np.full(len(s.albedo) - 1, 1.5, np.float32),
                            np.array([1.33], np.float32)])
    return s


def batch_scene_instance(rng) -> Scene:
    """Batching configuration instance: the box scene with the five walls'
    albedos uniform in [0.1, 0.9] (box_scene_instance) plus 0-3 spheres
    (radius uniform in [0.08, 0.15], resting on the floor at a uniform
    position inside the walls; icospheres subdivided 3 times), each a mirror
    or IOR-1.5 glass with probability 1/2."""
    s = box_scene_instance(rng)
    n_sph = int(rng.integers(0, 4))
    parts = [s.tris]
    albedo, types = [s.albedo], [s.mat_types]
    for _ in range(n_sph):
        r = float(rng.uniform(0.08, 0.15))
        x, z = rng.uniform(r, 1.0 - r, 2)
        glass = bool(rng.integers(0, 2))
        mat = sum(len(a) for a in albedo)
        parts.append(icosphere((x, r, z), r, 3, mat))
        albedo.append(np.ones((1, 3), np.float32))
        types.append(np.array([_lib.PM_MATERIAL_GLASS if glass else _lib.PM_MATERIAL_MIRROR],
                              np.int32))
    s.tris = np.concatenate(parts)
    s.albedo = np.concatenate(albedo)
    s.mat_types = np.concatenate(types)
    s.ior = np.full(len(s.albedo), 1.5, np.float32)
    return s


def batching_workload(n_inst: int = 64, seed: int = 0, spp: int = 16, ctx=None,
                      keep_photons: bool = False, photon_range=(1e4, 4e6),
                      resolutions=(128, 256, 512, 1024), ks=(32, 64, 100, 200)):
    """The batching configuration: n_inst box-scene instances of uneven size
    drawn with numpy seed `seed`, in this order per instance: the scene
    (batch_scene_instance: wall albedos uniform in [0.1, 0.9], 0-3 mirror or
    glass spheres), the stored photon count (log-uniform over photon_range),
    k (uniform over ks) and the square resolution (uniform over
    resolutions).  Photons are traced from the light (seed 1000 + i; the
    emitted paths are the divisor) and the queries are the final-gathering
    hits at spp samples per pixel (seed 1000 + i).  Returns one dict per
    instance: map (built PhotonMap), queries (device records), materials
    (device records), k, n_emit, n, res, and with keep_photons the host
    (pos, wi, flux) arrays."""
    import torch

    from .photon_map import Context, PhotonMap

    ctx = ctx or Context.default()
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_inst):
        s = batch_scene_instance(rng)
        lo, hi = math.log(photon_range[0]), math.log(photon_range[1])
        n = int(round(math.exp(rng.uniform(lo, hi))))
        k = int(rng.choice(list(ks)))
        res = int(rng.choice(list(resolutions)))
        with NativeScene(s, ctx) as ns:
            pos, wi, flux, n_paths = trace_photons(ns, n, seed=1000 + i)
            q = camera_queries(ns, res, res, spp, depth=1, seed=1000 + i)
        pmap = PhotonMap(ctx)
        pmap.set_photons_soa(pos, wi, flux)
        pmap.build()
        rec = dict(map=pmap, queries=q, materials=torch.from_numpy(s.materials()).to(q.device),
                   k=k, n_emit=float(n_paths), n=n, res=res, scene=s)
        if keep_photons:
            rec["photons"] = tuple(a.cpu().numpy() for a in (pos, wi, flux))
        del pos, wi, flux
        out.append(rec)
    return out


def _camera_struct(cam: dict) -> pmx_camera:
    c = pmx_camera()
    for k in ("position", "forward", "right", "up"):
        getattr(c, k)[:] = [float(v) for v in cam[k]]
    c.focal = float(cam["focal"])
    return c


def surrogate_photons(scene: Scene, n: int, seed: int = 0, device: str = "cuda"):
    """(pos, wi, flux) as (n, 3) float32 torch tensors on the device."""
    import torch
    from .photon_map import Context

    ctx = Context.default()
    soa = [torch.empty((3, n), dtype=torch.float32, device=device) for _ in range(3)]
    tris = np.ascontiguousarray(scene.tris)
    alb = np.ascontiguousarray(scene.albedo, np.float32)
    power = np.ascontiguousarray(scene.light_power, np.float32)
    call("pmx_surrogate_photons", ctx.handle, ptr(tris), len(tris), ptr(alb), len(alb),
         ptr(power), int(n), int(seed), ptr(soa[0]), ptr(soa[1]), ptr(soa[2]))
    return tuple(a.t() for a in soa)


class NativeScene:
    """The scene on the device with its BVH (pmx_scene_create): triangles,
    materials, IORs and the light radiance are copied at creation."""

    def __init__(self, scene: Scene, ctx=None):
        from .photon_map import Context

        self.ctx = ctx or Context.default()
        self.scene = scene
        tris = np.ascontiguousarray(scene.tris)
        mats = np.ascontiguousarray(scene.materials())
        ior = None if scene.ior is None else np.ascontiguousarray(scene.ior, np.float32)
        le = None if scene.light_le is None else np.ascontiguousarray(scene.light_le, np.float32)
        mle = None if scene.mat_le is None else np.ascontiguousarray(scene.mat_le, np.float32)
        h = C.c_void_p()
        call("pmx_scene_create", self.ctx.handle, ptr(tris), len(tris), ptr(mats), ptr(ior),
             len(mats), ptr(le), ptr(mle), C.byref(h))
        self.handle = h

    def info(self) -> dict:
        out = pmx_scene_info()
        call("pmx_scene_get_info", self.handle, C.byref(out))
        return {k: getattr(out, k) for k, _ in pmx_scene_info._fields_ if k != "reserved"}

    def close(self) -> None:
        if self.handle is not None:
            call("pmx_scene_destroy", self.handle)
            self.handle = None

    def __enter__(self) -> "NativeScene":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Native:
    """A NativeScene for the call: the one given, or a temporary one."""

    def __init__(self, scene):
        self.own = not isinstance(scene, NativeScene)
        self.ns = NativeScene(scene) if self.own else scene

    def __enter__(self) -> NativeScene:
        return self.ns

    def __exit__(self, *exc) -> None:
        if self.own:
            self.ns.close()


PHOTON_KINDS = {"global": 0, "caustics": 1}


def intersect(scene, org, dirs, brute_force: bool = False):
    """Closest hits of rays (pmx_scene_intersect): org and dirs (n, 3) torch
    float32 on the device -> (tri int32 (-1: miss), t, u, v) of shape (n,)."""
    import torch

    n = int(org.shape[0])
    o = org.to(torch.float32).t().contiguous()
    d = dirs.to(torch.float32).t().contiguous()
    tri = torch.empty((n,), dtype=torch.int32, device=org.device)
    t, u, v = (torch.empty((n,), dtype=torch.float32, device=org.device) for _ in range(3))
    with _Native(scene) as ns:
        call("pmx_scene_intersect", ns.handle, n, ptr(o), ptr(d), int(bool(brute_force)),
             ptr(tri), ptr(t), ptr(u), ptr(v), _lib.PM_MEM_DEVICE, None)
        torch.cuda.synchronize(org.device)
    return tri, t, u, v


def trace_photons(scene, n: int, seed: int = 0, max_depth: int = 100, kind: str = "global",
                  device: str = "cuda"):
    """Traced photon array (pmx_trace_photons): exactly n photons stored in
    (path, bounce) order.  Returns (pos, wi, flux) as (n, 3) float32 torch
    tensors on the device and the number of emitted paths (the estimate's
    divisor).  scene: a Scene or a NativeScene."""
    import torch

    soa = [torch.empty((3, n), dtype=torch.float32, device=device) for _ in range(3)]
    n_emit = C.c_int64()
    with _Native(scene) as ns:
        call("pmx_trace_photons", ns.handle, PHOTON_KINDS[kind], int(n), int(max_depth),
             int(seed), ptr(soa[0]), ptr(soa[1]), ptr(soa[2]), C.byref(n_emit))
    return tuple(a.t() for a in soa) + (int(n_emit.value),)


def trace_photon_map(scene, n_paths: int, seed: int = 0, max_depth: int = 100,
                     kind: str = "global"):
    """The reference's photon-map construction (pmx_trace_photon_map):
    n_paths paths emitted, every stored photon kept.  Returns (PhotonMap, not
    built yet, in the scene's context; number of photons stored)."""
    from .photon_map import PhotonMap

    h = C.c_void_p()
    n_stored = C.c_int64()
    with _Native(scene) as ns:
        call("pmx_trace_photon_map", ns.handle, PHOTON_KINDS[kind], int(n_paths),
             int(max_depth), int(seed), C.byref(h), C.byref(n_stored))
        ctx = ns.ctx
    return PhotonMap._adopt(ctx, h, n_stored.value), int(n_stored.value)


def camera_queries(scene, width: int, height: int, spp: int, depth: int,
                   seed: int = 0, device: str = "cuda", want_fg_mat: bool = False):
    """Packed (m, 10) query records on the device (see make_queries), m =
    width * height * spp; with want_fg_mat also the primary-hit material of
    each final-gathering query.  scene: a Scene or a NativeScene."""
    import torch

    m = width * height * spp
    out = torch.empty((m, 10), dtype=torch.float32, device=device)
    fg = torch.empty((m,), dtype=torch.int32, device=device) if want_fg_mat else None
    with _Native(scene) as ns:
        cam = _camera_struct(ns.scene.camera)
        call("pmx_fg_queries", ns.handle, C.byref(cam), width, height, spp, depth, int(seed),
             ptr(out), ptr(fg))
    return (out, fg) if want_fg_mat else out


def _render_params(width, height, spp, max_depth, fg_depth, k, n_emit_global, n_emit_caustics,
                   seed, batch_pixels):
    p = pmx_render_params()
    p.width, p.height, p.spp = int(width), int(height), int(spp)
    p.max_depth, p.final_gathering_depth, p.k = int(max_depth), int(fg_depth), int(k)
    p.n_emit_global, p.n_emit_caustics = float(n_emit_global), float(n_emit_caustics)
    p.seed, p.batch_pixels = int(seed), int(batch_pixels)
    return p


def _stats_dict(st: pmx_render_stats) -> dict:
    return {k: getattr(st, k) for k, _ in pmx_render_stats._fields_}


def render_photon_mapping(scene, global_map, n_emit_global: float, width: int, height: int,
                          spp: int, k: int = 100, caustics_map=None, n_emit_caustics: float = 1.0,
                          final_gathering_depth: int = 0, max_depth: int = 100, seed: int = 0,
                          batch_pixels: int = 0, device: str = "cuda"):
    """PhotonMapping::integrate over every pixel sample
    (pmx_render_photon_mapping).  The maps are built PhotonMaps.  Returns
    (mean, variance of the mean) as (height, width, 3) float32 device
    tensors and the render statistics."""
    import torch

    mean = torch.empty((height, width, 3), dtype=torch.float32, device=device)
    var = torch.empty_like(mean)
    p = _render_params(width, height, spp, max_depth, final_gathering_depth, k, n_emit_global,
                       n_emit_caustics, seed, batch_pixels)
    st = pmx_render_stats()
    with _Native(scene) as ns:
        cam = _camera_struct(ns.scene.camera)
        call("pmx_render_photon_mapping", ns.handle, C.byref(cam), global_map._handle,
             caustics_map._handle if caustics_map is not None else None, C.byref(p), ptr(mean),
             ptr(var), C.byref(st))
    return mean, var, _stats_dict(st)


def render_path_tracing(scene, width: int, height: int, spp: int, max_depth: int = 100,
                        seed: int = 0, batch_pixels: int = 0, device: str = "cuda"):
    """PathTracing::integrate over every pixel sample
    (pmx_render_path_tracing): (mean, variance of the mean, statistics)."""
    import torch

    mean = torch.empty((height, width, 3), dtype=torch.float32, device=device)
    var = torch.empty_like(mean)
    p = _render_params(width, height, spp, max_depth, 0, 1, 1.0, 1.0, seed, batch_pixels)
    st = pmx_render_stats()
    with _Native(scene) as ns:
        cam = _camera_struct(ns.scene.camera)
        call("pmx_render_path_tracing", ns.handle, C.byref(cam), C.byref(p), ptr(mean),
             ptr(var), C.byref(st))
    return mean, var, _stats_dict(st)


# ---- scene I/O (include/scene.h:15-52, 109-270; include/image.h:55-90) ------------

def _mtl_parse(path) -> dict:
    mats, cur = {}, None
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if not tok or tok[0].startswith("#"):
                continue
            if tok[0] == "newmtl":
                cur = " ".join(tok[1:])
                mats[cur] = {"Kd": (0.0, 0.0, 0.0), "Ke": (0.0, 0.0, 0.0), "illum": 0, "Ni": 1.0}
            elif cur is not None and tok[0] in ("Kd", "Ke"):
                mats[cur][tok[0]] = tuple(float(x) for x in tok[1:4])
            elif cur is not None and tok[0] == "illum":
                mats[cur]["illum"] = int(tok[1])
            elif cur is not None and tok[0] == "Ni":
                mats[cur]["Ni"] = float(tok[1])
    return mats


def load_obj(path, camera: dict | None = None) -> Scene:
    """Scene::loadModel (include/scene.h:109-270) for Wavefront OBJ + MTL.
    Polygons are fan-triangulated; every face gets its own three vertices.
    Normals: the face's vertex normals when all three have one, else the
    winding normal normalize(cross(normalize(v1 - v0), normalize(v2 - v0)))
    for all three (include/scene.h:182-189).  Materials (include/scene.h:15-38):
    illum 5 is a mirror of albedo 1, illum 7 glass of albedo 1 and IOR Ni,
    anything else Lambert with albedo Kd; faces without a material are
    Lambert 0.9.  A face whose material has a positive Ke is an area light
    of radiance Ke (include/scene.h:41-52).  MTL files are looked up next to
    the OBJ file."""
    import os

    base = os.path.dirname(os.path.abspath(path))
    V, N = [], []
    mtl = {}
    names = []            # material names in first-use order
    faces = []            # (vertex idx[3], normal idx[3] or None, material name)
    cur = None
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if not tok or tok[0].startswith("#"):
                continue
            if tok[0] == "v":
                V.append(tuple(float(x) for x in tok[1:4]))
            elif tok[0] == "vn":
                N.append(tuple(float(x) for x in tok[1:4]))
            elif tok[0] == "mtllib":
                for f in tok[1:]:
                    mtl.update(_mtl_parse(os.path.join(base, f)))
            elif tok[0] == "usemtl":
                cur = " ".join(tok[1:])
            elif tok[0] == "f":
                vi, ni = [], []
                for ref in tok[1:]:
                    parts = ref.split("/")
                    a = int(parts[0])
                    vi.append(a - 1 if a > 0 else len(V) + a)
                    if len(parts) >= 3 and parts[2]:
                        b = int(parts[2])
                        ni.append(b - 1 if b > 0 else len(N) + b)
                    else:
                        ni.append(None)
                for j in range(1, len(vi) - 1):
                    idx = (vi[0], vi[j], vi[j + 1])
                    nidx = (ni[0], ni[j], ni[j + 1])
                    faces.append((idx, None if None in nidx else nidx, cur))
                    if cur is not None and cur not in names:
                        names.append(cur)
    # material table: the used MTL materials, then the default Lambert 0.9
    albedo, types, ior, mat_le = [], [], [], []
    for nm in names:
        m = mtl.get(nm, {"Kd": (0.0, 0.0, 0.0), "Ke": (0.0, 0.0, 0.0), "illum": 0, "Ni": 1.0})
        if m["illum"] == 5:
            albedo.append((1.0, 1.0, 1.0)); types.append(_lib.PM_MATERIAL_MIRROR)
        elif m["illum"] == 7:
            albedo.append((1.0, 1.0, 1.0)); types.append(_lib.PM_MATERIAL_GLASS)
        else:
            albedo.append(m["Kd"]); types.append(_lib.PM_MATERIAL_LAMBERT)
        ior.append(m["Ni"])
        mat_le.append(m["Ke"])
    default = len(names)
    albedo.append((0.9, 0.9, 0.9)); types.append(_lib.PM_MATERIAL_LAMBERT)
    ior.append(1.0); mat_le.append((0.0, 0.0, 0.0))
    Vv = np.asarray(V, np.float32)
    Nn = np.asarray(N, np.float32) if N else np.zeros((0, 3), np.float32)
    tris = np.zeros(len(faces), TRI_DTYPE)
    for t, (idx, nidx, name) in enumerate(faces):
        p = Vv[list(idx)]
        tris[t]["v"] = p
        if nidx is not None:
            tris[t]["n"] = Nn[list(nidx)]
        else:
            e1 = p[1] - p[0]
            e2 = p[2] - p[0]
            e1 = e1 / np.linalg.norm(e1)
            e2 = e2 / np.linalg.norm(e2)
            n = np.cross(e1, e2)
            tris[t]["n"] = np.tile(n / np.linalg.norm(n), (3, 1))
        tris[t]["material"] = names.index(name) if name is not None else default
    le = np.asarray(mat_le, np.float32)
    emitting = (le > 0).any(1)
    light_area = 0.0
    for t in range(len(tris)):
        if emitting[tris[t]["material"]]:
            v = tris[t]["v"].astype(np.float64)
            light_area += 0.5 * float(np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0])))
    return Scene(tris, np.asarray(albedo, np.float32), np.asarray(types, np.int32), None,
                 light_area, camera or {}, np.asarray(ior, np.float32), le)


def write_obj(scene: Scene, path) -> None:
    """Writes the scene as OBJ + MTL (next to it, same stem) that load_obj
    reads back to the same triangles, normals and materials.  Emission: the
    material's mat_le, or light_le for the materials of is_light triangles."""
    import os

    stem = os.path.splitext(os.path.basename(path))[0]
    mtl_path = os.path.join(os.path.dirname(os.path.abspath(path)), stem + ".mtl")
    n_mat = len(scene.albedo)
    le = (np.zeros((n_mat, 3), np.float32) if scene.mat_le is None
          else np.asarray(scene.mat_le, np.float32).copy())
    if scene.light_le is not None:
        for t in scene.tris:
            if t["is_light"] and not (le[t["material"]] > 0).any():
                le[t["material"]] = scene.light_le
    ior = np.full(n_mat, 1.5, np.float32) if scene.ior is None else scene.ior
    with open(mtl_path, "w") as fh:
        for m in range(n_mat):
            illum = {_lib.PM_MATERIAL_MIRROR: 5, _lib.PM_MATERIAL_GLASS: 7}.get(
                int(scene.mat_types[m]), 2)
            kd = " ".join(repr(float(x)) for x in scene.albedo[m])
            ke = " ".join(repr(float(x)) for x in le[m])
            fh.write(f"newmtl m{m}\nKd {kd}\nKe {ke}\nillum {illum}\nNi {float(ior[m])!r}\n\n")
    with open(path, "w") as fh:
        fh.write(f"mtllib {stem}.mtl\n")
        cur = None
        for t, tri in enumerate(scene.tris):
            for v in tri["v"]:
                fh.write(f"v {float(v[0])!r} {float(v[1])!r} {float(v[2])!r}\n")
            for n in tri["n"]:
                fh.write(f"vn {float(n[0])!r} {float(n[1])!r} {float(n[2])!r}\n")
            m = int(tri["material"])
            if m != cur:
                fh.write(f"usemtl m{m}\n")
                cur = m
            b = 3 * t
            fh.write(f"f {b + 1}//{b + 1} {b + 2}//{b + 2} {b + 3}//{b + 3}\n")


def write_ppm(path, image, gamma: float = 2.2) -> None:
    """Image::gammaCorrection + writePPM (include/image.h:55-90): each
    channel c -> c^(1/gamma), then uint(255 c) clamped to [0, 255], written
    as ASCII P3, row 0 first.  image: (height, width, 3), numpy or torch."""
    img = image.detach().cpu().numpy() if hasattr(image, "detach") else np.asarray(image)
    img = np.asarray(img, np.float32)
    h, w, _ = img.shape
    g = np.power(np.maximum(img, 0.0), np.float32(1.0 / gamma)).astype(np.float32)
    q = np.clip(np.floor(np.float32(255.0) * g), 0, 255).astype(np.int64)
    with open(path, "w") as fh:
        fh.write(f"P3\n{w} {h}\n255\n")
        for row in q:
            fh.write("\n".join(f"{r} {g_} {b}" for r, g_, b in row))
            fh.write("\n")


def write_pfm(path, image) -> None:
    """Linear float image (height, width, 3) as a little-endian PFM, row 0 at
    the top of the picture (PFM stores bottom-up, so rows are reversed)."""
    img = image.detach().cpu().numpy() if hasattr(image, "detach") else np.asarray(image)
    img = np.asarray(img, np.float32)
    h, w, _ = img.shape
    with open(path, "wb") as fh:
        fh.write(f"PF\n{w} {h}\n-1.0\n".encode())
        fh.write(np.ascontiguousarray(img[::-1]).astype("<f4").tobytes())
