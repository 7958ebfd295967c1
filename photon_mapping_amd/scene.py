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

Surrogate photons (used until photon tracing runs on the device): uniform by
area over all 36 triangles, wi cosine-distributed about the face normal,
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


@dataclass
class Scene:
    tris: np.ndarray                   # TRI_DTYPE records
    albedo: np.ndarray                 # (n_mat, 3)
    mat_types: np.ndarray              # (n_mat,) int32
    light_le: np.ndarray               # (3,)
    light_area: float
    camera: dict = field(default_factory=dict)

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


def trace_photons(scene: Scene, n: int, seed: int = 0, max_depth: int = 100,
                  device: str = "cuda"):
    """Traced photon map (pmx_trace_photons): exactly n photons stored in
    (path, bounce) order.  Returns (pos, wi, flux) as (n, 3) float32 torch
    tensors on the device and the number of emitted paths (the estimate's
    divisor)."""
    import torch
    from .photon_map import Context

    ctx = Context.default()
    soa = [torch.empty((3, n), dtype=torch.float32, device=device) for _ in range(3)]
    tris = np.ascontiguousarray(scene.tris)
    mats = np.ascontiguousarray(scene.materials())
    le = np.ascontiguousarray(scene.light_le, np.float32)
    n_emit = C.c_int64()
    call("pmx_trace_photons", ctx.handle, ptr(tris), len(tris), ptr(mats), None, len(mats),
         ptr(le), int(n), int(max_depth), int(seed), ptr(soa[0]), ptr(soa[1]), ptr(soa[2]),
         C.byref(n_emit))
    return tuple(a.t() for a in soa) + (int(n_emit.value),)


def camera_queries(scene: Scene, width: int, height: int, spp: int, depth: int,
                   seed: int = 0, device: str = "cuda", want_fg_mat: bool = False):
    """Packed (m, 10) query records on the device (see make_queries), m =
    width * height * spp; with want_fg_mat also the primary-hit material of
    each final-gathering query."""
    import torch
    from .photon_map import Context

    ctx = Context.default()
    m = width * height * spp
    out = torch.empty((m, 10), dtype=torch.float32, device=device)
    fg = torch.empty((m,), dtype=torch.int32, device=device) if want_fg_mat else None
    tris = np.ascontiguousarray(scene.tris)
    cam = _camera_struct(scene.camera)
    call("pmx_fg_queries", ctx.handle, ptr(tris), len(tris), C.byref(cam), width, height, spp,
         depth, int(seed), ptr(out), ptr(fg))
    return (out, fg) if want_fg_mat else out
