"""CPU restatement of the scene's closest-hit query, in numpy float32.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the results of
pmx_scene_intersect.  The product library never loads it.

The reference intersects rays with Embree (include/scene.h:307-348: rays
start at the origin, hits count for t in (RAY_EPS, tmax) with RAY_EPS = 1e-5
from include/core.h:19, both faces, the closest hit wins).  Embree's own
triangle test is not in the reference tree, so its hit points are parity
unpinned; this oracle restates the library's stand-in instead, the two-sided
Moller-Trumbore test (Moller and Trumbore 1997, "Fast, minimum storage
ray/triangle intersection"), operation for operation in IEEE float32:

    pv = d x e2, det = e1 . pv          (|det| < 1e-12 rejects)
    inv = 1 / det, tv = o - v0
    u = (tv . pv) inv                   (reject outside [0, 1])
    qv = tv x e1, v = (d . qv) inv      (reject v < 0 or u + v > 1)
    t = (e2 . qv) inv                   (accept t > 1e-5)

with e1 = v1 - v0 and e2 = v2 - v0, dot products summed x, y, z left to
right, and ties in t resolved to the lower triangle index.  numpy float32
arithmetic rounds every operation to nearest and never fuses a multiply
with an add, matching the library's device code (built -ffp-contract=off).
"""

from __future__ import annotations

import numpy as np

RAY_EPS = np.float32(1e-5)
DET_EPS = np.float32(1e-12)


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def closest_hits(verts: np.ndarray, org: np.ndarray, dirs: np.ndarray, chunk: int = 256):
    """verts (T, 3, 3) float32; org, dirs (n, 3) float32 ->
    (tri int32 (-1: miss), t float32 (inf: miss), u, v)."""
    verts = np.asarray(verts, np.float32)
    org = np.asarray(org, np.float32)
    dirs = np.asarray(dirs, np.float32)
    n = org.shape[0]
    p0 = [verts[:, 0, a][None, :] for a in range(3)]
    e1 = [(verts[:, 1, a] - verts[:, 0, a])[None, :] for a in range(3)]
    e2 = [(verts[:, 2, a] - verts[:, 0, a])[None, :] for a in range(3)]
    out_tri = np.full(n, -1, np.int32)
    out_t = np.full(n, np.inf, np.float32)
    out_u = np.zeros(n, np.float32)
    out_v = np.zeros(n, np.float32)
    one = np.float32(1.0)
    zero = np.float32(0.0)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        for s in range(0, n, chunk):
            o = [org[s:s + chunk, a][:, None] for a in range(3)]
            d = [dirs[s:s + chunk, a][:, None] for a in range(3)]
            pv = _cross(d, e2)
            det = _dot(e1, pv)
            inv = one / det
            tv = tuple(o[a] - p0[a] for a in range(3))
            u = _dot(tv, pv) * inv
            qv = _cross(tv, e1)
            v = _dot(d, qv) * inv
            t = _dot(e2, qv) * inv
            ok = ~(np.abs(det) < DET_EPS)
            ok &= ~((u < zero) | (u > one))
            ok &= ~((v < zero) | (u + v > one))
            ok &= t > RAY_EPS
            tt = np.where(ok, t, np.float32(np.inf))
            best = np.argmin(tt, axis=1)          # first minimum: the lower index
            rows = np.arange(tt.shape[0])
            hit = np.isfinite(tt[rows, best])
            sl = slice(s, s + tt.shape[0])
            out_tri[sl] = np.where(hit, best, -1)
            out_t[sl] = np.where(hit, tt[rows, best], np.float32(np.inf))
            out_u[sl] = np.where(hit, u[rows, best], zero)
            out_v[sl] = np.where(hit, v[rows, best], zero)
    return out_tri, out_t, out_u, out_v
