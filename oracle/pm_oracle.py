"""ctypes wrapper of the CPU oracle (oracle/pm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, to check results.  The product library never
loads it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libpm_oracle.so"

_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    srcs = [HERE / "pm_oracle.cpp", HERE / "render_oracle.cpp", HERE / "oracle_core.hpp"]
    if not LIB.exists() or LIB.stat().st_mtime < max(s.stat().st_mtime for s in srcs):
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    lib = C.CDLL(str(LIB))
    VP, I64, I32, DBL = C.c_void_p, C.c_int64, C.c_int, C.c_double
    sig = {
        "orc_kdtree_new": (VP, [I64, VP, VP, VP]),
        "orc_kdtree_free": (None, [VP]),
        "orc_kdtree_num_nodes": (I64, [VP]),
        "orc_kdtree_knn": (I32, [VP, I64, VP, VP, VP, I32, VP, VP, I32]),
        "orc_brute_knn": (I32, [I64, VP, VP, VP, I64, VP, VP, VP, I32, VP, VP, I32]),
        "orc_radiance_fp64": (I32, [I64, VP, VP, I64, VP, VP, I32, I32, VP, VP, DBL, VP, I32]),
        "orc_radiance_ref32": (I32, [I64, VP, VP, I64, VP, VP, I32, I32, VP, VP, DBL, VP, I32]),
        "orc_gather_ref32": (I32, [VP, VP, VP, I64, VP, VP, I32, I32, DBL, VP, VP, I32]),
        "orc_adjoint_fp64": (I32, [I64, VP, VP, I64, VP, VP, I32, I32, VP, VP, DBL, VP, VP,
                                   VP]),
        "orc_render": (I64, [I32, VP, I32, VP, VP, I32, VP, VP, VP, I32, I32, I32, I32, I32, I32,
                             I64, VP, VP, VP, C.c_float, I64, VP, VP, VP, C.c_float,
                             C.c_uint64, I32, VP, VP]),
        "orc_closest_hits": (None, [VP, I32, I64, VP, VP, C.c_float, I32, I32, VP, VP]),
        "orc_fg_queries": (None, [VP, I32, VP, VP, I32, VP, VP, VP, I32, I32, I32, I32,
                                  C.c_uint64, I32, VP, VP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _soa(a):
    return np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, 3).T)


def _threads(t):
    return int(t) if t else (os.cpu_count() or 1)


class KdTree:
    """The reference kd-tree, restated (include/photon_map.h:46-184)."""

    def __init__(self, pos):
        lib = load()
        p = _soa(pos)
        self.n = p.shape[1]
        self._p = p
        self.h = lib.orc_kdtree_new(self.n, _p(p[0]), _p(p[1]), _p(p[2]))

    def knn(self, q, k: int, threads: int = 0):
        q = _soa(q)
        m = q.shape[1]
        idx = np.empty((m, k), np.int32)
        r2 = np.empty(m, np.float32)
        rc = load().orc_kdtree_knn(self.h, m, _p(q[0]), _p(q[1]), _p(q[2]), k, _p(idx), _p(r2),
                                   _threads(threads))
        assert rc == 0
        return idx, r2

    def gather_ref32(self, wi, flux, queries, mats, k, n_emit, threads: int = 0):
        w, f = _soa(wi), _soa(flux)
        q = np.ascontiguousarray(queries, np.float32)
        mt = np.ascontiguousarray(mats, np.float32)
        m = q.shape[0]
        rgb = np.empty((m, 3), np.float32)
        r2 = np.empty(m, np.float32)
        rc = load().orc_gather_ref32(self.h, _p(w), _p(f), m, _p(q), _p(mt), mt.shape[0], k,
                                     float(n_emit), _p(rgb), _p(r2), _threads(threads))
        assert rc == 0
        return rgb, r2

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_kdtree_free(self.h)
            self.h = None


def brute_knn(pos, q, k: int, threads: int = 0):
    """Exact k smallest (FP32 d^2, index) keys, farthest first, -1 padded."""
    p, qq = _soa(pos), _soa(q)
    n, m = p.shape[1], qq.shape[1]
    idx = np.empty((m, k), np.int32)
    r2 = np.empty(m, np.float32)
    rc = load().orc_brute_knn(n, _p(p[0]), _p(p[1]), _p(p[2]), m, _p(qq[0]), _p(qq[1]),
                              _p(qq[2]), k, _p(idx), _p(r2), _threads(threads))
    assert rc == 0
    return idx, r2


def radiance_fp64(wi, flux, queries, mats, idx, r2, n_emit, threads: int = 0):
    w, f = _soa(wi), _soa(flux)
    q = np.ascontiguousarray(queries, np.float32)
    mt = np.ascontiguousarray(mats, np.float32)
    idx = np.ascontiguousarray(idx, np.int32)
    r2 = np.ascontiguousarray(r2, np.float32)
    m, k = idx.shape
    out = np.empty((m, 3), np.float64)
    rc = load().orc_radiance_fp64(w.shape[1], _p(w), _p(f), m, _p(q), _p(mt), mt.shape[0], k,
                                  _p(idx), _p(r2), float(n_emit), _p(out), _threads(threads))
    assert rc == 0
    return out


def radiance_ref32(wi, flux, queries, mats, idx, r2, n_emit, threads: int = 0):
    w, f = _soa(wi), _soa(flux)
    q = np.ascontiguousarray(queries, np.float32)
    mt = np.ascontiguousarray(mats, np.float32)
    idx = np.ascontiguousarray(idx, np.int32)
    r2 = np.ascontiguousarray(r2, np.float32)
    m, k = idx.shape
    out = np.empty((m, 3), np.float32)
    rc = load().orc_radiance_ref32(w.shape[1], _p(w), _p(f), m, _p(q), _p(mt), mt.shape[0], k,
                                   _p(idx), _p(r2), float(n_emit), _p(out), _threads(threads))
    assert rc == 0
    return out


def adjoint_fp64(wi, flux, queries, mats, idx, r2, n_emit, dl_drgb):
    """Returns (dL/dflux (3, n) float64, dL/drho (n_mats, 3) float64)."""
    w, f = _soa(wi), _soa(flux)
    q = np.ascontiguousarray(queries, np.float32)
    mt = np.ascontiguousarray(mats, np.float32)
    idx = np.ascontiguousarray(idx, np.int32)
    r2 = np.ascontiguousarray(r2, np.float32)
    g = np.ascontiguousarray(dl_drgb, np.float32)
    m, k = idx.shape
    n = w.shape[1]
    dflux = np.zeros((3, n), np.float64)
    drho = np.zeros((mt.shape[0], 3), np.float64)
    rc = load().orc_adjoint_fp64(n, _p(w), _p(f), m, _p(q), _p(mt), mt.shape[0], k, _p(idx),
                                 _p(r2), float(n_emit), _p(g), _p(dflux), _p(drho))
    assert rc == 0
    return dflux, drho


def render(mode, tris, mats, ior, le, cam, width, height, spp, max_depth=100, fg_depth=0, k=100,
           global_map=None, n_emit_global=1.0, caustics_map=None, n_emit_caustics=1.0, seed=0,
           threads: int = 0, mat_le=None):
    """CPU restatement of the reference's render loop (render_oracle.cpp).
    mode "pt" (PathTracing) or "pm" (PhotonMapping).  tris: 80-byte
    triangle records; mats (n, 4) material records; ior (n,) or None; cam:
    13 floats (position, forward, right, up, focal); maps: (pos, wi, flux)
    arrays of shape (n, 3).  Returns (mean, variance of the mean) as
    (height, width, 3) float32 arrays and the number of dropped samples."""
    t = np.ascontiguousarray(tris)
    mt = np.ascontiguousarray(mats, np.float32)
    io = None if ior is None else np.ascontiguousarray(ior, np.float32)
    l = None if le is None else np.ascontiguousarray(le, np.float32)
    ml = None if mat_le is None else np.ascontiguousarray(mat_le, np.float32)
    c = np.ascontiguousarray(cam, np.float32)
    assert c.shape == (13,)

    def soa3(m):
        if m is None:
            z = np.zeros((3, 0), np.float32)
            return 0, z, z, z
        pos, wi, flux = (_soa(a) for a in m)
        return pos.shape[1], pos, wi, flux

    ng, gp, gw, gf = soa3(global_map)
    nc, cp, cw, cf = soa3(caustics_map)
    mean = np.empty((height, width, 3), np.float32)
    var = np.empty((height, width, 3), np.float32)
    dropped = load().orc_render(0 if mode == "pt" else 1, _p(t), len(t), _p(mt), _p(io),
                                mt.shape[0], _p(l), _p(ml), _p(c), width, height, spp, max_depth,
                                fg_depth, k, ng, _p(gp), _p(gw), _p(gf), float(n_emit_global),
                                nc, _p(cp), _p(cw), _p(cf), float(n_emit_caustics), int(seed),
                                _threads(threads), _p(mean), _p(var))
    return mean, var, int(dropped)


def fg_queries(tris, mats, ior, le, cam, width, height, spp, depth, seed=0, threads: int = 0,
               mat_le=None):
    """CPU restatement of the camera pass's gather query points
    (render_oracle.cpp orc_fg_queries): (m, 10) float32 records (pos, shading
    normal, wo, material as int32 bits) and the primary-hit material of each,
    m = width * height * spp."""
    t = np.ascontiguousarray(tris)
    mt = np.ascontiguousarray(mats, np.float32)
    io = None if ior is None else np.ascontiguousarray(ior, np.float32)
    l = None if le is None else np.ascontiguousarray(le, np.float32)
    ml = None if mat_le is None else np.ascontiguousarray(mat_le, np.float32)
    c = np.ascontiguousarray(cam, np.float32)
    assert c.shape == (13,)
    m = width * height * spp
    out = np.empty((m, 10), np.float32)
    fg = np.empty(m, np.int32)
    load().orc_fg_queries(_p(t), len(t), _p(mt), _p(io), mt.shape[0], _p(l), _p(ml), _p(c),
                          width, height, spp, depth, int(seed), _threads(threads), _p(out),
                          _p(fg))
    return out, fg


def closest_hits(tris, org, dirs, brute_force=False, tmax=3.0e38, threads: int = 0):
    """Closest hits of the render restatement's ray casting (BVH, or the
    brute-force loop): (triangle int32, -1 on a miss; t float32)."""
    t = np.ascontiguousarray(tris)
    o = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = o.shape[0]
    tri = np.empty(n, np.int32)
    tt = np.empty(n, np.float32)
    load().orc_closest_hits(_p(t), len(t), n, _p(o), _p(d), float(tmax), int(bool(brute_force)),
                            _threads(threads), _p(tri), _p(tt))
    return tri, tt
