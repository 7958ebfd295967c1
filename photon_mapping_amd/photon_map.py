"""Python mirror of the reference's photon-map API over the HIP C ABI.

``PhotonMap`` keeps the reference's method names and semantics
(include/photon_map.h:186-212): ``setPhotons``/``addPhoton`` copy photons in,
``build`` builds the search structure, ``queryKNearestPhotons`` returns the k
nearest photon indices farthest first together with the k-th squared
distance, ``getIthPhoton`` returns a photon by its input index.  Batched
methods (``knn``, ``gather_radiance``, ``gather_adjoint``) are the real entry
points: they take numpy arrays (host memory, the call stages and
synchronises) or torch tensors on the GPU (device memory, asynchronous on the
current stream).  Every call goes to libpm_gather.so; there is no CPU path.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Iterable, Sequence

import numpy as np

from . import _lib
from ._lib import call, ptr

QUERY_FIELDS = 10  # pos[3], ns[3], wo[3], material (int32 bits)


@dataclass
class Photon:
    """The reference's Photon record (include/photon_map.h:10-22)."""

    throughput: tuple
    position: tuple
    wi: tuple


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _stream_of(x):
    import torch

    return C.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)


def _soa(a):
    """(n, 3) -> contiguous (3, n) float32 of the same kind (numpy or torch)."""
    if _is_torch(a):
        import torch

        return a.to(torch.float32).t().contiguous()
    return np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, 3).T)


def make_queries(pos, ns, wo, material):
    """Packs query records (m, 10) float32: pos, shading normal, wo, material
    id stored as int32 bits (the 40-byte pm_query layout)."""
    if _is_torch(pos):
        import torch

        m = pos.shape[0]
        out = torch.empty((m, QUERY_FIELDS), dtype=torch.float32, device=pos.device)
        out[:, 0:3] = pos
        out[:, 3:6] = ns
        out[:, 6:9] = wo
        out.view(torch.int32)[:, 9] = torch.as_tensor(material, device=pos.device).to(torch.int32)
        return out
    pos = np.asarray(pos, np.float32)
    m = pos.shape[0]
    out = np.empty((m, QUERY_FIELDS), np.float32)
    out[:, 0:3] = pos
    out[:, 3:6] = ns
    out[:, 6:9] = wo
    out.view(np.int32)[:, 9] = np.asarray(material, np.int32)
    return out


def make_materials(albedo, types=None):
    """Packs material records (n, 4) float32: rho rgb, type (int32 bits)."""
    albedo = np.asarray(albedo, np.float32).reshape(-1, 3)
    out = np.zeros((albedo.shape[0], 4), np.float32)
    out[:, :3] = albedo
    t = np.zeros(albedo.shape[0], np.int32) if types is None else np.asarray(types, np.int32)
    out.view(np.int32)[:, 3] = t
    return out


class Context:
    """A pm_context on one device (photon maps and workspaces live in it)."""

    _default: dict[int, "Context"] = {}

    def __init__(self, device: int = 0):
        _lib.load()
        h = C.c_void_p()
        call("pm_context_create", device, C.byref(h))
        self.handle = h
        self.device = device

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        if device not in cls._default:
            cls._default[device] = Context(device)
        return cls._default[device]

    def set_option(self, option: int, value: int) -> None:
        call("pm_context_set_option", self.handle, option, int(value))

    def stats(self) -> dict:
        s = _lib.pm_stats()
        call("pm_context_get_stats", self.handle, C.byref(s))
        out = {f: getattr(s, f) for f, _ in s._fields_}
        out["phase_cycles"] = list(s.phase_cycles)
        return out

    def close(self) -> None:
        if self.handle:
            call("pm_context_destroy", self.handle)
            self.handle = None


class PhotonMap:
    """Photon map with the reference's interface (include/photon_map.h:186-212)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or Context.default()
        self._host: list[Photon] = []
        self._handle = None
        self._n = 0
        self._built = False
        # version of the flux the device map holds, and the array it came from
        self.flux_version = 0
        self._flux_ref = None

    @classmethod
    def _adopt(cls, ctx: Context, handle: C.c_void_p, n: int) -> "PhotonMap":
        """Wraps a device photon map the library created (not built yet)."""
        pmap = cls(ctx)
        pmap._handle = handle
        pmap._n = int(n)
        return pmap

    # ---- reference-named API -------------------------------------------------
    def getNPhotons(self) -> int:
        return self._n if self._handle is not None else len(self._host)

    def getIthPhoton(self, i: int) -> Photon:
        if self._handle is not None:
            out = (C.c_float * 9)()
            call("pm_photonmap_get_photon", self._handle, int(i), out)
            v = list(out)
            return Photon(tuple(v[0:3]), tuple(v[3:6]), tuple(v[6:9]))
        return self._host[i]

    def addPhoton(self, photon: Photon) -> None:
        self._sync_to_host()
        self._host.append(photon)
        self._drop_device()

    def setPhotons(self, photons: Sequence[Photon]) -> None:
        self._host = list(photons)
        self._drop_device()

    def build(self, leaf_size: int = 0) -> None:
        if self._handle is None:
            n = len(self._host)
            arr = np.zeros((n, 9), np.float32)
            for i, p in enumerate(self._host):
                arr[i, 0:3] = p.throughput
                arr[i, 3:6] = p.position
                arr[i, 6:9] = p.wi
            h = C.c_void_p()
            call("pm_photonmap_create_aos", self.ctx.handle, n, ptr(arr), _lib.PM_MEM_HOST,
                 C.byref(h))
            self._handle = h
            self._n = n
        call("pm_photonmap_build", self._handle, int(leaf_size), None)
        self._built = True

    def queryKNearestPhotons(self, p, k: int):
        """Returns (indices farthest first, max_dist2) like the reference
        (include/photon_map.h:166-183, 208-211)."""
        q = np.asarray(p, np.float32).reshape(1, 3)
        idx, r2 = self.knn(q, k)
        row = idx[0]
        return [int(i) for i in row if i >= 0], float(r2[0])

    # ---- batched API ---------------------------------------------------------------
    def set_photons_soa(self, pos, wi, flux) -> None:
        """Photons as (n, 3) arrays; numpy (host) or torch (device)."""
        n = int(pos.shape[0])
        self._drop_device()
        mem = _lib.PM_MEM_DEVICE if _is_torch(pos) else _lib.PM_MEM_HOST
        p, w, f = _soa(pos), _soa(wi), _soa(flux)
        h = C.c_void_p()
        call("pm_photonmap_create", self.ctx.handle, n, ptr(p), ptr(w), ptr(f), mem,
             C.byref(h))
        self._handle = h
        self._n = n
        self._host = []
        self.flux_version += 1
        self._flux_ref = flux

    def update_flux(self, flux) -> None:
        """Replaces every photon's flux ((n, 3), original order) without a
        rebuild; flux_version counts the replacements."""
        mem = _lib.PM_MEM_DEVICE if _is_torch(flux) else _lib.PM_MEM_HOST
        f = _soa(flux)
        stream = _stream_of(f) if mem == _lib.PM_MEM_DEVICE else None
        call("pm_photonmap_update_flux", self._handle, ptr(f), mem, stream)
        self.flux_version += 1
        self._flux_ref = flux

    def knn(self, queries, k: int):
        """queries (m, 3) -> (indices (m, k) int32 farthest first, -1 padded;
        r2 (m,) float32)."""
        m = int(queries.shape[0])
        if _is_torch(queries):
            import torch

            q = queries.to(torch.float32).t().contiguous()
            idx = torch.empty((m, k), dtype=torch.int32, device=queries.device)
            r2 = torch.empty((m,), dtype=torch.float32, device=queries.device)
            call("pm_knn", self._handle, m, ptr(q), int(k), ptr(idx), ptr(r2),
                 _lib.PM_MEM_DEVICE, _stream_of(q))
            return idx, r2
        q = np.ascontiguousarray(np.asarray(queries, np.float32).T)
        idx = np.empty((m, k), np.int32)
        r2 = np.empty((m,), np.float32)
        call("pm_knn", self._handle, m, ptr(q), int(k), ptr(idx), ptr(r2), _lib.PM_MEM_HOST,
             None)
        return idx, r2

    def gather_radiance(self, queries, materials, k: int, n_emit: float, want_key=True):
        """queries (m, 10) packed records, materials (n_mats, 4) packed
        records -> (rgb (m, 3), r2 (m,), key (m,) uint64 or None)."""
        m = int(queries.shape[0])
        n_mats = int(materials.shape[0])
        if _is_torch(queries):
            import torch

            dev = queries.device
            mats = torch.as_tensor(materials, device=dev)
            rgb = torch.empty((m, 3), dtype=torch.float32, device=dev)
            r2 = torch.empty((m,), dtype=torch.float32, device=dev)
            key = torch.empty((m,), dtype=torch.int64, device=dev) if want_key else None
            call("pm_gather_radiance", self._handle, m, ptr(queries), ptr(mats), n_mats,
                 int(k), float(n_emit), ptr(rgb), ptr(r2), ptr(key), _lib.PM_MEM_DEVICE,
                 _stream_of(queries))
            return rgb, r2, key
        q = np.ascontiguousarray(queries, np.float32)
        mats = np.ascontiguousarray(materials, np.float32)
        rgb = np.empty((m, 3), np.float32)
        r2 = np.empty((m,), np.float32)
        key = np.empty((m,), np.uint64) if want_key else None
        call("pm_gather_radiance", self._handle, m, ptr(q), ptr(mats), n_mats, int(k),
             float(n_emit), ptr(rgb), ptr(r2), ptr(key), _lib.PM_MEM_HOST, None)
        return rgb, r2, key

    def gather_adjoint(self, queries, materials, k: int, n_emit: float, dl_drgb, key=None,
                       fg_mat=None, dl_dflux=None, dl_dalbedo=None):
        """Accumulates dL/dflux (3, n) and dL/dalbedo (n_mats, 3); returns them."""
        m = int(queries.shape[0])
        n_mats = int(materials.shape[0])
        if _is_torch(queries):
            import torch

            dev = queries.device
            mats = torch.as_tensor(materials, device=dev)
            if dl_dflux is None:
                dl_dflux = torch.zeros((3, self._n), dtype=torch.float32, device=dev)
            if dl_dalbedo is None:
                dl_dalbedo = torch.zeros((n_mats, 3), dtype=torch.float32, device=dev)
            g = dl_drgb.to(torch.float32).contiguous()
            call("pm_gather_adjoint", self._handle, m, ptr(queries), ptr(mats), n_mats, int(k),
                 float(n_emit), ptr(g), ptr(key), ptr(fg_mat), ptr(dl_dflux), ptr(dl_dalbedo),
                 _lib.PM_MEM_DEVICE, _stream_of(queries))
            return dl_dflux, dl_dalbedo
        q = np.ascontiguousarray(queries, np.float32)
        mats = np.ascontiguousarray(materials, np.float32)
        if dl_dflux is None:
            dl_dflux = np.zeros((3, self._n), np.float32)
        if dl_dalbedo is None:
            dl_dalbedo = np.zeros((n_mats, 3), np.float32)
        g = np.ascontiguousarray(dl_drgb, np.float32)
        kk = None if key is None else np.ascontiguousarray(key, np.uint64)
        fg = None if fg_mat is None else np.ascontiguousarray(fg_mat, np.int32)
        call("pm_gather_adjoint", self._handle, m, ptr(q), ptr(mats), n_mats, int(k),
             float(n_emit), ptr(g), ptr(kk), ptr(fg), ptr(dl_dflux), ptr(dl_dalbedo),
             _lib.PM_MEM_HOST, None)
        return dl_dflux, dl_dalbedo

    # ---- internals -------------------------------------------------------------------
    def _sync_to_host(self) -> None:
        if self._handle is not None and not self._host:
            self._host = [self.getIthPhoton(i) for i in range(self._n)]

    def _drop_device(self) -> None:
        if self._handle is not None:
            call("pm_photonmap_destroy", self._handle)
        self._handle = None
        self._built = False

    @property
    def handle(self):
        return self._handle

    def __del__(self):
        try:
            self._drop_device()
        except Exception:
            pass


def autotune(pmap: PhotonMap, queries, materials, k: int, n_emit: float) -> dict:
    """pm_autotune on a device workload (torch tensors): picks the leaf size
    (rebuilding the map), the query-run split factor and the persistent waves
    per CU of the map's context by timing the gather; returns the report."""
    import torch

    mats = torch.as_tensor(materials, device=queries.device)
    rep = _lib.pm_tune_report()
    call("pm_autotune", pmap.handle, int(queries.shape[0]), ptr(queries), ptr(mats),
         int(mats.shape[0]), int(k), float(n_emit), _stream_of(queries), C.byref(rep))
    return {f: getattr(rep, f) for f, _ in rep._fields_ if f != "pad_"}


def gather_batched(instances, ctx: Context | None = None, want_key: bool = True):
    """Many independent radiance gathers in one persistent launch
    (pm_gather_batched).  instances: sequence of (PhotonMap, queries (m, 10)
    float32 records, materials (n_mats, 4) records, k, n_emit), queries and
    materials as torch tensors on the GPU.  Returns a list of (rgb (m, 3),
    r2 (m,), key (m,) or None) on the current stream, each equal bit for bit
    to PhotonMap.gather_radiance of that instance."""
    import torch

    instances = list(instances)
    if not instances:
        return []
    ctx = ctx or instances[0][0].ctx
    arr = (_lib.pm_instance * len(instances))()
    outs = []
    keep = []
    stream = None
    for i, (pmap, q, mats, k, n_emit) in enumerate(instances):
        dev = q.device
        stream = stream or _stream_of(q)
        m = int(q.shape[0])
        mats = torch.as_tensor(mats, device=dev)
        rgb = torch.empty((m, 3), dtype=torch.float32, device=dev)
        r2 = torch.empty((m,), dtype=torch.float32, device=dev)
        key = torch.empty((m,), dtype=torch.int64, device=dev) if want_key else None
        keep += [q, mats]
        a = arr[i]
        a.map = pmap.handle
        a.m = m
        a.q = ptr(q)
        a.mats = ptr(mats)
        a.n_mats = int(mats.shape[0])
        a.k = int(k)
        a.n_emit = float(n_emit)
        a.out_rgb = ptr(rgb)
        a.out_r2 = ptr(r2)
        a.out_key = ptr(key)
        outs.append((rgb, r2, key))
    call("pm_gather_batched", ctx.handle, len(instances), arr, stream)
    return outs


class BatchedPlan:
    """A prepared batched gather (pm_plan_create): the outputs are allocated
    here once, and ``launch()`` enqueues the whole gather on the current
    stream with kernels only, so it can be captured in a torch.cuda.CUDAGraph
    and replayed per frame.  ``outputs`` holds (rgb, r2, key) per instance."""

    def __init__(self, instances, ctx: Context | None = None):
        import torch

        instances = list(instances)
        if not instances:
            raise ValueError("a plan needs at least one instance")
        self.ctx = ctx or instances[0][0].ctx
        self._arr = (_lib.pm_instance * len(instances))()
        self._keep = []
        self.outputs = []
        for i, (pmap, q, mats, k, n_emit) in enumerate(instances):
            dev = q.device
            m = int(q.shape[0])
            mats = torch.as_tensor(mats, device=dev)
            rgb = torch.empty((m, 3), dtype=torch.float32, device=dev)
            r2 = torch.empty((m,), dtype=torch.float32, device=dev)
            key = torch.empty((m,), dtype=torch.int64, device=dev)
            self._keep += [pmap, q, mats]
            a = self._arr[i]
            a.map = pmap.handle
            a.m = m
            a.q = ptr(q)
            a.mats = ptr(mats)
            a.n_mats = int(mats.shape[0])
            a.k = int(k)
            a.n_emit = float(n_emit)
            a.out_rgb = ptr(rgb)
            a.out_r2 = ptr(r2)
            a.out_key = ptr(key)
            self.outputs.append((rgb, r2, key))
        h = C.c_void_p()
        call("pm_plan_create", self.ctx.handle, len(instances), self._arr,
             _stream_of(instances[0][1]), C.byref(h))
        self.handle = h
        self._device = instances[0][1].device

    def launch(self) -> None:
        import torch

        call("pm_plan_launch", self.handle,
             C.c_void_p(torch.cuda.current_stream(self._device).cuda_stream))

    def close(self) -> None:
        if getattr(self, "handle", None):
            call("pm_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
