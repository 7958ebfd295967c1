#!/usr/bin/env python3
"""Benchmark: kNN photon-gather queries/s (k = 100, 1,000,000 photons).

Workload (the GPU configuration of the box scene): 1,000,000 stored photons
traced from the box scene's light (paths in index order, seed 0, photons of
the paths concatenated until exactly 1,000,000 are stored; the number of
paths emitted is the estimate's divisor), k = 100, and the final-gathering
secondary-hit queries of a 512 x 512 camera at 64 samples per pixel, i.e.
16,777,216 gather queries per step.  --photon-source surrogate uses an
area-uniform surrogate array of the same size instead.  A step is one
pm_gather_radiance call over all queries with inputs resident in HBM: Morton
keys + radix sort of the queries, group anchors, and the persistent gather
kernel.  Prints one JSON line.

The roofline object describes the gather kernel (gather_kernel<0>) by the
resource that binds it, VALU issue: its VALU wave-instructions per launch
(rocprofv3 counter SQ_INSTS_VALU) over its HIP-event duration, against
1,024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction.  Beside it: the
HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, each scaled by what it reads
for a plain 1 GiB copy kernel in the same run) and their fraction of 8 TB/s,
the L2 hit rate (TCC_HIT / (TCC_HIT + TCC_MISS)), and the algorithmic bytes
per query (40 + 36 k + 16), which the kernel does not read from HBM once per
query, so their rate is reported but is no roofline fraction.  The counters
come from rocprofv3 --pmc child runs of the same workload started by this
script before it touches the GPU (one counter set per run); --no-pmc falls
back to the committed profile summaries.  The CPU baseline is the oracle's
restatement of the reference gather (kd-tree search + FP32 estimate, OpenMP
over queries) on a bounded sample of the same queries.
"""

from __future__ import annotations

import argparse
import csv
import glob
import json
import math
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "kNN gather queries/s (k=100, 1M photons), % HBM peak; fwd+adj vs fwd time"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_query(k: int) -> int:
    return 40 + 36 * k + 16


# MI355X: 256 CUs x 4 SIMDs, 2.4 GHz peak engine clock; a wave64 VALU
# instruction occupies a 16-lane SIMD for 4 cycles.
VALU_PEAK_G = 1024 * 2.4 / 4   # G wave-instructions/s

GATHER_KERNEL = re.compile(r"gather_kernel<0>")
COPY_KERNEL = re.compile(r"neg")
COPY_BYTES = 1 << 30           # the calibration copy reads and writes 1 GiB
PMC_PASSES = [
    ("valu", ["SQ_INSTS_VALU", "SQ_WAVES", "SQ_BUSY_CYCLES", "SQ_WAVE_CYCLES"]),
    ("fetch", ["FETCH_SIZE"]),
    ("write", ["WRITE_SIZE", "TCC_HIT_sum", "TCC_MISS_sum"]),
]
PROFILE_COUNTERS = ROOT / "profiles" / "gather_kernel_radiance_box1M_k100_fg512x64_counters.md"


def _read_counter_csv(path):
    """{(kernel name, dispatch): {counter: value summed over its rows}} from a
    rocprofv3 counter-collection CSV (columns found by name)."""
    out = {}
    with open(path, newline="") as fh:
        rd = csv.DictReader(fh)
        cols = rd.fieldnames or []

        def col(*names):
            for nm in names:
                for c in cols:
                    if c.lower() == nm.lower():
                        return c
            raise KeyError(f"{names} not among {cols}")

        kn = col("Kernel_Name", "Kernel-Name")
        cn = col("Counter_Name", "Counter-Name")
        cv = col("Counter_Value", "Counter-Value")
        di = col("Dispatch_Id", "Dispatch-Id", "Correlation_Id")
        for row in rd:
            key = (row[kn], int(row[di]))
            d = out.setdefault(key, {})
            d[row[cn]] = d.get(row[cn], 0.0) + float(row[cv])
    return out


def _per_dispatch(rows, kernel_re, counter, skip_first=0):
    vals = [v[counter] for (kn, _), v in sorted(rows.items(), key=lambda kv: kv[0][1])
            if kernel_re.search(kn) and counter in v]
    vals = vals[skip_first:] if len(vals) > skip_first else vals
    return sum(vals) / len(vals) if vals else None


def run_pmc_passes(args, save_dir=None):
    """rocprofv3 --pmc child runs of this script's gather workload (one
    counter set per run, before this process touches the GPU).  Returns the
    per-launch counter means of gather_kernel<0> and the copy calibration,
    or a dict with an 'error'."""
    rocprof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not Path(rocprof).exists():
        return {"error": "rocprofv3 not found"}
    child = [sys.executable, str(ROOT / "bench.py"), "--pmc-child",
             "--photons", str(args.photons), "--k", str(args.k), "--width", str(args.width),
             "--height", str(args.height), "--spp", str(args.spp), "--leaf", str(args.leaf),
             "--photon-source", args.photon_source, "--split", str(args.split),
             "--waves", str(args.waves)]
    env = dict(os.environ, TMPDIR="/tmp")
    res = {}
    for name, counters in PMC_PASSES:
        d = tempfile.mkdtemp(prefix=f"pm_pmc_{name}_", dir="/tmp")
        cmd = [rocprof, "--pmc", *counters, "-f", "csv", "-d", d, "-o", "run", "--", *child]
        log(f"pmc pass {name}: {' '.join(counters)}")
        t0 = time.perf_counter()
        try:
            proc = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE,
                                  stderr=subprocess.STDOUT, text=True, timeout=180)
        except subprocess.TimeoutExpired:
            return {"error": f"pmc pass {name} timed out"}
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            tail = "\n".join(proc.stdout.splitlines()[-5:])
            return {"error": f"pmc pass {name}: rc {proc.returncode}, no counter CSV: {tail}"}
        if save_dir:
            os.makedirs(save_dir, exist_ok=True)
            shutil.copy(files[0], os.path.join(save_dir, f"{name}_counter_collection.csv"))
        rows = _read_counter_csv(files[0])
        shutil.rmtree(d, ignore_errors=True)
        for c in counters:
            # the first gather dispatch also warms caches; the launches after
            # it are the measured kind
            res[c] = _per_dispatch(rows, GATHER_KERNEL, c, skip_first=1)
            if c in ("FETCH_SIZE", "WRITE_SIZE"):
                res[c + "_copy"] = _per_dispatch(rows, COPY_KERNEL, c)
        res[name + "_dispatches"] = sum(1 for (kn, _) in rows if GATHER_KERNEL.search(kn))
        log(f"pmc pass {name}: {time.perf_counter() - t0:.1f} s")
    return res


def pmc_from_profiles():
    """The committed counter summary of the same workload (when the live
    passes are off or fail)."""
    res = {}
    keys = {"valu_instructions_per_launch": "SQ_INSTS_VALU",
            "traffic_bytes_per_launch": "traffic_bytes", "l2_hit_rate": "l2_hit_rate"}
    try:
        for line in PROFILE_COUNTERS.read_text().splitlines():
            name, _, val = line.partition(":")
            if name in keys and val.strip():
                res[keys[name]] = float(val)
    except OSError:
        pass
    return res


def pmc_child(args) -> int:
    """The gather workload alone, for a counter pass: the inputs, the map,
    one call that also computes the radius bounds, two more calls, and three
    plain 1 GiB copies (torch.neg) that calibrate FETCH_SIZE / WRITE_SIZE."""
    import torch

    import photon_mapping_amd as pm
    from photon_mapping_amd import _lib
    from photon_mapping_amd import scene as sc

    ctx = pm.Context.default()
    s = sc.box_scene()
    ns = sc.NativeScene(s)
    if args.photon_source == "traced":
        pos, wi, flux, n_emit = sc.trace_photons(ns, args.photons, seed=0)
    else:
        pos, wi, flux = sc.surrogate_photons(s, args.photons, seed=0)
        n_emit = args.photons
    q = sc.camera_queries(ns, args.width, args.height, args.spp, depth=1, seed=0)
    if args.split > 0:
        ctx.set_option(_lib.PM_OPT_RUN_SPLIT, int(round(args.split * 1000)))
    if args.waves > 0:
        ctx.set_option(_lib.PM_OPT_WAVES_PER_CU, args.waves)
    pmap = pm.PhotonMap(ctx)
    pmap.set_photons_soa(pos, wi, flux)
    pmap.build(args.leaf)
    mats = torch.from_numpy(s.materials()).cuda()
    for _ in range(3):
        pmap.gather_radiance(q, mats, args.k, float(n_emit))
    torch.cuda.synchronize()
    a = torch.ones((COPY_BYTES // 4,), dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    print("pmc child done", flush=True)
    return 0


def roofline(k, m, gather_ms, pmc):
    """The roofline object of gather_kernel<0> from its HIP-event time and
    the counters (see the module docstring)."""
    sec = gather_ms * 1e-3
    out = {"bound": "valu", "kernel": "gather_kernel<0>", "kernel_ms": gather_ms,
           "unit": "G VALU wave-instructions/s", "peak": VALU_PEAK_G}
    valu = pmc.get("SQ_INSTS_VALU")
    out["valu_instructions_per_launch"] = valu
    out["achieved"] = valu / sec / 1e9 if valu else None
    out["frac"] = out["achieved"] / VALU_PEAK_G if valu else None
    traffic = pmc.get("traffic_bytes")
    if traffic is None and pmc.get("FETCH_SIZE") and pmc.get("WRITE_SIZE"):
        f_scale = COPY_BYTES / (pmc["FETCH_SIZE_copy"] * 1024) if pmc.get("FETCH_SIZE_copy") else 1.0
        w_scale = COPY_BYTES / (pmc["WRITE_SIZE_copy"] * 1024) if pmc.get("WRITE_SIZE_copy") else 1.0
        traffic = pmc["FETCH_SIZE"] * 1024 * f_scale + pmc["WRITE_SIZE"] * 1024 * w_scale
        out["traffic_fetch_bytes"] = pmc["FETCH_SIZE"] * 1024 * f_scale
        out["traffic_write_bytes"] = pmc["WRITE_SIZE"] * 1024 * w_scale
        out["counter_scale"] = {"FETCH_SIZE": f_scale, "WRITE_SIZE": w_scale}
    out["traffic"] = traffic
    out["hbm"] = {"achieved_GBs": traffic / sec / 1e9 if traffic else None,
                  "peak_GBs": HBM_PEAK_GBS,
                  "frac": traffic / sec / 1e9 / HBM_PEAK_GBS if traffic else None}
    hit, miss = pmc.get("TCC_HIT_sum"), pmc.get("TCC_MISS_sum")
    out["l2_hit_rate"] = (hit / (hit + miss) if hit is not None and miss is not None
                          and hit + miss > 0 else pmc.get("l2_hit_rate"))
    bq = bytes_per_query(k)
    out["algorithmic"] = {
        "bytes_per_query": bq, "rate_GBs": m * bq / sec / 1e9,
        "note": "40 + 36 k + 16 B per query over the kernel time; the kernel reads each "
                "candidate photon once per run of up to 64 queries, so this is not a "
                "fraction of the HBM roofline"}
    out["counters_source"] = pmc.get("source")
    return out


def cpu_baseline(pos, wi, flux, q, mats, k, n_emit, budget_s=12.0, m_total=None):
    """Times the oracle's reference gather on a bounded sample of queries."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import pm_oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    tree = pm_oracle.KdTree(pos)
    build_s = time.perf_counter() - t0
    rng = np.random.default_rng(0)
    probe = q[rng.choice(len(q), 4096, replace=False)]
    t0 = time.perf_counter()
    tree.gather_ref32(wi, flux, probe, mats, k, n_emit, threads)
    rate = 4096 / max(time.perf_counter() - t0, 1e-9)
    size = int(min(len(q), max(8192, rate * budget_s)))
    sample = q[rng.choice(len(q), size, replace=False)]
    t0 = time.perf_counter()
    tree.gather_ref32(wi, flux, sample, mats, k, n_emit, threads)
    dt = time.perf_counter() - t0
    return {
        "value": size / dt,
        "unit": "queries/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{size} of the {m_total or len(q)} queries drawn at random (seed 0); "
                   f"CPU restatement "
                   f"of the reference kd-tree search + FP32 estimate, g++ -O3 no -march, "
                   f"OpenMP dynamic over queries; single-threaded tree build over "
                   f"{len(pos)} photons took {build_s:.2f} s (not in the rate)"),
    }


def plain_config_bench(ctx, with_cpu, budget_s=10.0):
    """The plain method of the CPU-runnable configuration: the box scene, 100,000
    emitted photon paths (every stored photon kept: pmx_trace_photon_map), k = 100,
    256 x 256 x 16 spp, no final gathering.  Times the device frame
    (pmx_render_photon_mapping) and, in the CPU-baseline leg, the oracle's
    restatement of the reference integrator and render loop on the same photon
    array (a bounded number of rows of the same image)."""
    import numpy as np
    import torch

    from photon_mapping_amd import scene as sc

    s = sc.box_scene()
    w = h = 256
    spp, k, n_paths = 16, 100, 100_000
    with sc.NativeScene(s, ctx) as ns:
        gmap, n_stored = sc.trace_photon_map(ns, n_paths, seed=0)
        gmap.build()
        sc.render_photon_mapping(ns, gmap, n_paths, w, h, spp, k=k, seed=0)  # warm, same size
        torch.cuda.synchronize()
        img, _, st = sc.render_photon_mapping(ns, gmap, n_paths, w, h, spp, k=k, seed=0)
        torch.cuda.synchronize()
        out = {"workload": f"box scene, {n_paths} emitted paths ({n_stored} photons stored), "
                           f"k={k}, {w}x{h}x{spp} spp, no final gathering",
               "gpu_ms": st["ms"], "gpu_samples_per_s": w * h * spp / st["ms"] * 1e3,
               "image_mean": [float(x) for x in img.double().mean((0, 1)).cpu()]}
        if with_cpu:
            sys.path.insert(0, str(ROOT / "oracle"))
            import pm_oracle

            pos, wi, flux, n_emit = sc.trace_photons(ns, n_stored, seed=0)
            arrays = tuple(a.cpu().numpy() for a in (pos, wi, flux))
            cam = np.concatenate([np.asarray(s.camera[c], np.float32).ravel()
                                  for c in ("position", "forward", "right", "up")]
                                 + [np.array([s.camera["focal"]], np.float32)])
            threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
            rows = 4
            while True:
                t0 = time.perf_counter()
                pm_oracle.render("pm", s.tris, s.materials(), None, s.light_le, cam, w, rows,
                                 spp, k=k, global_map=arrays, n_emit_global=n_paths,
                                 threads=threads)
                dt = time.perf_counter() - t0
                if dt * (2 * rows) / rows > budget_s or rows >= h:
                    break
                rows *= 2
            rate = w * rows * spp / dt
            out["cpu"] = {"samples_per_s": rate, "cores": threads, "kind": "port",
                          "sample": f"the first {rows} rows at {spp} spp ({w * rows * spp} "
                                    f"samples); CPU restatement of PhotonMapping::integrate "
                                    f"and the render loop (brute-force ray casting, the "
                                    f"reference kd-tree), OpenMP over pixels",
                          "gpu_over_cpu": out["gpu_samples_per_s"] / rate}
    return out


def gpu_call_ms(fn, stream) -> float:
    """GPU time of one library call: HIP events on the call's stream around
    it, so every kernel and copy it enqueues, and the gaps between them,
    count."""
    import torch

    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    fn()
    e1.record(stream)
    e1.synchronize()
    return float(e0.elapsed_time(e1))


def water_diff_bench(ctx, dev, k=100, reps=3, frame=True):
    """The differentiation configuration on the water scene (131,072-triangle
    height field, glass IOR 1.33, walls cut into 64 albedo patches, 66
    materials): 1,000,000 traced global photons (seed 0) and the caustics map
    of 10,000,000 emitted caustic paths.  The forward pass is the two
    estimates of a 512 x 512 x 64 frame: the global map at the
    final-gathering hits (times the primary-hit albedo) and the caustics map
    at the primary diffuse hits (include/integrator.h:155-175, 239-244, 314);
    the adjoint runs both, with the forward keys reused and with the kNN
    recomputed; the adjoint with the forward keys follows its forward call and
    takes that call's query order.  HIP-event times of whole calls.
    With frame, also the rendered photon-mapping frame of the scene (both
    maps, final gathering, k = 100)."""
    import numpy as np
    import torch

    import photon_mapping_amd as pm
    from photon_mapping_amd import _lib
    from photon_mapping_amd import scene as sc

    s = sc.water_scene(0)
    out = {"workload": f"water scene ({len(s.tris)} triangles, {len(s.albedo)} materials: 64 "
                       f"wall patches), 1M global photons, 10M caustic paths, k={k}, "
                       "512x512x64 final-gathering queries + 512x512x64 caustics lookups"}
    ns = sc.NativeScene(s, ctx)
    t0 = time.perf_counter()
    pos, wi, flux, n_emit = sc.trace_photons(ns, 1_000_000, seed=0)
    torch.cuda.synchronize()
    out["global_trace_ms"] = 1e3 * (time.perf_counter() - t0)
    out["global_paths"] = int(n_emit)
    t0 = time.perf_counter()
    n_cpaths = 10_000_000
    cmap, n_c = sc.trace_photon_map(ns, n_cpaths, seed=0, kind="caustics")
    cmap.build()
    torch.cuda.synchronize()
    out["caustics_trace_build_ms"] = 1e3 * (time.perf_counter() - t0)
    out["caustic_photons"] = int(n_c)
    q, fg = sc.camera_queries(ns, 512, 512, 64, depth=1, seed=0, want_fg_mat=True)
    qp = sc.camera_queries(ns, 512, 512, 64, depth=0, seed=1)
    pmap = pm.PhotonMap(ctx)
    pmap.set_photons_soa(pos, wi, flux)
    pmap.build()
    mats = torch.from_numpy(s.materials()).to(dev)
    nm = mats.shape[0]
    m = int(q.shape[0])
    stream = torch.cuda.current_stream(dev)
    sptr = _lib.C.c_void_p(stream.cuda_stream)
    gen = torch.Generator(device=dev).manual_seed(0)
    bufs = {}
    for name, mp, qq, n_ph in (("global", pmap, q, 1_000_000), ("caustics", cmap, qp, n_c)):
        bufs[name] = dict(
            map=mp, q=qq, rgb=torch.empty((m, 3), dtype=torch.float32, device=dev),
            r2=torch.empty((m,), dtype=torch.float32, device=dev),
            key=torch.empty((m,), dtype=torch.int64, device=dev),
            g=torch.randn((m, 3), dtype=torch.float32, device=dev, generator=gen),
            dflux=torch.zeros((3, max(int(n_ph), 1)), dtype=torch.float32, device=dev))
    bufs["global"]["n_emit"], bufs["caustics"]["n_emit"] = float(n_emit), float(n_cpaths)
    bufs["global"]["fg"], bufs["caustics"]["fg"] = fg, None
    dalb = torch.zeros((nm, 3), dtype=torch.float32, device=dev)

    def fwd(b):
        _lib.call("pm_gather_radiance", b["map"].handle, m, _lib.ptr(b["q"]), _lib.ptr(mats),
                  nm, k, b["n_emit"], _lib.ptr(b["rgb"]), _lib.ptr(b["r2"]), _lib.ptr(b["key"]),
                  _lib.PM_MEM_DEVICE, sptr)

    def adj(b, reuse):
        _lib.call("pm_gather_adjoint", b["map"].handle, m, _lib.ptr(b["q"]), _lib.ptr(mats), nm,
                  k, b["n_emit"], _lib.ptr(b["g"]), _lib.ptr(b["key"]) if reuse else None,
                  _lib.ptr(b["fg"]), _lib.ptr(b["dflux"]), _lib.ptr(dalb), _lib.PM_MEM_DEVICE,
                  sptr)

    def call_ms(fn):
        return gpu_call_ms(fn, stream)

    for b in bufs.values():
        fwd(b)
        adj(b, True)
    torch.cuda.synchronize()
    times = {f"{n}_{w}": [] for n in bufs for w in ("forward", "reuse", "recompute")}
    for _ in range(reps):
        for name, b in bufs.items():
            times[f"{name}_forward"].append(call_ms(lambda: fwd(b)))
            times[f"{name}_reuse"].append(call_ms(lambda: adj(b, True)))
            times[f"{name}_recompute"].append(call_ms(lambda: adj(b, False)))
    t = {kk: float(np.mean(v)) for kk, v in times.items()}
    out.update({f"{kk}_ms": v for kk, v in t.items()})
    f = t["global_forward"] + t["caustics_forward"]
    out["forward_ms"] = f
    out["forward_queries_per_s"] = 2 * m / f * 1e3
    out["adjoint_reuse_ms"] = t["global_reuse"] + t["caustics_reuse"]
    out["adjoint_recompute_ms"] = t["global_recompute"] + t["caustics_recompute"]
    out["fwd_plus_adj_over_fwd_reuse"] = (f + out["adjoint_reuse_ms"]) / f
    out["fwd_plus_adj_over_fwd_recompute"] = (f + out["adjoint_recompute_ms"]) / f
    if frame:
        kw = dict(k=k, caustics_map=cmap, n_emit_caustics=float(n_cpaths),
                  final_gathering_depth=1, seed=0)
        sc.render_photon_mapping(ns, pmap, float(n_emit), 512, 512, 64, **kw)  # warm, same size
        torch.cuda.synchronize()
        img, _, st = sc.render_photon_mapping(ns, pmap, float(n_emit), 512, 512, 64, **kw)
        torch.cuda.synchronize()
        out["frame"] = _frame_row(st, img, 512, 512, 64, k)
    ns.close()
    return out


def _frame_row(st, img, w, h, spp, k):
    return {"resolution": f"{w}x{h}", "spp": spp, "k": k, "ms": st["ms"],
            "frames_per_s": 1e3 / st["ms"], "mrays_per_s": st["rays"] / st["ms"] / 1e3,
            "gathers_global": st["gathers_global"], "gathers_caustics": st["gathers_caustics"],
            "gathers_per_s": (st["gathers_global"] + st["gathers_caustics"]) / st["ms"] * 1e3,
            "rounds": st["iterations"], "batches": st["batches"],
            "dropped_samples": st["dropped_samples"],
            "image_mean": [float(x) for x in img.double().mean((0, 1)).cpu()]}


def caustics_frame_bench(ctx, k=100):
    """The caustics configuration's frame: the box walls and light with a
    mirror and an IOR-1.5 glass icosphere (5120 triangles each), 1,000,000
    emitted global paths and 10,000,000 emitted caustic paths (every stored
    photon kept, the reference's count semantics), 512 x 512 at 100 spp,
    final gathering, k = 100 for both maps.  Reports the trace + build time
    of the maps and the frame time."""
    import torch

    from photon_mapping_amd import scene as sc

    s = sc.caustics_scene()
    out = {"workload": "caustics scene (two 5120-triangle icospheres), 1M global + 10M caustic "
                       f"paths, k={k}, 512x512x100 spp, final gathering"}
    with sc.NativeScene(s, ctx) as ns:
        t0 = time.perf_counter()
        gmap, n_g = sc.trace_photon_map(ns, 1_000_000, seed=0)
        gmap.build()
        cmap, n_c = sc.trace_photon_map(ns, 10_000_000, seed=0, kind="caustics")
        cmap.build()
        torch.cuda.synchronize()
        out["maps_trace_build_ms"] = 1e3 * (time.perf_counter() - t0)
        out["global_photons"], out["caustic_photons"] = int(n_g), int(n_c)
        kw = dict(k=k, caustics_map=cmap, n_emit_caustics=1e7, final_gathering_depth=1, seed=0)
        # one warm frame at the same size (workspaces sized), then the timed one
        sc.render_photon_mapping(ns, gmap, 1e6, 512, 512, 100, **kw)
        torch.cuda.synchronize()
        img, _, st = sc.render_photon_mapping(ns, gmap, 1e6, 512, 512, 100, **kw)
        torch.cuda.synchronize()
        out["frame"] = _frame_row(st, img, 512, 512, 100, k)
    return out


def batched_bench(ctx, dev, n_inst=64, frames=3, seed=0):
    """Config-4 batching workload: n_inst box-scene instances of uneven size
    (wall albedos uniform in [0.1, 0.9]; stored photons log-uniform in
    [1e4, 4e6]; k from {32, 64, 100, 200}; square resolution from {128, 256,
    512, 1024}; 0-3 mirror or glass spheres; photons traced from the light
    (the emitted paths are the divisor); 16 spp final-gathering queries),
    timed three ways per frame: one plain pm_gather_radiance launch per
    instance, one pm_gather_batched launch, and the prepared plan replayed as
    a captured HIP graph."""
    import torch

    import photon_mapping_amd as pm
    from photon_mapping_amd import scene as sc

    work = sc.batching_workload(n_inst, seed, ctx=ctx)
    inst = [(w["map"], w["queries"], w["materials"], w["k"], w["n_emit"]) for w in work]
    sizes = [(w["n"], w["k"], w["res"]) for w in work]
    total_q = sum(int(q.shape[0]) for _, q, *_ in inst)
    torch.cuda.synchronize()

    def frame_plain():
        return [p.gather_radiance(q, mt, k, ne) for p, q, mt, k, ne in inst]

    def frame_batched():
        return pm.gather_batched(inst, ctx=ctx)

    def timed(fn):
        fn()  # warm-up (anchors for each map and k on the first pass)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(frames):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / frames

    t_plain = timed(frame_plain)
    t_batched = timed(frame_batched)
    plan = pm.BatchedPlan(inst, ctx=ctx)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.launch()
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        plan.launch()
    t_graph = timed(graph.replay)
    # the graph's last replay against one eager batched frame, bit for bit
    expect = frame_batched()
    torch.cuda.synchronize()
    same = all(torch.equal(a[0], b[0]) and torch.equal(a[2], b[2])
               for a, b in zip(plan.outputs, expect))
    del graph
    plan.close()
    return {
        "instances": n_inst,
        "queries_per_frame": total_q,
        "photons_total": int(sum(n for n, _, _ in sizes)),
        "plain_launches_ms_per_frame": 1e3 * t_plain,
        "batched_ms_per_frame": 1e3 * t_batched,
        "graph_ms_per_frame": 1e3 * t_graph,
        "plain_queries_per_s": total_q / t_plain,
        "batched_queries_per_s": total_q / t_batched,
        "graph_queries_per_s": total_q / t_graph,
        "batched_over_plain": t_plain / t_batched,
        "graph_over_plain": t_plain / t_graph,
        "frames_per_s_graph": 1.0 / t_graph,
        "graph_equals_eager": bool(same),
        "seed": 0,
    }


def ray_bench(ctx, dev, seed=0):
    """Closest-hit rays/s of the BVH (pmx_scene_intersect over 4M rays from
    uniform points in the unit box, isotropic directions; HIP events on the
    null stream) and caustic photons traced/s (1M stored), per scene: the box
    (36 triangles), the caustics scene (two 5120-triangle icospheres) and the
    water scene (131,202 triangles: the height field, 64 wall patches, the light)."""
    import torch

    from photon_mapping_amd import _lib
    from photon_mapping_amd import scene as sc

    out = {}
    n = 1 << 22
    g = torch.Generator(device=dev).manual_seed(seed)
    org = torch.rand((3, n), device=dev, generator=g)
    d = torch.randn((3, n), device=dev, generator=g)
    d /= d.norm(dim=0, keepdim=True)
    tri = torch.empty((n,), dtype=torch.int32, device=dev)
    t = torch.empty((n,), dtype=torch.float32, device=dev)
    for name, s in (("box", sc.box_scene()), ("caustics", sc.caustics_scene()),
                    ("water", sc.water_scene(0))):
        with sc.NativeScene(s, ctx) as ns:
            info = ns.info()

            def shoot():
                _lib.call("pmx_scene_intersect", ns.handle, n, _lib.ptr(org), _lib.ptr(d), 0,
                          _lib.ptr(tri), _lib.ptr(t), None, None, _lib.PM_MEM_DEVICE, None)

            shoot()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = 5
            e0.record()
            for _ in range(reps):
                shoot()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            hit = float((tri >= 0).float().mean())
            row = {"triangles": info["n_tri"], "bvh_nodes": info["bvh_nodes"],
                   "bvh_depth": info["bvh_depth"], "sah_cost": round(info["sah_cost"], 2),
                   "mrays_per_s": n / ms / 1e3, "hit_fraction": hit}
            if name != "box":
                sc.trace_photons(ns, 1 << 16, seed=0, kind="caustics")
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                _, _, _, paths = sc.trace_photons(ns, 1_000_000, seed=0, kind="caustics")
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                row["caustic_photons_per_s"] = 1e6 / dt
                row["caustic_paths_per_s"] = paths / dt
            out[name] = row
    return out


def render_bench(ns, pmap, n_emit, width, height, spp, k, seed=0):
    """Frames of the box scene at the benchmark's resolution: photon mapping
    with final gathering (pmx_render_photon_mapping: camera rays, direct light
    with shadow rays, final-gathering rays, the global-map gathers through the
    gather kernel; the box holds no specular surface, so its caustics map is
    empty) and the path tracer (pmx_render_path_tracing), one warm frame
    each, wall time of the call."""
    import photon_mapping_amd as pm
    import torch

    from photon_mapping_amd import scene as sc

    empty = pm.PhotonMap(pmap.ctx)
    z = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
    empty.set_photons_soa(z, z, z)
    empty.build()
    out = {}
    kw = dict(k=k, caustics_map=empty, n_emit_caustics=1.0, final_gathering_depth=1, seed=seed)
    sc.render_photon_mapping(ns, pmap, n_emit, width, height, spp, **kw)  # warm, same size
    torch.cuda.synchronize()
    img, _, st = sc.render_photon_mapping(ns, pmap, n_emit, width, height, spp, **kw)
    torch.cuda.synchronize()
    out["photon_mapping"] = {
        "resolution": f"{width}x{height}", "spp": spp, "k": k, "ms": st["ms"],
        "frames_per_s": 1e3 / st["ms"], "mrays_per_s": st["rays"] / st["ms"] / 1e3,
        "gathers": st["gathers_global"], "gathers_per_s": st["gathers_global"] / st["ms"] * 1e3,
        "rounds": st["iterations"], "batches": st["batches"],
        "dropped_samples": st["dropped_samples"],
        "image_mean": [float(x) for x in img.double().mean((0, 1)).cpu()]}
    sc.render_path_tracing(ns, width, height, spp, seed=seed)  # warm, same size
    torch.cuda.synchronize()
    img, _, st = sc.render_path_tracing(ns, width, height, spp, seed=seed)
    torch.cuda.synchronize()
    out["path_tracing"] = {
        "resolution": f"{width}x{height}", "spp": spp, "ms": st["ms"],
        "frames_per_s": 1e3 / st["ms"], "mrays_per_s": st["rays"] / st["ms"] / 1e3,
        "image_mean": [float(x) for x in img.double().mean((0, 1)).cpu()]}
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--photons", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=100)
    ap.add_argument("--width", type=int, default=512)
    ap.add_argument("--height", type=int, default=512)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--leaf", type=int, default=0)
    ap.add_argument("--photon-source", choices=["traced", "surrogate"], default="traced")
    ap.add_argument("--split", type=float, default=0.0,
                    help="query-run split factor (0: library default)")
    ap.add_argument("--waves", type=int, default=0, help="persistent waves per CU (0: default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint timing")
    ap.add_argument("--no-render", action="store_true",
                    help="skip the photon-mapping and path-tracing frame timings")
    ap.add_argument("--no-rays", action="store_true",
                    help="skip the BVH ray and caustic-tracing rates")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 64-instance batching measurement")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner run (reported beside the default setting)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a plain 1 GiB read + 1 GiB write kernel (torch.neg) three "
                         "times, the calibration of the HBM traffic counters")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-pmc", action="store_true",
                    help="skip the rocprofv3 counter passes (the roofline then reads the "
                         "committed profile summaries)")
    ap.add_argument("--pmc-dir", default="",
                    help="also keep the counter passes' CSV files in this directory")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the caustics-scene and water-scene frames")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the gather shards "
              "by query ranges across GPUs but no multi-GPU path is built.", file=sys.stderr)
        return 2
    if args.pmc_child:
        return pmc_child(args)

    # counter passes first: child processes, before this one touches the GPU
    pmc = {"error": "skipped (--no-pmc)"} if args.no_pmc else run_pmc_passes(args, args.pmc_dir)
    if "error" in pmc:
        log(f"counter passes: {pmc['error']}; reading the committed profile summaries")
        reason = pmc["error"]
        pmc = pmc_from_profiles()
        pmc["source"] = f"committed profiles/ summaries ({reason})"
    else:
        pmc["source"] = "rocprofv3 --pmc child runs of this workload in this invocation"

    import numpy as np
    import torch

    import photon_mapping_amd as pm
    from photon_mapping_amd import _lib
    from photon_mapping_amd import scene as sc

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    ctx = pm.Context.default()

    s = sc.box_scene()
    t0 = time.perf_counter()
    ns = sc.NativeScene(s)
    if args.photon_source == "traced":
        pos, wi, flux, n_paths = sc.trace_photons(ns, args.photons, seed=0)
        n_emit = float(n_paths)
    else:
        pos, wi, flux = sc.surrogate_photons(s, args.photons, seed=0)
        n_emit = float(args.photons)
    t_photons = time.perf_counter() - t0
    tracing = None
    if args.photon_source == "traced":
        # photons traced per second: the same 1M-photon map again, warm
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sc.trace_photons(ns, args.photons, seed=0)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        tracing = {"ms": 1e3 * dt, "stored_photons_per_s": args.photons / dt,
                   "paths_per_s": n_emit / dt, "emitted_paths": int(n_emit),
                   "note": "counter-based paths traced once, each path's first 16 photons "
                           "cached, the photons past them recorded, both then placed; "
                           "BVH closest hit"}
    q = sc.camera_queries(ns, args.width, args.height, args.spp, depth=1, seed=0)
    torch.cuda.synchronize()
    log(f"inputs: {args.photons} {args.photon_source} photons (divisor {n_emit:.0f}) in "
        f"{t_photons:.2f} s, {q.shape[0]} queries in {time.perf_counter() - t0:.2f} s")
    mats = torch.from_numpy(s.materials()).to(dev)
    m = q.shape[0]
    k = args.k

    if args.split > 0:
        ctx.set_option(_lib.PM_OPT_RUN_SPLIT, int(round(args.split * 1000)))
    if args.waves > 0:
        ctx.set_option(_lib.PM_OPT_WAVES_PER_CU, args.waves)
    # the map build: the process's first (the context's workspaces allocated
    # on the way), then a second map over the same photons, as a renderer that
    # traces a new map every frame builds it
    ctx.set_option(_lib.PM_OPT_TIMING, 1)
    first = pm.PhotonMap(ctx)
    first.set_photons_soa(pos, wi, flux)
    first.build(args.leaf)
    build_first_ms = ctx.stats()["build_ms"]
    pmap = pm.PhotonMap(ctx)
    pmap.set_photons_soa(pos, wi, flux)
    pmap.build(args.leaf)
    build_ms = ctx.stats()["build_ms"]
    ctx.set_option(_lib.PM_OPT_TIMING, 0)
    del first

    rgb = torch.empty((m, 3), dtype=torch.float32, device=dev)
    r2 = torch.empty((m,), dtype=torch.float32, device=dev)
    key = torch.empty((m,), dtype=torch.int64, device=dev)
    sptr = _lib.C.c_void_p(stream.cuda_stream)

    def step():
        _lib.call("pm_gather_radiance", pmap.handle, m, _lib.ptr(q), _lib.ptr(mats),
                  mats.shape[0], k, n_emit, _lib.ptr(rgb), _lib.ptr(r2), _lib.ptr(key),
                  _lib.PM_MEM_DEVICE, sptr)

    t0 = time.perf_counter()
    step()  # first call also computes the radius anchors for this k
    torch.cuda.synchronize()
    log(f"first call (anchors for k={k}): {time.perf_counter() - t0:.3f} s")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # ---- throughput pass --------------------------------------------------------
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = m * args.steps / wall

    # ---- kernel timing pass (HIP events on the call's stream) --------------------
    ctx.set_option(_lib.PM_OPT_TIMING, 1)
    gms, sms = [], []
    for _ in range(args.steps):
        step()
        st = ctx.stats()
        gms.append(st["gather_ms"])
        sms.append(st["sort_ms"])
    ctx.set_option(_lib.PM_OPT_TIMING, 0)
    # one more call with work counters (atomics) for the statistics
    ctx.set_option(_lib.PM_OPT_COUNT_STATS, 1)
    step()
    st = ctx.stats()
    ctx.set_option(_lib.PM_OPT_COUNT_STATS, 0)
    groups = max(int(st["groups"]), 1)
    work = {
        "groups": groups,
        "fallbacks": int(st["fallbacks"]),
        "refinement_rounds": int(st["refinements"]),
        "runs_per_group": st["runs"] / groups,
        "centre_candidates_per_group": st["candidates1"] / groups,
        "lane_candidates_per_run": st["candidates2"] / max(int(st["runs"]), 1),
        "pass_b_candidates_per_run": st["candidates3"] / max(int(st["runs"]), 1),
        "cut_lists": int(st["cut_lists"]),
        "split_runs": int(st["split_runs"]),
        "narrowed_runs": int(st["narrowed_runs"]),
        # share of wave time per phase of a group (statistics run)
        "phase_share": dict(zip(
            ["query+bound", "centre", "lists", "pass A", "thresholds", "pass B",
             "finish", "queue"],
            [round(c / max(sum(st["phase_cycles"]), 1), 4) for c in st["phase_cycles"]])),
        "wave_kcycles_per_group": sum(st["phase_cycles"]) / groups / 1e3,
    }
    gather_ms = float(np.mean(gms))
    sort_ms = float(np.mean(sms))

    # ---- adjoint: reuse the forward keys, and recompute the sets ------------------
    # Whole calls timed by HIP events on their stream: each round is a forward
    # call (fresh keys), the adjoint with those keys (which takes the forward
    # call's query order), then the adjoint that recomputes the sets.
    adj = None
    if not args.no_adjoint:
        g = torch.randn((m, 3), dtype=torch.float32, device=dev,
                        generator=torch.Generator(device=dev).manual_seed(0))
        dflux = torch.zeros((3, args.photons), dtype=torch.float32, device=dev)
        dalb = torch.zeros((mats.shape[0], 3), dtype=torch.float32, device=dev)

        def adjoint(kk):
            _lib.call("pm_gather_adjoint", pmap.handle, m, _lib.ptr(q), _lib.ptr(mats),
                      mats.shape[0], k, n_emit, _lib.ptr(g), _lib.ptr(kk), None,
                      _lib.ptr(dflux), _lib.ptr(dalb), _lib.PM_MEM_DEVICE, sptr)

        step()
        adjoint(key)
        adjoint(None)
        times = {"forward": [], "reuse": [], "recompute": []}
        for _ in range(max(2, args.steps // 2)):
            times["forward"].append(gpu_call_ms(step, stream))
            times["reuse"].append(gpu_call_ms(lambda: adjoint(key), stream))
            times["recompute"].append(gpu_call_ms(lambda: adjoint(None), stream))
        adj = {f"adjoint_{n}_ms": float(np.mean(times[n])) for n in ("reuse", "recompute")}
        fwd_ms = float(np.mean(times["forward"]))
        adj["forward_ms"] = fwd_ms
        adj["fwd_plus_adj_over_fwd_reuse"] = (fwd_ms + adj["adjoint_reuse_ms"]) / fwd_ms
        adj["fwd_plus_adj_over_fwd_recompute"] = (fwd_ms + adj["adjoint_recompute_ms"]) / fwd_ms
        log(f"adjoint: {adj}")
        adj["water_scene"] = water_diff_bench(ctx, dev, k, frame=not args.no_configs)
        log(f"adjoint (water scene): {adj['water_scene']}")

    batched = None
    if not args.no_batched:
        batched = batched_bench(ctx, dev)
        log(f"batched: {batched}")
    rays = None
    if not args.no_rays:
        rays = ray_bench(ctx, dev)
        log(f"rays: {rays}")
    render = None
    if not args.no_render:
        render = render_bench(ns, pmap, n_emit, args.width, args.height, args.spp, k)
        render["plain_config"] = plain_config_bench(ctx, not args.no_cpu)
        if not args.no_configs:
            render["caustics_config"] = caustics_frame_bench(ctx, k)
        log(f"render: {render}")

    tuned = None
    if not args.no_autotune:
        # on a context of its own, after every measurement above
        tctx = pm.Context(0)
        tmap = pm.PhotonMap(tctx)
        tmap.set_photons_soa(pos, wi, flux)
        tmap.build()
        tuned = pm.autotune(tmap, q, mats, k, n_emit)
        log(f"autotune: {tuned}")
        del tmap

    if args.copy_calib:
        a = torch.ones((1 << 28,), dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        log("copy calibration: 3 x torch.neg over 2^28 float32 (1 GiB read + 1 GiB write)")
        del a, b
    log(f"step {ms_per_step:.3f} ms | gather kernel {gather_ms:.3f} ms | sort+anchors "
        f"{sort_ms:.3f} ms | build {build_ms:.2f} ms | work {work}")

    cpu = None
    if not args.no_cpu:
        sel = torch.randperm(m, generator=torch.Generator().manual_seed(0))[:min(m, 2_000_000)]
        qs = q[sel.to(dev)].cpu().numpy()
        cpu = cpu_baseline(pos.cpu().numpy(), wi.cpu().numpy(), flux.cpu().numpy(), qs,
                           s.materials(), k, n_emit, args.cpu_budget, m_total=m)
        log(f"cpu baseline: {cpu['value']:.0f} queries/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"box scene, {args.photons} stored photons ({args.photon_source}), k={k}, "
                         f"final-gathering secondary-hit queries {args.width}x{args.height}x"
                         f"{args.spp} spp"),
            "photons": args.photons,
            "photon_source": args.photon_source,
            "emitted_paths": int(n_emit) if args.photon_source == "traced" else None,
            "queries": m,
            "k": k,
            "leaf_size": args.leaf or 8,
            "parallelism": "1 GPU",
        },
        "roofline": roofline(k, m, gather_ms, pmc),
        "cpu_baseline": cpu,
        "detail": {"sort_ms": sort_ms, "build_ms": build_ms, "build_first_ms": build_first_ms, **work, **(adj or {}),
                   "batched": batched, "autotune": tuned, "tracing": tracing, "rays": rays,
                   "render": render},
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
