#!/usr/bin/env python3
"""Writes the golden fixtures of this directory (run from the repository root:
``python tests/golden/make_golden.py``).

The reference holds no known-answer vectors for the gather path (its tests
are visual only), so the fixtures are produced by the oracle's brute-force
kNN under the reference's (FP32 d^2, index) order (include/photon_map.h:
25-41,107-126,166-183) and its FP64 restatement of the photon-density
estimate (include/integrator.h:133-152) and of its adjoint.  Each fixture is
data only: inputs and expected outputs, seeded, a few tens of KB each.  The
tests check the fixtures against an independent numpy computation (top-k by
np.lexsort over FP32 d^2 in the reference's operation order, and closed-form
sums in float64), the oracle against the fixtures, and the GPU path against
the fixtures.

Files:
  knn_<gen>.npz        pos (n,3) f32, q (m,3) f32, ks (K,) i32,
                       idx_k<k> (m,k) i32 farthest first (-1 padded),
                       r2_k<k> (m,) f32
  radiance.npz         pos, wi, flux (n,3) f32, queries (m,10) f32 records,
                       mats (M,4) f32 records, n_emit f64, ks,
                       rgb_k<k> (m,3) f64, r2_k<k> (m,) f32
  adjoint.npz          radiance inputs + dl_drgb (m,3) f32, k,
                       dflux (3,n) f64, dalbedo (M,3) f64
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT / "oracle"))

import pm_oracle  # noqa: E402
from util import (materials, photons_clustered, photons_cube, photons_lattice,  # noqa: E402
                  photons_surfaces, queries_for, query_records)

SEED = 20240601
N = 1000
M = 64
KS = (1, 8, 100)


def knn_case(name, pos, q, ks=KS):
    out = {"pos": pos, "q": q, "ks": np.asarray(ks, np.int32)}
    for k in ks:
        idx, r2 = pm_oracle.brute_knn(pos, q, k)
        out[f"idx_k{k}"] = idx.astype(np.int32)
        out[f"r2_k{k}"] = r2.astype(np.float32)
    np.savez_compressed(HERE / f"knn_{name}.npz", **out)


def main() -> None:
    pm_oracle.load()
    rng = np.random.default_rng(SEED)
    pos, _, _ = photons_cube(rng, N)
    knn_case("cube", pos, queries_for(rng, pos, M))
    pos, _, _, _ = photons_surfaces(rng, N)
    knn_case("surfaces", pos, queries_for(rng, pos, M))
    pos, _, _ = photons_clustered(rng, N)
    knn_case("clustered", pos, queries_for(rng, pos, M, jitter=0.005))
    # ties: an integer lattice (many equal d^2) and coincident photons
    pos, _, _ = photons_lattice(rng, 9)
    q = (pos[rng.integers(0, pos.shape[0], M)] + rng.integers(-1, 2, (M, 3)) * 0.5
         ).astype(np.float32)
    knn_case("lattice", pos, q, ks=(1, 7, 27, 100))
    pos = np.repeat(rng.random((40, 3), dtype=np.float32), 25, axis=0)
    knn_case("coincident", pos, queries_for(rng, pos, M), ks=(1, 25, 60))

    # radiance and adjoint on surface photons with Lambert and mirror materials
    pos, wi, flux, nrm = photons_surfaces(rng, N)
    pick = rng.integers(0, N, M)
    qpos = (pos[pick] + (nrm[pick] * 1e-3)).astype(np.float32)
    queries = query_records(rng, qpos, nrm[pick], n_mats=3)
    mats = materials(3, with_specular=True)
    n_emit = 5000.0
    rad = {"pos": pos, "wi": wi, "flux": flux, "queries": queries, "mats": mats,
           "n_emit": np.float64(n_emit), "ks": np.asarray((8, 100), np.int32)}
    for k in (8, 100):
        idx, r2 = pm_oracle.brute_knn(pos, qpos, k)
        rad[f"rgb_k{k}"] = pm_oracle.radiance_fp64(wi, flux, queries, mats, idx, r2, n_emit)
        rad[f"r2_k{k}"] = r2.astype(np.float32)
    np.savez_compressed(HERE / "radiance.npz", **rad)

    k = 100
    g = rng.normal(size=(M, 3)).astype(np.float32)
    idx, r2 = pm_oracle.brute_knn(pos, qpos, k)
    dflux, dalb = pm_oracle.adjoint_fp64(wi, flux, queries, mats, idx, r2, n_emit, g)
    adj = dict(rad)
    adj.update({"dl_drgb": g, "k": np.int32(k), "dflux": dflux, "dalbedo": dalb})
    np.savez_compressed(HERE / "adjoint.npz", **adj)
    for f in sorted(HERE.glob("*.npz")):
        print(f"{f.name}: {f.stat().st_size} bytes")


if __name__ == "__main__":
    main()
