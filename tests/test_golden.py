"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

Three layers, so that a fixture is never checked only against the code that
wrote it:
  * numpy recomputes every fixture independently: top-k by np.lexsort over
    FP32 d^2 in the reference's operation order (include/photon_map.h:33-41,
    Appendix-A key order (d^2, index)), and the estimate / adjoint sums in
    float64 (include/integrator.h:133-152);
  * the oracle's kd-tree restatement (include/photon_map.h:61-183) reproduces
    the kNN fixtures;
  * the HIP path (gpu) reproduces all of them: kNN bit-exact, the estimate
    within 1e-5 relative, the adjoint within 1e-4 relative.
"""

from pathlib import Path

import numpy as np
import pytest

from util import d2_ref, dot_ref

GOLDEN = Path(__file__).resolve().parent / "golden"
KNN_CASES = ["cube", "surfaces", "clustered", "lattice", "coincident"]
PI_F = float(np.float32(3.14159265359))  # include/core.h:7


def load(name):
    with np.load(GOLDEN / name) as z:
        return {k: z[k] for k in z.files}


def topk_numpy(pos, q, k):
    """Farthest-first k smallest (FP32 d^2, index) keys, -1 padded."""
    d2 = d2_ref(q[:, None, :], pos[None, :, :])
    n = pos.shape[0]
    idx = np.full((q.shape[0], k), -1, np.int32)
    r2 = np.zeros(q.shape[0], np.float32)
    ar = np.arange(n)
    for i in range(q.shape[0]):
        order = np.lexsort((ar, d2[i]))[: min(k, n)]
        idx[i, : order.size] = order[::-1]
        r2[i] = d2[i, order[-1]] if order.size else 0.0
    return idx, r2


def estimate_numpy(fx, idx, r2):
    wi, flux, qr, mats, n_emit = fx["wi"], fx["flux"], fx["queries"], fx["mats"], fx["n_emit"]
    mat = qr.view(np.int32)[:, 9]
    ns, wo = qr[:, 3:6], qr[:, 6:9]
    out = np.zeros((qr.shape[0], 3))
    for i in range(qr.shape[0]):
        members = idx[i][idx[i] >= 0]
        m = mat[i]
        if members.size == 0 or not (0 <= m < mats.shape[0]) or mats.view(np.int32)[m, 3] != 0:
            continue
        cos_o = dot_ref(wo[i], ns[i])
        cos_i = dot_ref(wi[members], ns[i][None, :])
        ok = ~((cos_o < 0) | (cos_i < 0))
        s = flux[members][ok].astype(np.float64).sum(axis=0)
        out[i] = (mats[m, :3].astype(np.float64) / PI_F) * s / (float(n_emit) * PI_F * float(r2[i]))
    return out


@pytest.mark.parametrize("case", KNN_CASES)
def test_knn_fixture_matches_numpy_topk(case):
    fx = load(f"knn_{case}.npz")
    for k in fx["ks"]:
        idx, r2 = topk_numpy(fx["pos"], fx["q"], int(k))
        np.testing.assert_array_equal(fx[f"idx_k{k}"], idx)
        np.testing.assert_array_equal(fx[f"r2_k{k}"].view(np.uint32), r2.view(np.uint32))


@pytest.mark.parametrize("case", KNN_CASES)
def test_knn_fixture_matches_oracle_kdtree(oracle, case):
    """The kd-tree restatement returns the fixture's r^2 bit-exactly and the
    same multiset of d^2 (index choice among exact d^2 ties at the k-th
    position is the reference's measure-zero pruning corner)."""
    fx = load(f"knn_{case}.npz")
    tree = oracle.KdTree(fx["pos"])
    for k in fx["ks"]:
        idx, r2 = tree.knn(fx["q"], int(k))
        np.testing.assert_array_equal(r2.view(np.uint32), fx[f"r2_k{k}"].view(np.uint32))
        for i in range(fx["q"].shape[0]):
            got = np.sort(d2_ref(fx["q"][i], fx["pos"][idx[i][idx[i] >= 0]]))
            exp = np.sort(d2_ref(fx["q"][i], fx["pos"][fx[f"idx_k{k}"][i][fx[f"idx_k{k}"][i] >= 0]]))
            np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))


def test_radiance_fixture_matches_numpy():
    fx = load("radiance.npz")
    q = fx["queries"][:, :3]
    for k in fx["ks"]:
        idx, r2 = topk_numpy(fx["pos"], q, int(k))
        np.testing.assert_array_equal(r2.view(np.uint32), fx[f"r2_k{k}"].view(np.uint32))
        np.testing.assert_allclose(fx[f"rgb_k{k}"], estimate_numpy(fx, idx, r2), rtol=1e-12,
                                   atol=0)


def test_adjoint_fixture_is_the_gradient_of_the_estimate():
    """The estimate is linear in flux and in albedo, so the adjoint equals
    the directional derivatives exactly: <dL/dflux, v> = <g, L(v)>."""
    fx = load("adjoint.npz")
    k = int(fx["k"])
    idx, r2 = topk_numpy(fx["pos"], fx["queries"][:, :3], k)
    g = fx["dl_drgb"].astype(np.float64)
    rng = np.random.default_rng(7)
    for _ in range(3):
        v = rng.normal(size=fx["flux"].shape)
        lv = estimate_numpy({**fx, "flux": v}, idx, r2)
        np.testing.assert_allclose(np.sum(fx["dflux"] * v.T), np.sum(g * lv), rtol=1e-9)
    base = estimate_numpy(fx, idx, r2)
    mat = fx["queries"].view(np.int32)[:, 9]
    for m in range(fx["mats"].shape[0]):
        for c in range(3):
            rho = float(fx["mats"][m, c])
            expect = np.sum(g[mat == m, c] * base[mat == m, c]) / rho if rho else 0.0
            assert fx["dalbedo"][m, c] == pytest.approx(expect, rel=1e-9, abs=1e-15)


@pytest.mark.gpu
@pytest.mark.parametrize("case", KNN_CASES)
@pytest.mark.parametrize("leaf", [8, 32])
def test_gpu_knn_matches_fixture(gpu, case, leaf):
    fx = load(f"knn_{case}.npz")
    print(f"golden knn {case} leaf {leaf}", flush=True)
    pmap = gpu.PhotonMap()
    pmap.set_photons_soa(fx["pos"], fx["pos"] * 0 + 1, fx["pos"] * 0 + 1)
    pmap.build(leaf)
    for k in fx["ks"]:
        idx, r2 = pmap.knn(fx["q"], int(k))
        np.testing.assert_array_equal(idx, fx[f"idx_k{k}"])
        np.testing.assert_array_equal(r2.view(np.uint32), fx[f"r2_k{k}"].view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("leaf", [8, 32])
def test_gpu_radiance_matches_fixture(gpu, leaf):
    fx = load("radiance.npz")
    print(f"golden radiance leaf {leaf}", flush=True)
    pmap = gpu.PhotonMap()
    pmap.set_photons_soa(fx["pos"], fx["wi"], fx["flux"])
    pmap.build(leaf)
    for k in fx["ks"]:
        rgb, r2, key = pmap.gather_radiance(fx["queries"], fx["mats"], int(k), float(fx["n_emit"]))
        np.testing.assert_array_equal(r2.view(np.uint32), fx[f"r2_k{k}"].view(np.uint32))
        np.testing.assert_allclose(rgb, fx[f"rgb_k{k}"], rtol=1e-5,
                                   atol=1e-7 * np.abs(fx[f"rgb_k{k}"]).max())


@pytest.mark.gpu
def test_gpu_adjoint_matches_fixture(gpu):
    fx = load("adjoint.npz")
    print("golden adjoint", flush=True)
    pmap = gpu.PhotonMap()
    pmap.set_photons_soa(fx["pos"], fx["wi"], fx["flux"])
    pmap.build()
    k = int(fx["k"])
    _, _, key = pmap.gather_radiance(fx["queries"], fx["mats"], k, float(fx["n_emit"]))
    for kk in (key, None):
        dflux, dalb = pmap.gather_adjoint(fx["queries"], fx["mats"], k, float(fx["n_emit"]),
                                          fx["dl_drgb"], key=kk)
        np.testing.assert_allclose(dflux, fx["dflux"], rtol=1e-4,
                                   atol=1e-6 * np.abs(fx["dflux"]).max())
        np.testing.assert_allclose(dalb, fx["dalbedo"], rtol=1e-4,
                                   atol=1e-6 * np.abs(fx["dalbedo"]).max())
