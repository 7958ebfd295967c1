"""kNN parity on the GPU: neighbour sets bit-exact against the oracle.

Brute force gives the exact (FP32 d^2, index) order.  On continuous data the
restated reference kd-tree returns the same lists, which these tests also
check; on exact ties (lattices, coincident photons) the GPU follows the exact
order.
"""

import numpy as np
import pytest

from util import (photons_clustered, photons_cube, photons_lattice, photons_surfaces,
                  queries_for)

pytestmark = pytest.mark.gpu


def _map(pm, pos, wi, flux, leaf=0):
    m = pm.PhotonMap()
    m.set_photons_soa(pos, wi, flux)
    m.build(leaf)
    return m


@pytest.mark.parametrize("gen", ["cube", "surfaces", "clustered"])
@pytest.mark.parametrize("k", [1, 7, 100])
def test_knn_bruteforce_parity(gpu, oracle, rng, gen, k):
    n = 20000
    if gen == "cube":
        pos, wi, flux = photons_cube(rng, n)
    elif gen == "surfaces":
        pos, wi, flux, _ = photons_surfaces(rng, n)
    else:
        pos, wi, flux = photons_clustered(rng, n)
    q = queries_for(rng, pos, 2500)
    print(f"knn parity gen={gen} k={k} n={n} m={len(q)}", flush=True)
    pmap = _map(gpu, pos, wi, flux)
    idx, r2 = pmap.knn(q, k)
    bidx, br2 = oracle.brute_knn(pos, q, k)
    np.testing.assert_array_equal(idx, bidx)
    np.testing.assert_array_equal(r2.view(np.uint32), br2.view(np.uint32))
    kidx, kr2 = oracle.KdTree(pos).knn(q, k)
    np.testing.assert_array_equal(idx, kidx)


@pytest.mark.parametrize("k", [1, 100, 500])
def test_knn_dense_cluster_beside_sparse_region(gpu, oracle, rng, k):
    """A dense cluster inside a sparse cloud: the k-NN spheres of sparse
    photons near the cluster reach into it (their radius bounds must come
    from small spheres, not from sweeping the cluster), and queries sit in
    the cluster, around it and far from it."""
    from util import unit

    dense = (rng.random((27000, 3)) * 0.02 + 0.49).astype(np.float32)
    sparse = rng.random((3000, 3)).astype(np.float32)
    pos = np.concatenate([dense, sparse])
    wi = unit(rng.normal(size=pos.shape))
    flux = rng.uniform(0.1, 1.0, pos.shape).astype(np.float32)
    q = np.concatenate([queries_for(rng, pos, 1500),
                        rng.random((1000, 3)).astype(np.float32)])
    print(f"knn dense cluster k={k}", flush=True)
    idx, r2 = _map(gpu, pos, wi, flux).knn(q, k)
    bidx, br2 = oracle.brute_knn(pos, q, k)
    np.testing.assert_array_equal(idx, bidx)
    np.testing.assert_array_equal(r2.view(np.uint32), br2.view(np.uint32))


@pytest.mark.parametrize("k", [600, 1000])
def test_knn_spheres_covering_most_of_a_small_map(gpu, oracle, rng, k):
    """k close to n on a 256-leaf map: candidate spheres cover most leaves,
    so leaf-only list walks overflow and the restarted walks (whole inner
    nodes, then a depth cap) must still list every photon."""
    pos, wi, flux = photons_cube(rng, 2048)
    q = np.concatenate([queries_for(rng, pos, 300), rng.random((200, 3)).astype(np.float32)])
    print(f"knn covering spheres k={k}", flush=True)
    idx, r2 = _map(gpu, pos, wi, flux, leaf=8).knn(q, k)
    bidx, br2 = oracle.brute_knn(pos, q, k)
    np.testing.assert_array_equal(idx, bidx)
    np.testing.assert_array_equal(r2.view(np.uint32), br2.view(np.uint32))


@pytest.mark.parametrize("stride", [1, 3, 64, 4096])
def test_knn_any_anchor_stride(gpu, oracle, rng, stride):
    """Radius anchors only bound the search: neighbour sets stay exact for
    any photons-per-anchor setting, including one anchor per photon and a
    single anchor for the whole map, at k = 1 and k = 100."""
    from photon_mapping_amd import _lib

    ctx = gpu.Context(0)
    try:
        with pytest.raises(_lib.PmError):
            ctx.set_option(_lib.PM_OPT_ANCHOR_STRIDE, 4097)
        ctx.set_option(_lib.PM_OPT_ANCHOR_STRIDE, stride)
        pos, wi, flux, _ = photons_surfaces(rng, 4000)
        q = queries_for(rng, pos, 600)
        pmap = gpu.PhotonMap(ctx)
        pmap.set_photons_soa(pos, wi, flux)
        pmap.build()
        print(f"knn anchor stride {stride}", flush=True)
        for k in (1, 100):
            idx, r2 = pmap.knn(q, k)
            bidx, br2 = oracle.brute_knn(pos, q, k)
            np.testing.assert_array_equal(idx, bidx)
            np.testing.assert_array_equal(r2.view(np.uint32), br2.view(np.uint32))
        del pmap
    finally:
        ctx.close()


@pytest.mark.parametrize("n", [1, 2, 3, 31, 63, 64, 65, 100, 101, 257])
@pytest.mark.parametrize("k", [1, 64, 100, 130])
def test_knn_small_maps_and_k_over_n(gpu, oracle, rng, n, k):
    pos, wi, flux = photons_cube(rng, n)
    q = queries_for(rng, pos, 97, jitter=0.3)
    print(f"knn small n={n} k={k}", flush=True)
    pmap = _map(gpu, pos, wi, flux)
    idx, r2 = pmap.knn(q, k)
    bidx, br2 = oracle.brute_knn(pos, q, k)
    np.testing.assert_array_equal(idx, bidx)
    np.testing.assert_array_equal(r2, br2)


def test_knn_empty_map(gpu):
    pmap = gpu.PhotonMap()
    z = np.zeros((0, 3), np.float32)
    pmap.set_photons_soa(z, z, z)
    pmap.build()
    idx, r2 = pmap.knn(np.zeros((5, 3), np.float32), 3)
    assert (idx == -1).all() and (r2 == 0).all()


@pytest.mark.parametrize("side,k", [(6, 1), (6, 7), (9, 27), (9, 100)])
def test_knn_lattice_ties(gpu, oracle, rng, side, k):
    pos, wi, flux = photons_lattice(rng, side)
    q = np.vstack([pos[rng.integers(0, len(pos), 200)],
                   pos[rng.integers(0, len(pos), 200)] + np.float32(0.5)]).astype(np.float32)
    print(f"knn lattice side={side} k={k}", flush=True)
    idx, r2 = _map(gpu, pos, wi, flux).knn(q, k)
    bidx, br2 = oracle.brute_knn(pos, q, k)
    np.testing.assert_array_equal(idx, bidx)
    np.testing.assert_array_equal(r2, br2)


def test_knn_coincident_photons(gpu, oracle, rng):
    """Many photons on identical positions: the selection must split ties
    by index (exercises the refinement of an overfull histogram bin)."""
    base = rng.random((50, 3), dtype=np.float32)
    pos = base[rng.integers(0, 50, 6000)]
    wi = np.tile(np.float32([[0, 1, 0]]), (6000, 1))
    flux = np.ones((6000, 3), np.float32)
    q = np.vstack([base[:20], queries_for(rng, base, 200, jitter=0.05)]).astype(np.float32)
    print("knn coincident", flush=True)
    for k in (1, 5, 100, 333):
        idx, r2 = _map(gpu, pos, wi, flux).knn(q, k)
        bidx, br2 = oracle.brute_knn(pos, q, k)
        np.testing.assert_array_equal(idx, bidx)
        np.testing.assert_array_equal(r2, br2)


@pytest.mark.parametrize("leaf", [4, 16, 32, 64, 256])
def test_knn_leaf_sizes(gpu, oracle, rng, leaf):
    pos, wi, flux, _ = photons_surfaces(rng, 12000)
    q = queries_for(rng, pos, 1500)
    print(f"knn leaf={leaf}", flush=True)
    idx, r2 = _map(gpu, pos, wi, flux, leaf).knn(q, 50)
    bidx, br2 = oracle.brute_knn(pos, q, 50)
    np.testing.assert_array_equal(idx, bidx)
    np.testing.assert_array_equal(r2, br2)


def test_knn_device_tensors_and_reference_api(gpu, oracle, rng):
    import torch

    pos, wi, flux, _ = photons_surfaces(rng, 8000)
    q = queries_for(rng, pos, 700)
    pmap = gpu.PhotonMap()
    pmap.set_photons_soa(torch.from_numpy(pos).cuda(), torch.from_numpy(wi).cuda(),
                         torch.from_numpy(flux).cuda())
    pmap.build()
    idx, r2 = pmap.knn(torch.from_numpy(q).cuda(), 33)
    torch.cuda.synchronize()
    bidx, br2 = oracle.brute_knn(pos, q, 33)
    np.testing.assert_array_equal(idx.cpu().numpy(), bidx)
    np.testing.assert_array_equal(r2.cpu().numpy(), br2)
    # the reference-named single-query call
    lst, d2 = pmap.queryKNearestPhotons(q[3], 33)
    assert lst == bidx[3].tolist() and np.float32(d2) == br2[3]
    # getIthPhoton returns the input photon
    ph = pmap.getIthPhoton(123)
    np.testing.assert_array_equal(np.float32(ph.position), pos[123])
    np.testing.assert_array_equal(np.float32(ph.throughput), flux[123])


def test_knn_setphotons_addphoton_semantics(gpu, oracle, rng):
    pos, wi, flux = photons_cube(rng, 300)
    photons = [gpu.Photon(tuple(flux[i]), tuple(pos[i]), tuple(wi[i])) for i in range(300)]
    pmap = gpu.PhotonMap()
    pmap.setPhotons(photons[:200])
    pmap.build()
    assert pmap.getNPhotons() == 200
    for p in photons[200:]:
        pmap.addPhoton(p)
    pmap.build()
    assert pmap.getNPhotons() == 300
    q = queries_for(rng, pos, 50)
    idx, _ = pmap.knn(q, 10)
    bidx, _ = oracle.brute_knn(pos, q, 10)
    np.testing.assert_array_equal(idx, bidx)


def test_knn_rejects_bad_input(gpu):
    from photon_mapping_amd import PmError

    pmap = gpu.PhotonMap()
    pos = np.zeros((4, 3), np.float32)
    pos[2, 1] = np.nan
    with pytest.raises(PmError):
        pmap.set_photons_soa(pos, pos, pos)
    ok = np.random.default_rng(0).random((10, 3), dtype=np.float32)
    pmap.set_photons_soa(ok, ok, ok)
    with pytest.raises(PmError):
        pmap.knn(ok, 3)  # not built
    pmap.build()
    with pytest.raises(PmError):
        pmap.knn(ok, 0)


@pytest.mark.parametrize("k", [1025, 3000, 9000])
def test_knn_large_k(gpu, oracle, rng, k):
    """k beyond 1024 (the reference has no cap): rows up to 8192 sort in LDS,
    longer rows through a device scratch; both equal brute force."""
    pos, wi, flux = photons_clustered(rng, 20000)
    q = queries_for(rng, pos, 60, jitter=0.02)
    print(f"knn large k={k}", flush=True)
    pmap = _map(gpu, pos, wi, flux)
    idx, r2 = pmap.knn(q, k)
    bidx, br2 = oracle.brute_knn(pos, q, k)
    np.testing.assert_array_equal(r2, br2)
    np.testing.assert_array_equal(idx, bidx)
