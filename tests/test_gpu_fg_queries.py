"""The device query producer (pmx_fg_queries) against the oracle's
restatement of the camera pass (oracle/render_oracle.cpp orc_fg_queries).

Both draw independent samples of the same distribution: the primary diffuse
hit of a camera ray (depth 0), or the Lambert hit of a cosine-sampled
final-gathering ray followed through specular surfaces (depth 1,
include/integrator.h:215-252).  Query points are binned by (material, cell of
an 8^3 grid over the scene, dominant axis of the shading normal), the
final-gathering queries also by the primary hit's material, and separately by
(8 x 8 block of the image, material, primary material); each pair of
histograms is compared with a two-sample chi-square test (bins with fewer
than 10 expected counts pooled); the test fails below p = 1e-4, so a
misplaced bounce, a wrong cosine sampling or a mirrored camera fails it
while sampling noise does not.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cam(s):
    return np.concatenate([np.asarray(s.camera[c], np.float32).ravel()
                           for c in ("position", "forward", "right", "up")]
                          + [np.float32([s.camera["focal"]])])


def _bins(q, fg, lo, hi):
    mat = q.view(np.int32)[:, 9].astype(np.int64)
    cell = np.clip(((q[:, :3] - lo) / (hi - lo) * 8).astype(np.int64), 0, 7)
    ax = np.argmax(np.abs(q[:, 3:6]), axis=1)
    sg = (np.take_along_axis(q[:, 3:6], ax[:, None], 1)[:, 0] > 0).astype(np.int64)
    key = (((mat + 1) * 512 + cell[:, 0] * 64 + cell[:, 1] * 8 + cell[:, 2]) * 6
           + ax * 2 + sg) * 64 + (fg.astype(np.int64) + 1)
    return key


def _bins_image(q, fg, w, h, spp):
    """(8 x 8 block of the image, material, primary material): the records are
    pixel-major, so this sees which pixel a query belongs to."""
    pix = np.arange(q.shape[0]) // spp
    blk = (pix // w * 8 // h) * 8 + (pix % w * 8 // w)
    mat = q.view(np.int32)[:, 9].astype(np.int64)
    return (blk * 64 + (mat + 1)) * 64 + (fg.astype(np.int64) + 1)


def _chi2_two_sample(a_keys, b_keys):
    from scipy import stats

    keys, inv = np.unique(np.concatenate([a_keys, b_keys]), return_inverse=True)
    a = np.bincount(inv[: len(a_keys)], minlength=len(keys)).astype(np.float64)
    b = np.bincount(inv[len(a_keys):], minlength=len(keys)).astype(np.float64)
    # pool sparse bins into one
    small = (a + b) < 20
    a = np.append(a[~small], a[small].sum())
    b = np.append(b[~small], b[small].sum())
    keep = (a + b) > 0
    a, b = a[keep], b[keep]
    ka, kb = np.sqrt(b.sum() / a.sum()), np.sqrt(a.sum() / b.sum())
    chi2 = float(np.sum((ka * a - kb * b) ** 2 / (a + b)))
    dof = len(a) - 1
    return chi2, dof, float(stats.chi2.sf(chi2, dof))


def _specular_box():
    from photon_mapping_amd import _lib
    from photon_mapping_amd import scene as sc

    s = sc.box_scene()
    mirror, glass = len(s.albedo), len(s.albedo) + 1
    s.tris["material"][12:24] = mirror
    s.tris["material"][24:36] = glass
    s.albedo = np.concatenate([s.albedo, np.ones((2, 3), np.float32)])
    s.mat_types = np.concatenate([s.mat_types, np.array([_lib.PM_MATERIAL_MIRROR,
                                                         _lib.PM_MATERIAL_GLASS], np.int32)])
    s.ior = np.full(len(s.albedo), 1.5, np.float32)
    return s


@pytest.mark.parametrize("name,depth", [("box", 0), ("box", 1), ("specular", 1),
                                        ("caustics", 1)])
def test_fg_queries_match_oracle_distribution(gpu, oracle, name, depth):
    import torch

    from photon_mapping_amd import scene as sc

    s = {"box": sc.box_scene, "specular": _specular_box, "caustics": sc.caustics_scene}[name]()
    w, h, spp = (64, 64, 16) if name != "caustics" else (32, 32, 16)
    print(f"fg queries {name} depth={depth}", flush=True)
    qg, fgg = sc.camera_queries(s, w, h, spp, depth=depth, seed=9, want_fg_mat=True)
    torch.cuda.synchronize()
    qg, fgg = qg.cpu().numpy(), fgg.cpu().numpy()
    qc, fgc = oracle.fg_queries(s.tris, s.materials(), s.ior, s.light_le, _cam(s), w, h, spp,
                                depth, seed=10, mat_le=s.mat_le)
    if depth == 0:
        fgg = np.full_like(fgg, -1)
    lo, hi = np.float32(-0.01), np.float32(1.01)
    chi2, dof, p = _chi2_two_sample(_bins(qg, fgg, lo, hi), _bins(qc, fgc, lo, hi))
    print(f"  world cells: chi2 {chi2:.1f} on {dof} dof, p = {p:.3g}", flush=True)
    assert dof > 50
    assert p > 1e-4, (chi2, dof, p)
    chi2, dof, p = _chi2_two_sample(_bins_image(qg, fgg, w, h, spp),
                                    _bins_image(qc, fgc, w, h, spp))
    print(f"  image blocks: chi2 {chi2:.1f} on {dof} dof, p = {p:.3g}", flush=True)
    assert dof > 50
    assert p > 1e-4, (chi2, dof, p)
    # placeholders (pixels whose 64 draws find no query point: edge pixels
    # that see past the open box, pixels that see only a specular box) are
    # as frequent on both sides
    pg = float((qg.view(np.int32)[:, 9] < 0).mean())
    pc = float((qc.view(np.int32)[:, 9] < 0).mean())
    print(f"  placeholders: device {pg:.4f}, restatement {pc:.4f}", flush=True)
    assert abs(pg - pc) <= 3e-3
