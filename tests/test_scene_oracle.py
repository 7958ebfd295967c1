"""Pins of the numpy closest-hit oracle (oracle/scene_oracle.py): worked
cases with exact answers, two-sidedness, the t > 1e-5 rule, ties to the lower
index, and agreement with an independent float64 plane/barycentric
computation on random rays away from edges and ties."""

import numpy as np

from oracle import scene_oracle as so

TRI = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], np.float32)


def _one(verts, o, d):
    tri, t, u, v = so.closest_hits(np.asarray(verts, np.float32),
                                   np.asarray([o], np.float32), np.asarray([d], np.float32))
    return int(tri[0]), float(t[0]), float(u[0]), float(v[0])


def test_worked_hit_is_exact():
    # u weights v1 = (1, 0, 0), v weights v2 = (0, 1, 0)
    assert _one(TRI, (0.25, 0.125, 2.0), (0, 0, -1)) == (0, 2.0, 0.25, 0.125)
    assert _one(TRI, (0.5, 0.0, 3.0), (0, 0, -0.5)) == (0, 6.0, 0.5, 0.0)


def test_two_sided():
    assert _one(TRI, (0.25, 0.125, -2.0), (0, 0, 1)) == (0, 2.0, 0.25, 0.125)


def test_misses():
    assert _one(TRI, (0.75, 0.75, 2.0), (0, 0, -1))[0] == -1      # u + v > 1
    assert _one(TRI, (-0.1, 0.25, 2.0), (0, 0, -1))[0] == -1      # u < 0
    assert _one(TRI, (0.25, 0.25, 2.0), (0, 0, 1))[0] == -1       # behind the origin
    assert _one(TRI, (0.25, 0.25, 2.0), (1, 0, 0))[0] == -1       # parallel
    tri, t, _, _ = _one(TRI, (0.25, 0.25, 2.0), (0, 0, 1))
    assert t == np.inf


def test_ray_epsilon():
    assert _one(TRI, (0.25, 0.25, 5e-6), (0, 0, -1))[0] == -1     # t = 5e-6 <= 1e-5
    assert _one(TRI, (0.25, 0.25, 2e-5), (0, 0, -1))[0] == 0


def test_closest_and_ties():
    lower = TRI.copy()
    lower[:, :, 2] = -1.0
    two = np.concatenate([lower, TRI])
    assert _one(two, (0.25, 0.25, 2.0), (0, 0, -1))[:2] == (1, 2.0)
    dup = np.concatenate([TRI, TRI, lower])
    assert _one(dup, (0.25, 0.25, 2.0), (0, 0, -1))[0] == 0
    dup2 = np.concatenate([lower, TRI, TRI])
    assert _one(dup2, (0.25, 0.25, 2.0), (0, 0, -1))[0] == 1


def test_random_against_float64_geometry():
    rng = np.random.default_rng(5)
    T, n = 40, 3000
    verts = rng.uniform(-1, 1, (T, 3, 3)).astype(np.float32)
    org = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    dirs = rng.normal(size=(n, 3)).astype(np.float32)
    tri, t, u, v = so.closest_hits(verts, org, dirs)
    # independent float64: plane intersection, then barycentrics by area
    V = verts.astype(np.float64)
    O = org.astype(np.float64)[:, None, :]
    D = dirs.astype(np.float64)[:, None, :]
    nrm = np.cross(V[:, 1] - V[:, 0], V[:, 2] - V[:, 0])
    den = np.einsum("rti,ti->rt", np.broadcast_to(D, (n, T, 3)), nrm)
    tt = np.einsum("rti,ti->rt", V[None, :, 0] - O, nrm) / den
    P = O + tt[..., None] * D
    # weights of v1 and v2: P - v0 = b1 (v1 - v0) + b2 (v2 - v0)
    w1 = np.einsum("rti,ti->rt", np.cross(P - V[None, :, 0], V[None, :, 2] - V[None, :, 0]), nrm)
    w2 = np.einsum("rti,ti->rt", np.cross(V[None, :, 1] - V[None, :, 0], P - V[None, :, 0]), nrm)
    nn = np.einsum("ti,ti->t", nrm, nrm)
    b1, b2 = w1 / nn, w2 / nn
    inside = (b1 >= 0) & (b2 >= 0) & (b1 + b2 <= 1) & (tt > 1e-5)
    t64 = np.where(inside, tt, np.inf)
    best = np.argmin(t64, axis=1)
    hit64 = np.isfinite(t64[np.arange(n), best])
    # rays whose decision is robust: every candidate's barycentrics far from the
    # edges, parallel rays excluded, and a clear gap to the runner-up hit
    margin = np.minimum(np.minimum(np.abs(b1), np.abs(b2)), np.abs(1 - b1 - b2))
    robust = (margin > 1e-4).all(1) & (np.abs(den) > 1e-3).all(1)
    srt = np.sort(t64, axis=1)
    robust &= ~np.isfinite(srt[:, 1]) | (srt[:, 1] - srt[:, 0] > 1e-4 * np.abs(srt[:, 0]) + 1e-6)
    assert robust.sum() > n // 2
    r = robust
    assert (tri[r] == np.where(hit64, best, -1)[r]).all()
    h = r & hit64
    assert h.sum() > 100
    np.testing.assert_allclose(t[h], t64[np.arange(n), best][h], rtol=1e-4)
    np.testing.assert_allclose(u[h], b1[np.arange(n), best][h], atol=1e-4)
    np.testing.assert_allclose(v[h], b2[np.arange(n), best][h], atol=1e-4)
