"""A scene written as OBJ + MTL and loaded back (per-material emission from
Ke instead of the is_light flags) traces the same photons and renders the
same image, bit for bit, as the scene it came from."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_obj_round_trip_traces_and_renders_identically(gpu, tmp_path):
    from photon_mapping_amd import scene as sc

    s = sc.box_scene()
    sc.write_obj(s, tmp_path / "box.obj")
    back = sc.load_obj(tmp_path / "box.obj", camera=s.camera)
    assert back.light_le is None and back.mat_le is not None
    a = sc.trace_photons(s, 20_000, seed=3)
    b = sc.trace_photons(back, 20_000, seed=3)
    assert a[3] == b[3]
    for x, y in zip(a[:3], b[:3]):
        assert bool((x == y).all())
    ia, _, _ = sc.render_path_tracing(s, 16, 16, 16, seed=2)
    ib, _, _ = sc.render_path_tracing(back, 16, 16, 16, seed=2)
    assert bool((ia == ib).all())


def test_lights_of_different_emission(gpu):
    """Two light materials of different Ke: a path tracer image of a closed
    box lit only by them scales linearly with each emission."""
    from photon_mapping_amd import scene as sc

    s = sc.box_scene()
    n_mat = len(s.albedo)
    s.tris["is_light"] = 0
    s.tris["material"][10] = n_mat          # one triangle of the light quad
    s.tris["material"][11] = n_mat + 1      # the other
    s.albedo = np.concatenate([s.albedo, np.full((2, 3), 0.75, np.float32)])
    s.mat_types = np.concatenate([s.mat_types, np.zeros(2, np.int32)])
    s.light_le = None
    le = np.zeros((n_mat + 2, 3), np.float32)
    le[n_mat] = [1.0, 0.0, 0.0]
    le[n_mat + 1] = [0.0, 0.0, 2.0]
    s.mat_le = le
    img, _, _ = sc.render_path_tracing(s, 16, 16, 64, seed=1)
    m = img.double().mean((0, 1)).cpu().numpy()
    assert m[0] > 0 and m[2] > 0 and m[1] == 0.0
    s.mat_le = le * 3
    img3, _, _ = sc.render_path_tracing(s, 16, 16, 64, seed=1)
    np.testing.assert_allclose(img3.cpu().numpy(), 3 * img.cpu().numpy(), rtol=1e-5, atol=1e-6)
