"""Scene I/O: OBJ + MTL loading with the reference's semantics
(include/scene.h:15-52, 109-270), OBJ writing, and PPM / PFM output
(include/image.h:55-90).  CPU only."""

import numpy as np

from photon_mapping_amd import _lib
from photon_mapping_amd import scene as sc

OBJ = """# a quad light, a smooth triangle, a mirror, a glass face and a bare face
mtllib test.mtl
v 0 1 0
v 1 1 0
v 1 1 1
v 0 1 1
v 0 0 0
v 2 0 0
v 0 3 0
vn 0 0 1
vn 0 0.6 0.8
vn 0.6 0 0.8
usemtl light
f 1 2 3 4
usemtl red
f 5//1 6//2 7//3
usemtl mirror
f 5 6 7
usemtl glass
f -3/1/1 -2/1/2 -1/1/3
f 1 3 2
"""

MTL = """newmtl light
Kd 0.75 0.75 0.75
Ke 34 24 8
illum 2

newmtl red
Kd 0.63 0.065 0.05
illum 2

newmtl mirror
Kd 0.2 0.2 0.2
illum 5

newmtl glass
Kd 0.1 0.1 0.1
Ni 1.33
illum 7
"""


def _write(tmp_path):
    (tmp_path / "test.obj").write_text(OBJ)
    (tmp_path / "test.mtl").write_text(MTL)
    return tmp_path / "test.obj"


def test_load_obj_semantics(tmp_path):
    s = sc.load_obj(_write(tmp_path))
    # the quad is fan-triangulated: 2 + 1 + 1 + 1 + 1 faces
    assert len(s.tris) == 6
    assert np.array_equal(s.tris["v"][0], [[0, 1, 0], [1, 1, 0], [1, 1, 1]])
    assert np.array_equal(s.tris["v"][1], [[0, 1, 0], [1, 1, 1], [0, 1, 1]])
    # winding normal of the quad: cross((1,0,0), (1,0,1)/sqrt2) -> (0, -1, 0)
    np.testing.assert_allclose(s.tris["n"][0], np.tile([0, -1, 0], (3, 1)), atol=1e-7)
    # vertex normals when every vertex has one
    np.testing.assert_array_equal(s.tris["n"][2],
                                  np.array([[0, 0, 1], [0, 0.6, 0.8], [0.6, 0, 0.8]], np.float32))
    # materials in first-use order, then the default Lambert 0.9
    assert list(s.mat_types) == [_lib.PM_MATERIAL_LAMBERT, _lib.PM_MATERIAL_LAMBERT,
                                 _lib.PM_MATERIAL_MIRROR, _lib.PM_MATERIAL_GLASS,
                                 _lib.PM_MATERIAL_LAMBERT]
    np.testing.assert_allclose(s.albedo, [[0.75] * 3, [0.63, 0.065, 0.05], [1] * 3, [1] * 3,
                                          [0.9] * 3], rtol=1e-7)
    assert s.ior[3] == np.float32(1.33)
    np.testing.assert_array_equal(s.mat_le[0], [34, 24, 8])
    assert not s.mat_le[1:].any()
    assert list(s.tris["material"]) == [0, 0, 1, 2, 3, 3]
    assert s.light_area == 1.0
    assert s.light_le is None


def test_negative_indices_and_mixed_normals(tmp_path):
    s = sc.load_obj(_write(tmp_path))
    # "f -3/1/1 -2/1/2 -1/1/3" refers to vertices 5, 6, 7 with normals 1..3
    np.testing.assert_array_equal(s.tris["v"][4], s.tris["v"][2])
    np.testing.assert_array_equal(s.tris["n"][4], s.tris["n"][2])


def test_write_obj_round_trip(tmp_path):
    for scene in (sc.box_scene(), sc.caustics_scene()):
        p = tmp_path / "scene.obj"
        sc.write_obj(scene, p)
        back = sc.load_obj(p)
        np.testing.assert_array_equal(back.tris["v"], scene.tris["v"])
        np.testing.assert_array_equal(back.tris["n"], scene.tris["n"])
        np.testing.assert_array_equal(back.tris["material"], scene.tris["material"])
        np.testing.assert_array_equal(back.albedo[:len(scene.albedo)],
                                      np.where(scene.mat_types[:, None] == 0, scene.albedo, 1.0))
        np.testing.assert_array_equal(back.mat_types[:len(scene.albedo)], scene.mat_types)
        lights = back.mat_le[back.tris["material"]].any(1)
        np.testing.assert_array_equal(lights, scene.tris["is_light"] == 1)
        assert abs(back.light_area - scene.light_area) < 1e-9


def test_write_ppm_and_pfm(tmp_path):
    img = np.zeros((2, 3, 3), np.float32)
    img[0, 0] = [1.0, 0.5, 0.0]
    img[1, 2] = [4.0, 0.001, 0.25]
    sc.write_ppm(tmp_path / "a.ppm", img)
    lines = (tmp_path / "a.ppm").read_text().split("\n")
    assert lines[:3] == ["P3", "3 2", "255"]
    # 255 * 0.5^(1/2.2) = 186.0 -> 186; 0.001^(1/2.2) * 255 = 11.7 -> 11;
    # 0.25^(1/2.2) * 255 = 135.3 -> 135; 4^(1/2.2) > 1 -> 255
    assert lines[3] == "255 186 0"
    assert lines[3 + 5] == "255 11 135"
    sc.write_pfm(tmp_path / "a.pfm", img)
    raw = (tmp_path / "a.pfm").read_bytes()
    head = b"PF\n3 2\n-1.0\n"
    assert raw.startswith(head)
    back = np.frombuffer(raw[len(head):], "<f4").reshape(2, 3, 3)[::-1]
    np.testing.assert_array_equal(back, img)
