"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: this runs without a GPU)."""

import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(pmx?_[a-z0-9_]+)\s*\(", re.M)


def declared_functions():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in DECL.finditer(text):
            if not m.group(0).lstrip().startswith(("typedef", "#define")):
                names.add(m.group(1))
    return names


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ("pm_photonmap_create", "pm_photonmap_build", "pm_knn", "pm_gather_radiance",
                 "pm_gather_adjoint", "pm_gather_batched", "pm_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from photon_mapping_amd import _lib

    if not _lib.LIB_PATH.exists():
        pytest.fail("libpm_gather.so is not built; run __graft_entry__.build()")
    lib = C.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_bindings_cover_the_headers():
    from photon_mapping_amd import _lib

    assert set(_lib.SIGNATURES) == declared_functions()


def test_record_layouts():
    from photon_mapping_amd import _lib
    from photon_mapping_amd.scene import TRI_DTYPE

    assert C.sizeof(_lib.pm_query) == 40
    assert C.sizeof(_lib.pm_material) == 16
    assert TRI_DTYPE.itemsize == 80
    assert C.sizeof(_lib.pm_instance) == 72


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from photon_mapping_amd import _lib

    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "nope.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="missing"):
        _lib.load()
