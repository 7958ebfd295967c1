"""The header-only C++ drop-in (include/pm/photon_map.hpp) compiles against
the C ABI and, on the GPU, answers like the reference's PhotonMap."""

import subprocess
from pathlib import Path

import numpy as np
import pytest

from util import photons_cube, queries_for

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "cpp" / "dropin_knn.cpp"


def _compile(tmp_path) -> Path:
    exe = tmp_path / "dropin_knn"
    lib_dir = ROOT / "photon_mapping_amd"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(SRC),
           "-o", str(exe), f"-L{lib_dir}", "-l:libpm_gather.so", f"-Wl,-rpath,{lib_dir}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_dropin_header_compiles(tmp_path):
    assert _compile(tmp_path).exists()


@pytest.mark.gpu
def test_dropin_knn_matches_oracle(tmp_path, oracle, rng):
    exe = _compile(tmp_path)
    pos, wi, flux = photons_cube(rng, 3000)
    q = queries_for(rng, pos, 64)
    k = 25
    lines = [str(len(pos))]
    lines += [" ".join(float(v).hex() for v in np.concatenate([pos[i], wi[i], flux[i]]))
              for i in range(len(pos))]
    lines += [str(len(q))] + [" ".join(float(v).hex() for v in row) for row in q] + [str(k)]
    print("dropin cpp knn", flush=True)
    out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = out.stdout.strip().splitlines()
    assert len(rows) == 2 * len(q)
    bidx, br2 = oracle.brute_knn(pos, q, k)
    for i in range(len(q)):
        for row in (rows[i], rows[len(q) + i].lstrip("B")):
            parts = row.split()
            assert int(parts[0], 16) == int(br2[i].view(np.uint32))
            assert [int(x) for x in parts[1:]] == bidx[i].tolist()
