"""Build recipe for libpm_gather.so (gfx950) and the CPU oracle.

``build_library()`` compiles every HIP translation unit under ``csrc/`` with
hipcc for gfx950 and links them into ``photon_mapping_amd/libpm_gather.so``
(in-tree, so the built library travels with the source tree).
``build_oracle()`` runs ``make`` in ``oracle/`` (test infrastructure only).
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
INCLUDE = REPO_DIR / "include"
BUILD_DIR = PKG_DIR / "_build"
LIB_PATH = PKG_DIR / "libpm_gather.so"
ORACLE_DIR = REPO_DIR / "oracle"

HIP_SOURCES = ["pm_capi.hip", "pm_build.hip", "pm_sort.hip", "pm_gather.hip",
               "pm_scene.hip", "pm_bvh.hip", "pm_render.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # d^2 and the sign tests must round exactly like the reference's
    # non-FMA x86-64 build; denormals must be kept.
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    # packed FP32 instructions are disabled for this device target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libpm_gather.so)")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "is not a recognized feature for this target" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def build_library(verbose: bool = False, out: Path | None = None) -> Path:
    """Builds the library (to ``out``, default photon_mapping_amd/libpm_gather.so).
    PM_HIPCC_DEFINES (space-separated -D flags) passes tuning macros such as
    -DPM_GATHER_MIN_WAVES=3 to every translation unit."""
    hipcc = _hipcc()
    defines = os.environ.get("PM_HIPCC_DEFINES", "").split()
    lib_path = Path(out) if out else LIB_PATH
    BUILD_DIR.mkdir(exist_ok=True)
    sources = [CSRC / s for s in HIP_SOURCES]
    objects = []
    cmds = []
    for src in sources:
        obj = BUILD_DIR / (src.stem + ".o")
        objects.append(obj)
        cmds.append([hipcc, *HIPCC_FLAGS, *defines, f"-I{INCLUDE}", f"-I{CSRC}", "-c",
                     str(src), "-o", str(obj)])
    workers = min(len(cmds), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for cmd, fut in zip(cmds, [ex.submit(_run, c) for c in cmds]):
            if verbose:
                print(" ".join(cmd), flush=True)
            fut.result()
    tmp = lib_path.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
          *[str(o) for o in objects]])
    os.replace(tmp, lib_path)
    return lib_path


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ORACLE_DIR)])
    return ORACLE_DIR / "libpm_oracle.so"


if __name__ == "__main__":
    v = "-v" in sys.argv
    print(build_library(verbose=v))
    print(build_oracle())
