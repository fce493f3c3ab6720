"""Build libpymdp_hip.so for gfx950 with hipcc.

The MI355X counterpart of the reference's scripts/ffi_build.py (which drives
CMake, pymdp/ffi/CMakeLists.txt).  Here the library is a handful of HIP
translation units compiled straight into one shared object, in-tree, so the
built file travels with the source tree.

    python -m pymdp_amd.build            # build (always recompiles)
    python -m pymdp_amd.build --check    # print the output path only
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
LIB_DIR = PKG_DIR / "_lib"
LIB_NAME = "libpymdp_hip.so"
OFFLOAD_ARCH = os.environ.get("PYMDP_HIP_ARCH", "gfx950")

SOURCES = ["fpi.hip", "efe.hip", "ragged.hip", "adjoint.hip", "learning.hip", "vfe.hip", "seq.hip", "scratch.hip",
           "history.cpp"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={OFFLOAD_ARCH}",
    # Packed FP32 instructions are disabled for all device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    # MFMA accumulators in ordinary VGPRs: with the default AGPR form the
    # compiler copies a loop-carried accumulator AGPR <-> VGPR around every
    # MFMA of a k loop (sixteen moves and a wait per step).
    "-mllvm", "-amdgpu-mfma-vgpr-form=true",
    "-I", str(REPO_DIR / "include"),
    "-I", str(CSRC),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def lib_path() -> Path:
    return LIB_DIR / LIB_NAME


def _compile(src: str, obj_dir: Path) -> Path:
    s = CSRC / src
    o = obj_dir / (src + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(s), "-o", str(o)]
    if src.endswith(".cpp"):
        cmd = [hipcc(), "-O3", "-std=c++17", "-fPIC", "-x", "c++", "-I", str(REPO_DIR / "include"),
               "-I", str(CSRC), "-c", str(s), "-o", str(o)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return o


def build(verbose: bool = False) -> Path:
    LIB_DIR.mkdir(exist_ok=True)
    obj_dir = LIB_DIR / "_obj"
    obj_dir.mkdir(exist_ok=True)
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, obj_dir), srcs))
    out = lib_path()
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={OFFLOAD_ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true")
    a = ap.parse_args(argv)
    if a.check:
        print(lib_path())
        return 0
    build(verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
