"""Build libneon_amd.so (HIP kernels for gfx950 + the C ABI) in-tree.

    python neon_amd/build.py            # compile every source, link the library

Each translation unit is compiled separately (in parallel) with hipcc and the
objects are linked into neon_amd/lib/libneon_amd.so; the programs under
neon_amd/apps/ (on include/neon_amd.hpp) are then built into neon_amd/lib/:
*.cpp (host code) with g++, *.hip (a caller's own device containers) with
hipcc and the library's device flags.  Flags:
  * --offload-arch=gfx950       MI355X only
  * -ffp-contract=off           no FMA contraction: binary64 kernels match a CPU
                                evaluation of the same expression tree bit for bit
  * -Xclang -target-feature -Xclang -packed-fp32-ops
                                no packed FP32 (v_pk_*_f32) instructions
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parent
REPO = ROOT.parent
CSRC = ROOT / "csrc"
OBJ = ROOT / "_build"
LIB = ROOT / "lib" / "libneon_amd.so"
APPS = ROOT / "apps"     # C++ programs on include/neon_amd.hpp, linked against LIB

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    f"-I{REPO / 'include'}", f"-I{CSRC}",
    "-Wno-unused-result",
]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources() -> list[pathlib.Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: pathlib.Path) -> pathlib.Path:
    out = OBJ / (src.stem + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, "-x", "hip", *DEVICE, *COMMON, "-c", str(src), "-o", str(out)]
    else:
        cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = True) -> pathlib.Path:
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJ.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    for app in sorted(list(APPS.glob("*.cpp")) + list(APPS.glob("*.hip"))):
        exe = LIB.parent / app.stem
        if app.suffix == ".hip":
            # a caller's own containers: device code from include/neon_amd.hpp
            cmd = [HIPCC, "-x", "hip", *DEVICE, "-O3", "-std=c++17", "-ffp-contract=off", "-Wall",
                   f"-I{REPO / 'include'}", str(app), "-o", str(exe) + ".tmp",
                   f"-L{LIB.parent}", "-lneon_amd", "-Wl,-rpath,$ORIGIN"]
        else:
            cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{REPO / 'include'}", str(app), "-o", str(exe) + ".tmp",
                   f"-L{LIB.parent}", "-lneon_amd", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"app build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(str(exe) + ".tmp", exe)
        if verbose:
            print(f"built {exe}")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
