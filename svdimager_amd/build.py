"""Build libsvdimager.so (HIP, gfx950) and the sharedlibrary_gpu CLI in-tree.

Every .hip / .cpp under csrc/ is compiled with hipcc for gfx950 with packed-FP32 instructions
disabled; fits.cpp, msreader.cpp and the two CLIs are plain host C++.  Outputs:
svdimager_amd/libsvdimager.so, svdimager_amd/bin/sharedlibrary_gpu and
svdimager_amd/bin/svd_integrated_imager_gpu.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libsvdimager.so")
BIN = os.path.join(PKG, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SVDI_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["preprocess.hip", "plan.hip", "grid.hip", "tiles.hip", "windows.hip", "fft.hip", "image.hip", "api.cpp",
               "service.cpp"]
HOST_SOURCES = ["fits.cpp", "msreader.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src in HIP_SOURCES:
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", path, "-o", out])
    else:
        _run(["g++", *HOST_FLAGS, "-c", path, "-o", out])
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    sources = HIP_SOURCES + HOST_SOURCES
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    clis = []
    for name, src in (("sharedlibrary_gpu", "cli_sharedlibrary_gpu.cpp"),
                      ("svd_integrated_imager_gpu", "cli_integrated_imager_gpu.cpp")):
        cli = os.path.join(BIN, name)
        _run(["g++", *HOST_FLAGS, os.path.join(CSRC, src), "-o", cli, f"-L{PKG}", "-lsvdimager",
              "-Wl,-rpath,$ORIGIN/.."])
        clis.append(cli)
    if verbose:
        print(f"built {LIB} and {' '.join(clis)}")
    return LIB


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)
    for p in (LIB, os.path.join(BIN, "sharedlibrary_gpu"), os.path.join(BIN, "svd_integrated_imager_gpu")):
        if os.path.exists(p):
            os.remove(p)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "clean":
        clean()
    else:
        build(verbose=True)
