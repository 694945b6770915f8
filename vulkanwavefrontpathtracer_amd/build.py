"""Builds libwfpt.so (HIP kernels + C ABI + C++ host runtime) for gfx950.

Every translation unit under csrc/ is compiled by hipcc into an object and the
objects are linked into one shared library next to this file.  The oracle
(test-only CPU restatement) has its own recipe under oracle/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libwfpt.so"
ROOT = PKG.parent

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -ffp-contract=off: fused multiply-adds only where written as fmaf(), so the
# CPU restatement reproduces the arithmetic.  Packed fp32 instructions are
# disabled on this target.
COMMON_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]


def sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *COMMON_FLAGS, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [HIPCC, *COMMON_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr}")
    return obj


TORCH_SRC = PKG / "torch_ext" / "render_autograd.cpp"
TORCH_LIB = PKG / "_wfpt_torch.so"


def build_torch_ext(verbose: bool = True) -> Path:
    """The libtorch autograd Function (torch_ext/render_autograd.cpp) as the
    Python module _wfpt_torch, linked against libwfpt.so next to it."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce

    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    inc = [f"-I{p}" for p in ce.include_paths()] + [f"-I{sysconfig.get_paths()['include']}",
                                                    f"-I{ROOT / 'include'}"]
    libdirs = ce.library_paths()
    tmp = TORCH_LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           "-DTORCH_EXTENSION_NAME=_wfpt_torch", "-DTORCH_API_INCLUDE_EXTENSION_H", *inc,
           str(TORCH_SRC), "-o", str(tmp), *[f"-L{d}" for d in libdirs], f"-L{PKG}",
           "-lwfpt", "-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python",
           "-Wl,-rpath,$ORIGIN", *[f"-Wl,-rpath,{d}" for d in libdirs]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"torch extension build failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, TORCH_LIB)
    if verbose:
        print(f"[build] {TORCH_LIB}", flush=True)
    return TORCH_LIB


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"[build] {LIB} ({len(objs)} objects)", flush=True)
    return LIB


if __name__ == "__main__":
    build()
    build_torch_ext()
    sys.exit(0)
