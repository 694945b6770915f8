"""Recipe for the test-only CPU oracle: gcc on wfpt_oracle.c -> liboracle.so.

The reference (GLSL compute shaders + Vulkan + glm) cannot be compiled in this
image: there is no Vulkan loader, no glslangValidator and the glm submodule is
empty, so there is no oracle/_ref build; the restatement here is the oracle.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "wfpt_oracle.c"
LIB = HERE / "liboracle.so"


def build(verbose: bool = True) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           "-fexcess-precision=standard", "-fopenmp", "-o", str(tmp), str(SRC), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stderr}")
    tmp.replace(LIB)
    if verbose:
        print(f"[build] {LIB}", flush=True)
    return LIB


if __name__ == "__main__":
    build()
