"""Build libspi_hip.so (gfx950) in-tree with hipcc.

The library lands in ``spi-active_amd/lib/``; it is git-ignored but travels with the
source tree.  ``build_library()`` always recompiles (it is the "does it build" check);
loaders call ``ensure_built()`` which compiles only when the library is missing.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
from typing import List

from .model import generate_header, load_model

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))   # spi-active_amd/
REPO_ROOT = os.path.dirname(PKG_ROOT)
CSRC = os.path.join(PKG_ROOT, "csrc")
INCLUDE = os.path.join(REPO_ROOT, "include")
LIB_DIR = os.path.join(PKG_ROOT, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libspi_hip.so")
SOURCES = ["spi_kernels.hip", "spi_cem.hip", "spi_sens.hip", "spi_sens_lane.hip", "spi_sens_pair.hip",
           "spi_fisher.hip",
           "spi_commands.hip", "spi_data.hip", "spi_adjoint.hip", "spi_adjoint_lane.hip", "spi_policy.hip",
           "spi_mlp.hip"]
# Per-file flags: the dynamics kernels (primal and tangent) may assume finite arithmetic (x*0 folds away; their NaN
# and Inf guards test exponent bits).  The primal, policy, four-lane tangent, tangent-pair and
# four-lane adjoint sources may also reassociate floating-point sums, which lets the compiler turn separate products and sums into
# fused multiply-adds (primal: 3,748 -> 3,454 instructions per articulated-body sub-interval,
# tangent pairs: 6,718 -> 5,867 per kernel); the four-lane sums pin their operands so that
# reassociation cannot make them lane-dependent.  The one-lane adjoint and tangent kernels
# (spi_adjoint_lane.hip, spi_sens_lane.hip) do without it: they already spill, and
# reassociation keeps more values live (one-lane adjoint 461 -> 540 ms).
# CEM, the cost reduction and the Fisher/logdet kernels keep strict IEEE semantics.
DYN_FLAGS = ["-ffinite-math-only"]
DYN_REASSOC_FLAGS = DYN_FLAGS + ["-fassociative-math", "-DSPI_REASSOCIATE=1"]
FILE_FLAGS = {"spi_kernels.hip": DYN_REASSOC_FLAGS, "spi_cem.hip": [], "spi_sens.hip": DYN_REASSOC_FLAGS,
              "spi_sens_lane.hip": DYN_FLAGS,
              "spi_sens_pair.hip": DYN_REASSOC_FLAGS,
              "spi_fisher.hip": [], "spi_commands.hip": [], "spi_data.hip": [],
              "spi_adjoint.hip": DYN_REASSOC_FLAGS, "spi_adjoint_lane.hip": DYN_FLAGS,
              "spi_policy.hip": DYN_REASSOC_FLAGS,
              "spi_mlp.hip": []}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in generated code.
# -fno-signed-zeros: lets x*0 / x+0 with compile-time zeros of the model fold away.
CFLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-I", INCLUDE, "-I", CSRC,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fno-signed-zeros",
          "-Wall", "-Wno-unused-function"]


def write_model_header() -> str:
    """Regenerate the compile-time model header from the committed model file."""
    path = os.path.join(CSRC, "spi_go2_model.h")
    text = generate_header(load_model())
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)
    return path


def _compile(src: str, obj: str, extra: List[str]) -> None:
    cmd = [HIPCC] + CFLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")


def build_library(extra_flags: List[str] = (), verbose: bool = False) -> str:
    write_model_header()
    os.makedirs(LIB_DIR, exist_ok=True)
    objs = [os.path.join(LIB_DIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        list(ex.map(lambda so: _compile(so[0], so[1], FILE_FLAGS[so[0]] + list(extra_flags)), zip(SOURCES, objs)))
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc link failed:\n{r.stderr}")
    os.replace(tmp, LIB_PATH)
    for o in objs:
        os.remove(o)
    if verbose:
        print("built", LIB_PATH)
    return LIB_PATH


def ensure_built() -> str:
    if not os.path.exists(LIB_PATH):
        build_library()
    return LIB_PATH


if __name__ == "__main__":
    build_library(verbose=True)
