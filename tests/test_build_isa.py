"""The four-lane sums stay butterflies in the build.  The sources compiled with reassociation
(-fassociative-math) pin the operands of their quad sums; this compiles each of them to gfx950
assembly with the build's own flags and checks that every DPP instruction is a butterfly step
``v_add_f32_dpp vD, vS, vS quad_perm:...`` — the lane's value plus its partner's value of the
same register — and never a separate DPP move whose result the compiler could add in a
lane-dependent order (which is what reassociation does to unpinned sums).  CPU only: needs
hipcc, not a GPU."""
import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from spi_amd import build

REASSOC = [src for src, flags in build.FILE_FLAGS.items() if "-fassociative-math" in flags]
DPP_ADD = re.compile(r"^\s*v_add_f32_dpp\s+(v\d+),\s*(v\d+),\s*(v\d+)\s+quad_perm:\[(1,0,3,2|2,3,0,1)\]")


@pytest.fixture(scope="module")
def device_asm(tmp_path_factory):
    if not os.path.exists(build.HIPCC):
        pytest.skip("no hipcc")
    out_dir = tmp_path_factory.mktemp("isa")

    def one(src):
        out = out_dir / (src + ".s")
        cmd = ([build.HIPCC] + build.CFLAGS + build.FILE_FLAGS[src] +
               ["--cuda-device-only", "-S", os.path.join(build.CSRC, src), "-o", str(out)])
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        return src, out.read_text()

    with ThreadPoolExecutor(max_workers=min(4, len(REASSOC))) as ex:
        return dict(ex.map(one, REASSOC))


def test_reassociated_sources_are_the_leg_parallel_ones():
    assert set(REASSOC) >= {"spi_kernels.hip", "spi_sens_pair.hip", "spi_adjoint.hip"}
    for src in REASSOC:
        assert "-DSPI_REASSOCIATE=1" in build.FILE_FLAGS[src]    # turns the operand pins on


@pytest.mark.parametrize("src", REASSOC)
def test_quad_sums_are_butterflies(device_asm, src):
    dpp = [ln for ln in device_asm[src].splitlines() if "_dpp" in ln and not ln.lstrip().startswith(";")]
    assert dpp, "no four-lane sums found"
    for ln in dpp:
        m = DPP_ADD.match(ln)
        assert m, f"{src}: not a butterfly step: {ln.strip()}"
        assert m.group(2) == m.group(3), f"{src}: operands differ: {ln.strip()}"
    n1 = sum("quad_perm:[1,0,3,2]" in ln for ln in dpp)
    n2 = sum("quad_perm:[2,3,0,1]" in ln for ln in dpp)
    assert n1 == n2                                  # every sum has both stages


def test_finiteness_tests_survive_finite_math(tmp_path):
    """The dynamics sources are built with -ffinite-math-only, under which the compiler may
    assume that no float is NaN or infinite and fold a bit test of one to "finite" (a NaN
    candidate would then escape the divergence checks).  spi_launch::finite_bits and
    below_limit pass the bits through an empty asm: compiled with the dynamics' flags, a
    kernel that tests a loaded value must still mask the exponent and compare it."""
    if not os.path.exists(build.HIPCC):
        pytest.skip("no hipcc")
    src = tmp_path / "probe.hip"
    src.write_text(
        '#include <hip/hip_runtime.h>\n#include "spi_launch.h"\n#include "spi_dyn.h"\n'
        "__global__ void probe_bits(const float* __restrict__ x, int* __restrict__ o) {\n"
        "  o[threadIdx.x] = spi_launch::finite_bits(x[threadIdx.x] * 2.0f + 1.0f); }\n"
        "__global__ void probe_limit(const float* __restrict__ x, int* __restrict__ o) {\n"
        "  o[threadIdx.x] = spi::below_limit(x[threadIdx.x] * 2.0f + 1.0f, 1e4f); }\n")
    out = tmp_path / "probe.s"
    flags = build.FILE_FLAGS["spi_kernels.hip"]
    assert "-ffinite-math-only" in flags
    cmd = [build.HIPCC] + build.CFLAGS + flags + ["--cuda-device-only", "-S", str(src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    asm = out.read_text()
    for fn in ("probe_bits", "probe_limit"):
        body = asm[asm.index(f"{fn}PKfPi:"):]
        body = body[:body.index("s_endpgm")]
        assert "0x7f800000" in body and "v_cmp" in body, f"{fn}: the finiteness test was folded away"
