"""Register and scratch use of the column kernels, from the compiler's gfx950
code-object metadata (the build's own flags).

The forward + VJP kernels at 7 and 8 levels run without scratch: every value
stays in registers (VGPRs and accumulation registers).  Before, they spilled
0.3-0.8 KB per lane, which was 2.9-3.5x the algorithmic HBM traffic of the
call, and builds of those spilling kernels gave wrong results on the GPU after
unrelated code changes.  The forward kernels at 7 and 8 levels fit in
registers too, except the batched forward kernels at 8 levels: they are
built for two waves per SIMD (at most 256 registers a lane), which costs them
a few dwords of scratch a lane and makes the T425 forward 1.45x faster than
one wave per SIMD without scratch.  Their scratch is bounded here instead.  A
change that brings scratch back fails here, on the CPU, before any GPU run.
(Sixteen-level kernels still spill; they are not checked.)
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "jax-gcm_amd", "csrc", "speedy_kernels.hip")

# mangled-name prefixes of the kernels that must not use scratch
NO_SCRATCH = {
    "fused forward + VJP, 7 and 8 levels": re.compile(r"^_ZN6speedy10fwd_kernelILi[78]ELb0ELb1ELb1ELb1E"),
    "batched forward + VJP, 7 and 8 levels": re.compile(r"^_ZN6speedy16fwd_batch_kernelILi[78]E"),
    "forward, 7 and 8 levels": re.compile(r"^_ZN6speedy(10fwd_kernelILi[78]ELb\dELb1ELb0E|21fwd_diag_batch_kernelILi[78]E|"
                                          r"\d+fwd_(only|plain)_batch_kernelILi7E)"),
}

# kernels built for two waves per SIMD: {pattern: largest scratch bytes per lane allowed}
TWO_WAVES = {
    "batched forward, 7 and 8 levels": (re.compile(r"^_ZN6speedy\d+fwd_(only|plain)_batch_kernelILi[78]E"), 64),
}


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


def kernel_resources():
    """{mangled kernel name: (next free VGPR incl. accumulation registers, scratch bytes per lane)}."""
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not found")
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    flags = [f for f in ge.HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    with tempfile.TemporaryDirectory() as d:
        asm = os.path.join(d, "k.s")
        subprocess.run([hipcc] + flags + ["--cuda-device-only", "-S", "-I", os.path.join(ROOT, "include"), "-o", asm,
                                          SRC], check=True, capture_output=True)
        text = open(asm).read()
    out = {}
    for blk in re.finditer(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", text, re.S):
        body = blk.group(2)
        get = lambda k: int(re.search(r"\.amdhsa_" + k + r"\s+(\d+)", body).group(1))
        out[blk.group(1)] = (get("next_free_vgpr"), get("private_segment_fixed_size"))
    return out


@pytest.fixture(scope="module")
def resources():
    return kernel_resources()


@pytest.mark.parametrize("what", list(NO_SCRATCH))
def test_kernels_without_scratch(resources, what):
    pat = NO_SCRATCH[what]
    hits = {k: v for k, v in resources.items() if pat.match(k)}
    assert hits, f"no kernel matched {what}"
    spilling = {k: v for k, v in hits.items() if v[1] != 0}
    assert not spilling, f"{what}: scratch bytes per lane {spilling}"


@pytest.mark.parametrize("what", list(TWO_WAVES))
def test_two_wave_kernels(resources, what):
    pat, max_scratch = TWO_WAVES[what]
    hits = {k: v for k, v in resources.items() if pat.match(k)}
    assert len(hits) == 4, f"{what}: {sorted(hits)}"
    for k, (vgpr, scratch) in hits.items():
        assert vgpr <= 256, f"{k}: {vgpr} registers a lane, more than two waves per SIMD allow"
        assert scratch <= max_scratch, f"{k}: scratch {scratch} bytes per lane"
