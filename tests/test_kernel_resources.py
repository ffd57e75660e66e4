"""Register and scratch use of the column kernels, from the compiler's gfx950
code-object metadata (the build's own flags).

The fused forward + VJP kernels at 7 and 8 levels run without scratch: every
value stays in registers (VGPRs and accumulation registers).  Before, they
spilled 0.3-0.8 KB per lane, which was 2.9-3.5x the algorithmic HBM traffic of
the call, and builds of those spilling kernels gave wrong results on the GPU
after unrelated code changes.  The fused and diagnostics forward kernels at 7
and 8 levels fit in registers too.

The batched kernels at 7 and 8 levels (forward-only and forward + VJP) are
built for two waves per SIMD instead, at most 256 registers a lane: the
forward-only ones without scratch, the forward + VJP one with some (132 B a
lane at 8 levels, none at 7), which makes it 1.26x faster at T425 than one
wave per SIMD without scratch.  Their register count and scratch are bounded here.  A change that brings scratch back or raises it fails
here, on the CPU, before any GPU run.  (Sixteen-level kernels still spill; they
are not checked.)
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
    "forward, 7 and 8 levels": re.compile(r"^_ZN6speedy(10fwd_kernelILi[78]ELb\dELb1ELb0E|21fwd_diag_batch_kernelILi[78]E|"
                                          r"\d+fwd_(only|plain)_batch_kernelILi7E)"),
}

# kernels built for two waves per SIMD: {pattern: largest scratch bytes per lane allowed}
TWO_WAVES = {
    "batched forward, 7 and 8 levels": (re.compile(r"^_ZN6speedy\d+fwd_(only|plain)_batch_kernelILi[78]E"), 0, 4),
    "batched forward + VJP, 7 and 8 levels": (re.compile(r"^_ZN6speedy16fwd_batch_kernelILi[78]E"), 160, 2),
}


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


def _device_asm(src):
    """gfx950 assembly of one HIP source, compiled with the build's flags."""
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not found")
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    flags = [f for f in ge.HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    with tempfile.TemporaryDirectory() as d:
        asm = os.path.join(d, "k.s")
        subprocess.run([hipcc] + flags + ["--cuda-device-only", "-S", "-I", os.path.join(ROOT, "include"), "-o", asm,
                                          src], check=True, capture_output=True)
        return open(asm).read()


def kernel_resources(text):
    """{mangled kernel name: (next free VGPR incl. accumulation registers, scratch bytes per lane)}."""
    out = {}
    for blk in re.finditer(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", text, re.S):
        body = blk.group(2)
        get = lambda k: int(re.search(r"\.amdhsa_" + k + r"\s+(\d+)", body).group(1))
        out[blk.group(1)] = (get("next_free_vgpr"), get("private_segment_fixed_size"))
    return out


@pytest.fixture(scope="module")
def kernels_asm():
    return _device_asm(SRC)


@pytest.fixture(scope="module")
def resources(kernels_asm):
    return kernel_resources(kernels_asm)


@pytest.mark.parametrize("what", list(NO_SCRATCH))
def test_kernels_without_scratch(resources, what):
    pat = NO_SCRATCH[what]
    hits = {k: v for k, v in resources.items() if pat.match(k)}
    assert hits, f"no kernel matched {what}"
    spilling = {k: v for k, v in hits.items() if v[1] != 0}
    assert not spilling, f"{what}: scratch bytes per lane {spilling}"


@pytest.mark.parametrize("what", list(TWO_WAVES))
def test_two_wave_kernels(resources, what):
    pat, max_scratch, count = TWO_WAVES[what]
    hits = {k: v for k, v in resources.items() if pat.match(k)}
    assert len(hits) == count, f"{what}: {sorted(hits)}"
    for k, (vgpr, scratch) in hits.items():
        assert vgpr <= 256, f"{k}: {vgpr} registers a lane, more than two waves per SIMD allow"
        assert scratch <= max_scratch, f"{k}: scratch {scratch} bytes per lane"


def _valu_counts(asm_text):
    """{mangled kernel name: static count of VALU instructions} of an assembly file."""
    out = {}
    for m in re.finditer(r"^(_Z\w+):.*\n", asm_text, re.M):
        end = asm_text.find(".Lfunc_end", m.end())
        if end < 0:
            continue
        ops = [ln.split()[0] for ln in (x.strip() for x in asm_text[m.end():end].split("\n"))
               if ln and not ln.startswith((".", ";")) and not ln.endswith(":")]
        out[m.group(1)] = sum(1 for o in ops if o.startswith("v_"))
    return out


def test_term_census(kernels_asm):
    """Static VALU instructions per term of the 8-level column code (gfx950,
    the build's flags), from tests/host_column/term_census.hip: each term's kernel
    minus the loads and stores alone.  Printed as a table; the terms together
    must account for most of the fused forward kernel's instructions (the rest
    is addressing, the level-set exchange and the output sums)."""
    census = _valu_counts(_device_asm(os.path.join(ROOT, "tests", "host_column", "term_census.hip")))
    kernels = _valu_counts(kernels_asm)
    terms = {re.sub(r"^_Z\d+k_(\w+?)4Args.*", r"\1", k): v for k, v in census.items()}
    base = terms.pop("copy")
    terms = {k: v - base for k, v in terms.items()}
    assert len(terms) == 12 and all(v >= 0 for v in terms.values()), terms
    fused = kernels["_ZN6speedy10fwd_kernelILi8ELb0ELb1ELb0ELb0EEEvNS_5KArgsE"]
    total = sum(terms.values())
    for k, v in sorted(terms.items(), key=lambda kv: -kv[1]):
        print(f"{k:18s} {v:5d}  {100.0 * v / fused:5.1f} %")
    print(f"{'terms':18s} {total:5d}  {100.0 * total / fused:5.1f} % of fwd_kernel<8> ({fused} VALU)")
    assert 0.75 * fused <= total <= fused
