"""The forward + VJP kernels are pinned to the gfx950 code that passed the GPU suite.

These kernels (fwd_kernel<KX, DIAG, FUSED, VJP=true, EXTRA=true> and
fwd_batch_kernel<KX>) spill several hundred registers per lane.  Builds of them
with a different register allocation (-O3, and -O2 builds after unrelated
changes to shared code) have given wrong gradients or tendencies on the GPU,
while the same source passed every host check.  So changes elsewhere in the
library are made such that these kernels compile to the same code, and this test
checks that: it compiles the library's device code to assembly with the build's
flags and compares a hash of each of these functions (block labels normalised,
comments dropped) with the pinned one in golden/vjp_kernels_isa.json.

A deliberate change to these kernels needs the GPU VJP tests to pass, then a new
pin: ``python tests/test_vjp_isa_pin.py --write``.  The pin names the compiler it
was made with; under another compiler the test is skipped."""
import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PIN = os.path.join(ROOT, "tests", "golden", "vjp_kernels_isa.json")
SRC = os.path.join(ROOT, "jax-gcm_amd", "csrc", "speedy_kernels.hip")
PINNED = re.compile(r"^_ZN6speedy(10fwd_kernelILi\d+ELb\dELb\dELb1ELb1E|16fwd_batch_kernelILi\d+E)\S*$")


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


def _compiler_id(hipcc):
    out = subprocess.run([hipcc, "--version"], capture_output=True, text=True).stdout
    return " | ".join(l.strip() for l in out.splitlines() if "version" in l.lower())


def _device_flags():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    return [f for f in ge.HIPCC_FLAGS if f not in ("-shared", "-fPIC")]


def isa_hashes(hipcc):
    """sha256 of each pinned kernel's gfx950 assembly (labels normalised)."""
    with tempfile.TemporaryDirectory() as d:
        asm = os.path.join(d, "k.s")
        subprocess.run([hipcc] + _device_flags() + ["--cuda-device-only", "-S", "-I", os.path.join(ROOT, "include"),
                                                    "-o", asm, SRC], check=True, capture_output=True)
        lines = open(asm).read().split("\n")
    out, name, body = {}, None, []
    for line in lines:
        m = re.match(r"^(_Z\S+):", line)
        if m:
            name, body = m.group(1), []
            continue
        if name and line.startswith(".Lfunc_end"):
            if PINNED.match(name):
                norm = [re.sub(r"\.LBB\d+_", ".LBB_", x) for x in body if not x.strip().startswith(";")]
                out[name] = hashlib.sha256("\n".join(norm).encode()).hexdigest()
            name = None
            continue
        if name:
            body.append(line.split(";")[0].rstrip())
    return out


def test_vjp_kernels_compile_to_the_pinned_code():
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not found")
    pin = json.load(open(PIN))
    if pin["compiler"] != _compiler_id(hipcc):
        pytest.skip(f"pinned with another compiler: {pin['compiler']}")
    got = isa_hashes(hipcc)
    assert sorted(got) == sorted(pin["sha256"]), "the set of forward + VJP kernels changed"
    changed = [k for k in got if got[k] != pin["sha256"][k]]
    assert not changed, f"forward + VJP kernels compile to different code: {changed}"


if __name__ == "__main__":
    if "--write" not in sys.argv:
        sys.exit("usage: python tests/test_vjp_isa_pin.py --write   (after the GPU VJP tests pass)")
    hipcc = _hipcc()
    pin = {"compiler": _compiler_id(hipcc), "sha256": isa_hashes(hipcc)}
    json.dump(pin, open(PIN, "w"), indent=1, sort_keys=True)
    open(PIN, "a").write("\n")
    print(f"pinned {len(pin['sha256'])} kernels")
