"""Regression test for a GPU memory fault: scalar loads (s_load_*) through a
non-dword-aligned address.

Scalar memory instructions ignore the two low bits of their address
components.  Op records that packed byte fields beside a 32-bit field were
compiled into a ushort/ubyte vector load plus `s_load_dword sX, s[base], 0x2`
from a base two bytes past the record, so the "coefficient offset" came back
as the record's byte fields and the kernel indexed far past its coefficient
array (an illegal memory access on the GPU).  Every record read in a uniform
loop now has 32-bit fields only; this test compiles each HIP source to gfx950
assembly and rejects any scalar load whose immediate offset is not a multiple
of 4.
"""
import os
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "subquokka_dev_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
UNALIGNED = re.compile(r"^\s*s_load_dword\w*\s+s\[?[0-9:]+\]?,\s*s\[[0-9:]+\],\s*0x([0-9a-f]+)\s*$")


def device_asm(src, out_dir):
    # the same per-source flags as the library build
    from subquokka_dev_amd.build import SOURCE_FLAGS
    out = os.path.join(out_dir, src + ".s")
    cmd = [HIPCC, "-std=c++17", "-O3", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "--cuda-device-only", "-S",
           "-o", out, os.path.join(CSRC, src)] + SOURCE_FLAGS.get(src, [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return out


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_no_scalar_load_at_unaligned_offset(tmp_path):
    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        outs = list(ex.map(lambda s: device_asm(s, str(tmp_path)), SOURCES))
    bad = []
    for src, path in zip(SOURCES, outs):
        for i, line in enumerate(open(path)):
            m = UNALIGNED.match(line)
            if m and int(m.group(1), 16) % 4:
                bad.append(f"{src}.s:{i + 1}: {line.strip()}")
    assert not bad, "\n".join(bad[:20])
    assert shutil.which(HIPCC) or os.path.exists(HIPCC)


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_register_pass_state_stays_in_registers(tmp_path):
    # k_pass keeps a wave's 2^KH x 64 vectors and the op's coefficients in
    # VGPRs; an array the compiler cannot split (e.g. one indexed through a
    # selected address) moves to scratch or is promoted to LDS, and every op
    # then pays a memory round trip.  Every instance must use neither.
    path = device_asm("sq_pass.hip", str(tmp_path))
    sizes, name = {}, None
    for line in open(path):
        t = line.split()
        if len(t) >= 2 and t[0] == ".amdhsa_kernel":
            name = t[1]
        elif name and len(t) == 2 and t[0] in (".amdhsa_group_segment_fixed_size",
                                               ".amdhsa_private_segment_fixed_size"):
            sizes.setdefault(name, {})[t[0]] = int(t[1])
    passes = {k: v for k, v in sizes.items() if "k_pass" in k}
    assert len(passes) == 10, sorted(sizes)   # complex128 / complex64, 1-5 register qubits
    bad = {k: v for k, v in passes.items() if any(v.values())}
    assert not bad, bad
