"""Memory safety of the engine's launches, checked on the CPU.

tests/cpu_emul/build.sh compiles the engine sources for the host with a CPU
stand-in for the HIP runtime and AddressSanitizer; every launch then runs
block by block and any out-of-bounds access aborts the process.  This covers
the subcircuit sweep (with the LDS tile pass turned on) and the VSWAP
permutations, whose GPU tests are skipped (see test_gpu_parity.py).
"""
import os
import shutil
import subprocess

import pytest

import subquokka_dev_amd as sq
from helpers import GOLDEN, ROOT, random_circuit

EMUL_DIR = os.path.join(ROOT, "tests", "cpu_emul")
EMUL = os.path.join(EMUL_DIR, "emul")
CLANG = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not os.path.exists(CLANG), reason="clang++ for the emulation build not found")


@pytest.fixture(scope="module")
def emul():
    subprocess.run(["bash", os.path.join(EMUL_DIR, "build.sh")], check=True)
    return EMUL


def run_case(emul, tmp_path, mode, n, chunk, dtype, text):
    f = tmp_path / "c.txt"
    f.write_text(text)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([emul, mode, str(n), str(chunk), dtype, str(f)], capture_output=True, text=True,
                       timeout=900, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert r.stdout.startswith("ok")


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_sub_fixture_qft21_tiles(emul, tmp_path, dtype):
    text = open(os.path.join(GOLDEN, "ref_circuits", "sub_qft21.txt")).read()
    run_case(emul, tmp_path, "sub", 21, 12, dtype, text)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_planned_blocks_tiles(emul, tmp_path, dtype):
    n = 14
    text = sq.plan(random_circuit(n, 200, seed=3), n, 1, 5)
    run_case(emul, tmp_path, "sub", n, 12, dtype, text)


def test_wide_block_split(emul, tmp_path):
    n = 16
    text = random_circuit(n, 150, seed=5)
    run_case(emul, tmp_path, "sub", n, 12, "c128", f"{len(text.strip().splitlines())}\n" + text)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_finder_plan_with_vswaps(emul, tmp_path, dtype):
    n = 16
    text = sq.plan(random_circuit(n, 150, seed=7), n, 3, 10)
    assert "VSWAP_2_2" in text
    run_case(emul, tmp_path, "sub", n, 10, dtype, text)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [4, 9, 13])
def test_plain_all_gate_kinds(emul, tmp_path, dtype, n):
    text = random_circuit(n, 120, seed=n) + "VSWAP_1_1 0 %d\nVSWAP_2_2 0 1 %d %d\n" % (n - 1, n - 2, n - 1)
    run_case(emul, tmp_path, "plain", n, n, dtype, text)
