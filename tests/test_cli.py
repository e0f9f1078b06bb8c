"""The quokka CLI and C++ Simulator: the reference's -i/-c interface, INI keys
and runner selection (main.cpp:13-20, common.cpp:5-32, simulator.cpp:88-257).

Like the reference's harness (gate_check_MEM.py + test_util.py), tests write an
ini and a circuit file, run the binary and compare the dumped state.
"""
import os
import subprocess

import numpy as np
import pytest

import subquokka_dev_amd as sq
from helpers import GOLDEN, random_circuit, random_state

QUOKKA = os.path.join(os.path.dirname(sq.native.LIB_PATH), "quokka")


def write_ini(path, **kv):
    lines = ["[system]"] + [f"{k}={v}" for k, v in kv.items()]
    open(path, "w").write("\n".join(lines) + "\n")


def run(args, cwd):
    return subprocess.run([QUOKKA] + args, cwd=cwd, capture_output=True, text=True, timeout=600)


# ---------------------------- CPU: argument and config errors ----------------------------

def test_unknown_option_and_stray_argument(tmp_path):
    r = run(["-x"], tmp_path)
    assert r.returncode == 1 and "[Unknown Option]: x" in r.stdout
    r = run(["foo"], tmp_path)
    assert r.returncode == 1 and "[Non-option Argument]: foo" in r.stdout


def test_missing_ini(tmp_path):
    open(tmp_path / "c.txt", "w").write("H 0\n")
    r = run(["-i", "nope.ini", "-c", "c.txt"], tmp_path)
    assert r.returncode == 1 and "Error parsing config.ini!" in r.stderr


@pytest.mark.parametrize("kv,msg", [
    (dict(total_qbit=4, chunk_qbit=0, runner_type="MEM"), "chunk_qbit"),
    (dict(total_qbit=4, chunk_qbit=3, file_qbit=2, runner_type="MEM"), "total_qbit"),
    (dict(total_qbit=6, chunk_qbit=2, runner_type="RDMA"), "RDMA runner not found"),
    (dict(total_qbit=6, chunk_qbit=2, mpi_qbit=1, runner_type="MPI_MEM"), "mpi_qbit"),
    (dict(total_qbit=6, chunk_qbit=2, runner_type="MEM", precision="c32"), "precision"),
    (dict(total_qbit=6, chunk_qbit=2, runner_type="MEM", fused_k=7), "fused_k"),
    (dict(total_qbit=10, chunk_qbit=2, runner_type="MEM", fused_passes=1), "fused_passes"),
    (dict(total_qbit=10, chunk_qbit=2, runner_type="MEM", precision="c64", fused_passes=2), "fused_passes"),
    (dict(total_qbit=8, chunk_qbit=2, file_qbit=2, runner_type="DirectIO", state_paths="a,b"), "state_paths"),
])
def test_config_errors(tmp_path, kv, msg):
    write_ini(tmp_path / "a.ini", **kv)
    open(tmp_path / "c.txt", "w").write("H 0\n")
    r = run(["-i", "a.ini", "-c", "c.txt"], tmp_path)
    assert r.returncode == 1 and msg in r.stderr, r.stderr


# ---------------------------- GPU: full runs ----------------------------

def dumped(path, n):
    raw = np.fromfile(path, dtype=np.float64)
    assert raw.size == 2 << n
    return raw[0::2] + 1j * raw[1::2]


@pytest.mark.gpu
@pytest.mark.parametrize("runner", ["MEM", "GPU", "MPI_MEM"])
def test_cli_plain_matches_oracle(tmp_path, oracle, runner):
    n = 12
    text = random_circuit(n, 120, seed=3)
    open(tmp_path / "c.txt", "w").write(text)
    write_ini(tmp_path / "a.ini", total_qbit=n, file_qbit=2, chunk_qbit=5, runner_type=runner,
              is_subcircuit=0, init_seed=7, dump_binary="out.bin")
    r = run(["-i", "a.ini", "-c", "c.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "[Config File]: a.ini" in r.stdout and r.stdout.strip().splitlines()[-1].endswith("s")
    s0 = sq.gen_state(n, 7)
    np.testing.assert_allclose(dumped(tmp_path / "out.bin", n), oracle.run(s0, n, text, chunk=5),
                               rtol=0, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("runner", ["DirectIO", "IO"])
def test_cli_state_files_hold_the_result(tmp_path, oracle, odirect_state_dir, runner):
    # runner_type DirectIO / IO with state_paths: the state lives in 2^file_qbit
    # files (setupFiles, simulator.cpp:102-129) and holds the result afterwards
    n = 16
    text = random_circuit(n, 120, seed=5)
    open(tmp_path / "c.txt", "w").write(text)
    sdir = cli_state_dir(tmp_path, runner, odirect_state_dir)
    paths = ",".join(os.path.join(sdir, f"path{i}") for i in range(4))
    write_ini(tmp_path / "a.ini", total_qbit=n, file_qbit=2, chunk_qbit=5, runner_type=runner,
              is_subcircuit=0, init_seed=9, resident_qbit=14, state_paths=paths)
    r = run(["-i", "a.ini", "-c", "c.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert ("DIO Mode" if runner == "DirectIO" else "IO Mode") in r.stdout
    got = np.concatenate([np.fromfile(os.path.join(sdir, f"path{i}"), dtype=np.complex128) for i in range(4)])
    np.testing.assert_allclose(got, oracle.run(sq.gen_state(n, 9), n, text, chunk=5), rtol=0, atol=1e-12)


def cli_state_dir(tmp_path, runner, odirect):
    """Directory for the state files: for DirectIO the first candidate whose
    filesystem accepts O_DIRECT (skip, naming what refused, if none)."""
    if runner != "DirectIO":
        return str(tmp_path / "state")
    d, seen = odirect
    if d is None:
        pytest.skip("no directory on this host accepts O_DIRECT: " + "; ".join(seen))
    print(f"O_DIRECT state files in {d}", flush=True)
    return os.path.join(d, "state")


@pytest.mark.gpu
@pytest.mark.parametrize("runner,files", [("DirectIO", True), ("IO", True), ("IO", False)])
def test_cli_subcircuit_on_the_out_of_core_runner(tmp_path, oracle, odirect_state_dir, runner, files):
    # runner_type DirectIO / IO at the default is_subcircuit=1 (simulator.cpp:149):
    # DIO_Runner::run(subcircuits) (DIO_Runner.cpp:399) on a chunk-12 finder plan
    n = 17
    text = sq.plan(random_circuit(n, 160, seed=6), n, 0, 12)
    assert "VSWAP_2_2" in text
    open(tmp_path / "c.txt", "w").write(text)
    kv = dict(total_qbit=n, file_qbit=2, chunk_qbit=12, runner_type=runner, init_seed=11, resident_qbit=15,
              dump_binary="out.bin")
    if files:
        sdir = cli_state_dir(tmp_path, runner, odirect_state_dir)
        kv["state_paths"] = ",".join(os.path.join(sdir, f"path{i}") for i in range(4))
    write_ini(tmp_path / "a.ini", **kv)
    r = run(["-i", "a.ini", "-c", "c.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert ("DIO Mode" if runner == "DirectIO" else "IO Mode") in r.stdout
    want = oracle.run(sq.gen_state(n, 11), n, text, chunk=12, sub=True)
    np.testing.assert_allclose(dumped(tmp_path / "out.bin", n), want, rtol=0, atol=1e-12)
    if files:
        got = np.concatenate([np.fromfile(os.path.join(sdir, f"path{i}"), dtype=np.complex128) for i in range(4)])
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


@pytest.mark.gpu
def test_cli_subcircuit_fixture_and_text_dump(tmp_path, oracle):
    n = 21
    text = open(os.path.join(GOLDEN, "ref_circuits", "sub_qft21.txt")).read()
    open(tmp_path / "c.txt", "w").write(text)
    write_ini(tmp_path / "a.ini", total_qbit=n, file_qbit=6, chunk_qbit=12, runner_type="MEM",
              is_subcircuit=1, dump_file="res.txt", MPI_testing=1, dump_binary="out.bin")
    r = run(["-i", "a.ini", "-c", "c.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    s = np.zeros(1 << n, dtype=complex)
    s[0] = 1
    want = oracle.run(s, n, text, chunk=12, sub=True)
    np.testing.assert_allclose(dumped(tmp_path / "out.bin", n), want, rtol=0, atol=1e-12)
    # simulator.cpp:699-702: "<first 3 chars>0.txt", re and im on separate lines, 11 decimals
    lines = open(tmp_path / "res0.txt").read().split("\n")
    assert len(lines[0].split(".")[1]) == 11
    assert abs(float(lines[0]) - want[0].real) < 1e-11 and abs(float(lines[1]) - want[0].imag) < 1e-11


@pytest.mark.gpu
def test_cli_c64_precision_key(tmp_path, oracle):
    n = 14
    text = random_circuit(n, 100, seed=8)
    open(tmp_path / "c.txt", "w").write(text)
    write_ini(tmp_path / "a.ini", total_qbit=n, chunk_qbit=4, runner_type="MEM", is_subcircuit=0,
              precision="c64", init_seed=3, dump_binary="out.bin")
    r = run(["-i", "a.ini", "-c", "c.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    want = oracle.run(sq.gen_state(n, 3), n, text)
    assert np.max(np.abs(dumped(tmp_path / "out.bin", n) - want)) < 2e-6


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("c128", 1e-12), ("c64", 2e-6)])
def test_cli_fused_k_blocks(tmp_path, oracle, precision, tol):
    # fused_k=5: the plain circuit is planned into 5-qubit blocks on MFMA
    n = 13
    text = random_circuit(n, 150, seed=12)
    open(tmp_path / "c.txt", "w").write(text)
    write_ini(tmp_path / "a.ini", total_qbit=n, chunk_qbit=4, runner_type="MEM", is_subcircuit=0,
              precision=precision, fused_k=5, init_seed=4, dump_binary="out.bin")
    r = run(["-i", "a.ini", "-c", "c.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    want = oracle.run(sq.gen_state(n, 4), n, text)
    assert np.max(np.abs(dumped(tmp_path / "out.bin", n) - want)) < tol


@pytest.mark.gpu
def test_cli_fused_passes(tmp_path, oracle):
    # fused_passes=1 (complex64): the plain circuit as fused passes
    n = 14
    text = random_circuit(n, 300, seed=13)
    open(tmp_path / "c.txt", "w").write(text)
    write_ini(tmp_path / "a.ini", total_qbit=n, chunk_qbit=4, runner_type="MEM", is_subcircuit=0,
              precision="c64", fused_passes=1, init_seed=5, dump_binary="out.bin")
    r = run(["-i", "a.ini", "-c", "c.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    want = oracle.run(sq.gen_state(n, 5), n, text)
    got = dumped(tmp_path / "out.bin", n)
    assert np.max(np.abs(got - want)) <= 1e-4 * np.max(np.abs(want))


@pytest.mark.gpu
def test_cli_bad_gate_line_exits_1(tmp_path):
    open(tmp_path / "c.txt", "w").write("H 0\nCZ 0 1\n")
    write_ini(tmp_path / "a.ini", total_qbit=6, chunk_qbit=2, runner_type="MEM", is_subcircuit=0)
    r = run(["-i", "a.ini", "-c", "c.txt"], tmp_path)
    assert r.returncode == 1 and "[setGate]: Not implemented yet." in r.stderr
