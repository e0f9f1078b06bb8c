"""Out-of-core runner (DIO_Runner role): state in pinned host memory, streamed
through HBM pass by pass; results must equal the HBM-resident engine."""
import numpy as np
import pytest

import subquokka_dev_amd as sq
from helpers import GOLDEN, high_qubit_subcircuit, random_circuit, random_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,w", [(12, 9), (16, 12), (20, 14)])
def test_out_of_core_matches_oracle(oracle, dtype, n, w):
    text = sq.gen_circuit("random", n, 3, seed=n) + random_circuit(n, 80, seed=w)
    s0 = random_state(n, w)
    eng = sq.Engine(n, dtype=dtype, resident_qbit=w)
    try:
        eng.set_state(s0.astype(np.complex128 if dtype == "c128" else np.complex64))
        eng.apply_text(text)
        got = eng.get_state()
        norm = eng.norm2()
    finally:
        eng.close()
    want = oracle.run(s0, n, text)
    tol = 1e-12 if dtype == "c128" else 2e-6
    assert np.max(np.abs(got - want)) <= tol
    # H uses the reference's literal 0.70710678118, so the norm drifts from 1
    assert abs(norm - np.vdot(want, want).real) < (1e-12 if dtype == "c128" else 1e-5)


def test_out_of_core_24q_matches_hbm_engine():
    # 24 qubits streamed as 2^20-amplitude groups vs the same circuit in HBM
    n, w = 24, 20
    text = sq.gen_circuit("random", n, 6, seed=5)
    a = sq.Engine(n, dtype="c128", resident_qbit=w)
    b = sq.Engine(n, dtype="c128")
    try:
        a.apply_text(text)
        b.apply_text(text)
        assert np.max(np.abs(a.get_state() - b.get_state())) < 1e-12
        assert abs(a.expect_zz_ring() - b.expect_zz_ring()) < 1e-10
    finally:
        a.close()
        b.close()


def test_out_of_core_rejects_hbm_only_entry_points():
    n, w = 12, 9
    eng = sq.Engine(n, resident_qbit=w)
    try:
        with pytest.raises(sq.SqError):
            eng.device_state()
        with pytest.raises(sq.SqError):
            eng.hea_energy(1, np.zeros(2 * n))
    finally:
        eng.close()


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_pageable_tier_matches_oracle(oracle, dtype):
    # state in ordinary host memory, groups packed through pinned staging
    n, w = 16, 12
    text = sq.gen_circuit("random", n, 3, seed=2) + random_circuit(n, 100, seed=21)
    s0 = random_state(n, 5)
    eng = sq.Engine(n, dtype=dtype, resident_qbit=w, host_tier="pageable")
    try:
        eng.set_state(s0.astype(np.complex128 if dtype == "c128" else np.complex64))
        eng.apply_text(text)
        got = eng.get_state()
        passes, _, tier = eng.stream_info()
    finally:
        eng.close()
    assert tier == "pageable" and passes >= 1
    tol = 1e-12 if dtype == "c128" else 2e-6
    assert np.max(np.abs(got - oracle.run(s0, n, text))) <= tol


def read_state_files(paths, dtype=np.complex128):
    return np.concatenate([np.fromfile(p, dtype=dtype) for p in paths])


def state_dir(tmp_path, direct, odirect):
    """tmp_path, or for O_DIRECT the first candidate directory whose
    filesystem accepts it (the test is skipped, naming what refused, if none)."""
    if not direct:
        return tmp_path
    d, seen = odirect
    if d is None:
        pytest.skip("no directory on this host accepts O_DIRECT: " + "; ".join(seen))
    print(f"O_DIRECT state files in {d}", flush=True)
    import pathlib
    return pathlib.Path(d)


@pytest.mark.parametrize("direct", [False, True])
def test_file_tier_matches_oracle_and_resumes(oracle, tmp_path, odirect_state_dir, direct):
    # state in 4 files of the reference's layout; a second engine reopens the
    # files and continues the circuit from them
    n, w = 16, 14
    d = state_dir(tmp_path, direct, odirect_state_dir)
    paths = [str(d / f"path{i}") for i in range(4)]
    t1 = sq.gen_circuit("random", n, 2, seed=3) + random_circuit(n, 60, seed=31)
    t2 = random_circuit(n, 60, seed=32)
    s0 = random_state(n, 6)
    eng = sq.Engine(n, resident_qbit=w, state_files=paths, direct_io=direct)
    try:
        eng.set_state(s0)
        eng.apply_text(t1)
        assert eng.stream_info()[2] == "files"
    finally:
        eng.close()
    mid = oracle.run(s0, n, t1)
    assert np.max(np.abs(read_state_files(paths) - mid)) <= 1e-12
    eng = sq.Engine(n, resident_qbit=w, state_files=paths, direct_io=direct, keep_files=True)
    try:
        eng.apply_text(t2)
        got = eng.get_state()
        zz = eng.expect_zz_ring()
    finally:
        eng.close()
    want = oracle.run(s0, n, t1 + t2)
    assert np.max(np.abs(got - want)) <= 1e-12
    assert np.max(np.abs(read_state_files(paths) - want)) <= 1e-12
    assert abs(zz - oracle.expect_zz_ring(want, n)) < 1e-10


def run_blocks_text(eng, text, n):
    gates, blocks, _, nb = sq.parse_circuit(text, n, is_subcircuit=True)
    eng.apply_blocks(gates, blocks, nb)


@pytest.mark.parametrize("tier", ["auto", "pageable", "files"])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,w", [(16, 14), (20, 16)])
def test_out_of_core_subcircuit_plan_matches_oracle(oracle, tmp_path, dtype, n, w, tier):
    # DIO_Runner::run(subcircuits) (DIO_Runner.cpp:399), the DIO runner's
    # default mode (is_subcircuit = 1, simulator.cpp:149): a chunk-12 finder
    # plan -- gate blocks with relabelled targets between VSWAP_2_2 /
    # VSWAP_1_1 passes -- on the out-of-core engine, against the oracle's
    # subcircuit sweep
    text = sq.plan(random_circuit(n, 200, seed=n) + sq.gen_circuit("random", n, 2, seed=n), n, 0, 12)
    assert "VSWAP_2_2" in text and "VSWAP_1_1" in text
    s0 = random_state(n, n + 1)
    kw = dict(resident_qbit=w)
    if tier == "files":
        kw["state_files"] = [str(tmp_path / f"s{i}") for i in range(4)]
    else:
        kw["host_tier"] = tier
    eng = sq.Engine(n, dtype=dtype, **kw)
    try:
        eng.set_state(s0.astype(np.complex128 if dtype == "c128" else np.complex64))
        run_blocks_text(eng, text, n)
        got = eng.get_state()
        passes, _, _ = eng.stream_info()
    finally:
        eng.close()
    assert passes >= 1
    want = oracle.run(s0, n, text, chunk=12, sub=True)
    tol = 1e-12 if dtype == "c128" else 2e-6
    assert np.max(np.abs(got - want)) <= tol


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_out_of_core_vswap_above_the_fixed_run(oracle, dtype):
    # n = resident_qbit + 1: VSWAP_2_2 blocks whose four qubits all lie above
    # a pass's fixed low run need four free slots in one pass
    n, w = 13, 12
    text = high_qubit_subcircuit(n, n - 2, seed=77)
    s0 = random_state(n, 3)
    eng = sq.Engine(n, dtype=dtype, resident_qbit=w)
    try:
        eng.set_state(s0.astype(np.complex128 if dtype == "c128" else np.complex64))
        run_blocks_text(eng, text, n)
        got = eng.get_state()
    finally:
        eng.close()
    want = oracle.run(s0, n, text, chunk=n - 2, sub=True)
    assert np.max(np.abs(got - want)) <= (1e-12 if dtype == "c128" else 2e-6)


@pytest.mark.parametrize("direct", [False, True])
def test_out_of_core_reference_sub_fixture(oracle, tmp_path, odirect_state_dir, direct):
    # the reference's own 21-qubit subcircuit fixture (src/correctness/circuit/sub/21)
    # from |0...0> on state files, O_DIRECT for DirectIO
    import os
    n, w = 21, 18
    text = open(os.path.join(GOLDEN, "ref_circuits", "sub_qft21.txt")).read()
    d = state_dir(tmp_path, direct, odirect_state_dir)
    paths = [str(d / f"sub{i}") for i in range(8)]
    eng = sq.Engine(n, resident_qbit=w, state_files=paths, direct_io=direct)
    try:
        run_blocks_text(eng, text, n)
        got = eng.get_state()
    finally:
        eng.close()
    s = np.zeros(1 << n, dtype=complex)
    s[0] = 1
    want = oracle.run(s, n, text, chunk=12, sub=True)
    assert np.max(np.abs(got - want)) <= 1e-12
    assert np.max(np.abs(read_state_files(paths) - want)) <= 1e-12
