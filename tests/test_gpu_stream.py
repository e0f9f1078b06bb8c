"""Out-of-core runner (DIO_Runner role): state in pinned host memory, streamed
through HBM pass by pass; results must equal the HBM-resident engine."""
import numpy as np
import pytest

import subquokka_dev_amd as sq
from helpers import random_circuit, random_state

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


@pytest.mark.parametrize("direct", [False, True])
def test_file_tier_matches_oracle_and_resumes(oracle, tmp_path, direct):
    # state in 4 files of the reference's layout; a second engine reopens the
    # files and continues the circuit from them
    n, w = 16, 14
    paths = [str(tmp_path / f"path{i}") for i in range(4)]
    t1 = sq.gen_circuit("random", n, 2, seed=3) + random_circuit(n, 60, seed=31)
    t2 = random_circuit(n, 60, seed=32)
    s0 = random_state(n, 6)
    try:
        eng = sq.Engine(n, resident_qbit=w, state_files=paths, direct_io=direct)
    except sq.SqError as e:
        if direct and e.code == -9:
            pytest.skip("O_DIRECT is not supported on this filesystem")
        raise
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
