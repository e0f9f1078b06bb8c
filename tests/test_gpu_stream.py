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
