"""Fixtures written by the reference's own compiled gate code
(tests/golden/make_ref_vectors.py, oracle/_ref): the oracle on the CPU and the
engine on the GPU must reproduce them.  The reference is built with -Ofast,
the oracle without contraction, so agreement is to 1e-12, not bit for bit."""
import glob
import os

import numpy as np
import pytest

import subquokka_dev_amd as sq
from helpers import GOLDEN

CASES = sorted(os.path.splitext(os.path.basename(p))[0]
               for p in glob.glob(os.path.join(GOLDEN, "ref_vectors", "*.npz")))


def load(name):
    z = np.load(os.path.join(GOLDEN, "ref_vectors", name + ".npz"), allow_pickle=False)
    text = open(os.path.join(GOLDEN, "ref_vectors", name + ".txt")).read()
    return int(z["n"]), int(z["chunk"]), bool(z["sub"]), z["state_in"], z["state_out"], text


def test_fixture_set_is_present():
    assert len(CASES) >= 6


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_reference_vectors(oracle, name):
    n, chunk, sub, s_in, s_out, text = load(name)
    got = oracle.run(s_in, n, text, chunk=chunk, sub=sub)
    assert np.max(np.abs(got - s_out)) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("sweep_mode", [0, 1])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("name", CASES)
def test_engine_reproduces_reference_vectors(name, dtype, sweep_mode):
    # sweep_mode 1: one launch per gate (plain lists) and LDS tiles for every gate block
    n, chunk, sub, s_in, s_out, text = load(name)
    eng = sq.Engine(n, dtype=dtype, chunk=chunk, sweep_mode=sweep_mode)
    try:
        eng.set_state(s_in.astype(np.complex128 if dtype == "c128" else np.complex64))
        gates, blocks, ng, nb = sq.parse_circuit(text, n, is_subcircuit=sub)
        if sub:
            eng.apply_blocks(gates, blocks, nb)
        else:
            eng.apply_gates(gates, ng)
        got = eng.get_state()
    finally:
        eng.close()
    assert np.max(np.abs(got - s_out)) <= (1e-12 if dtype == "c128" else 2e-6)
