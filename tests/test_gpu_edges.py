"""Edge cases of the drop-in entry points on the GPU: empty and ragged inputs,
subcircuit files made only of VSWAPs, and a VSWAP undone by a later one, in
both sweep modes (0: VSWAPs relabelled on the host; 1: VSWAP permutation sweeps
and LDS tile passes).  Permutations and empty inputs must leave amplitudes
bit-identical; gate blocks are compared with the oracle at the parity bar."""
import numpy as np
import pytest

import subquokka_dev_amd as sq
from helpers import infidelity, random_state

pytestmark = pytest.mark.gpu

N, CHUNK = 14, 8
NP = {"c128": np.complex128, "c64": np.complex64}


def check(got, want, dtype, what):
    # the GPU parity bar: c128 within 1e-12 max-abs; c64 within 2e-6 and 1e-6 infidelity
    err = float(np.max(np.abs(got - want)))
    if dtype == "c128":
        assert err <= 1e-12, f"{what}: max abs err {err:.3e}"
    else:
        assert err <= 2e-6, f"{what}: max abs err {err:.3e}"
        assert infidelity(got, want) <= 1e-6, what


def engine(dtype, sweep_mode, s0):
    eng = sq.Engine(N, dtype=dtype, chunk=CHUNK, sweep_mode=sweep_mode)
    eng.set_state(s0.astype(NP[dtype]))
    return eng


def apply(eng, text, sub):
    gates, blocks, ng, nb = sq.parse_circuit(text, N, is_subcircuit=sub)
    if sub:
        eng.apply_blocks(gates, blocks, nb)
    else:
        eng.apply_gates(gates, ng)


@pytest.mark.parametrize("sweep_mode", [0, 1])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_empty_inputs_leave_the_state_bit_identical(dtype, sweep_mode):
    s0 = random_state(N, 3)
    eng = engine(dtype, sweep_mode, s0)
    try:
        apply(eng, "", False)               # no gates
        apply(eng, "0\n", True)             # one empty block
        apply(eng, "0\n0\n0\n", True)       # only empty blocks
        assert len(eng.sample(0, 5)) == 0
        assert np.array_equal(eng.get_state(NP[dtype]), s0.astype(NP[dtype]))
    finally:
        eng.close()


@pytest.mark.parametrize("sweep_mode", [0, 1])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_vswap_then_the_same_vswap_is_the_identity(dtype, sweep_mode):
    s0 = random_state(N, 5)
    eng = engine(dtype, sweep_mode, s0)
    try:
        apply(eng, "1\nVSWAP_2_2 0 3 9 12\n0\n1\nVSWAP_1_1 5 13\n1\nVSWAP_1_1 5 13\n1\nVSWAP_2_2 0 3 9 12\n", True)
        assert np.array_equal(eng.get_state(NP[dtype]), s0.astype(NP[dtype]))
    finally:
        eng.close()


@pytest.mark.parametrize("sweep_mode", [0, 1])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_only_vswaps_match_the_oracle_bit_for_bit(oracle, dtype, sweep_mode):
    # a pure permutation: the oracle's amplitudes, moved, in the state's own precision
    s0 = random_state(N, 7).astype(NP[dtype])
    text = "1\nVSWAP_2_2 0 3 9 12\n1\nVSWAP_1_1 5 13\n1\nVSWAP_2_2 1 2 8 10\n1\nVSWAP_1_1 0 11\n"
    want = oracle.run(s0.astype(np.complex128), N, text, chunk=CHUNK, sub=True).astype(NP[dtype])
    eng = engine(dtype, sweep_mode, s0)
    try:
        apply(eng, text, True)
        assert np.array_equal(eng.get_state(NP[dtype]), want)
    finally:
        eng.close()


@pytest.mark.parametrize("sweep_mode", [0, 1])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_ragged_blocks_match_the_oracle(oracle, dtype, sweep_mode):
    s0 = random_state(N, 9)
    u = "0.6 0.0 0.0 0.8 0.8 0.0 0.0 -0.6"
    text = ("2\nH 0\nX 1\n0\n1\nVSWAP_1_1 2 9\n"
            f"4\nRY 1 0.3\nCP 2 7 0.7\nU1 4 {u}\nRZZ 0 6 1.1\n0\n1\nVSWAP_2_2 1 4 8 13\n"
            f"3\nH 4\nU2 1 3 {' '.join(['0.5 0.0'] * 16)}\nSWAP 0 6\n0\n")
    want = oracle.run(s0, N, text, chunk=CHUNK, sub=True)
    eng = engine(dtype, sweep_mode, s0)
    try:
        apply(eng, text, True)
        check(eng.get_state(np.complex128), want, dtype, "ragged blocks")
    finally:
        eng.close()
