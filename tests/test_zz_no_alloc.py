"""sq_apply_gates_noalloc: the plain sweep with its register-pass program in a
fixed device array (no allocation) gives sq_apply_gates's result, call after
call, and rejects a program larger than the array before anything runs."""
import numpy as np
import pytest

import subquokka_dev_amd as sq
from helpers import random_circuit, random_state

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 2e-6}


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_noalloc_matches_apply_gates_on_two_calls(oracle, dtype):
    n = 20
    first = sq.gen_circuit("random", n, 8, seed=4) + random_circuit(n, 300, seed=41)
    second = random_circuit(n, 40, seed=42)   # a smaller list: the array keeps the first call's tail
    s0 = random_state(n, 9)
    host = np.complex128 if dtype == "c128" else np.complex64
    a = sq.Engine(n, dtype=dtype)
    b = sq.Engine(n, dtype=dtype)
    try:
        a.set_state(s0.astype(host))
        b.set_state(s0.astype(host))
        for text in (first, second):
            gates, _, ng, _ = sq.parse_circuit(text, n)
            print(f"noalloc {dtype} n={n} gates={ng}", flush=True)
            a.apply_gates_noalloc(gates, ng)
            b.apply_gates(gates, ng)
            assert np.max(np.abs(a.get_state() - b.get_state())) <= TOL[dtype]
        want = oracle.run(s0, n, first + second)
        assert np.max(np.abs(a.get_state() - want)) <= TOL[dtype]
    finally:
        a.close()
        b.close()


def test_noalloc_small_state_and_empty_list(oracle):
    # below the register-pass width every gate is its own launch; an empty
    # list leaves the state bit for bit
    n = 8
    text = random_circuit(n, 50, seed=5)
    s0 = random_state(n, 2)
    eng = sq.Engine(n)
    try:
        eng.set_state(s0)
        gates, _, ng, _ = sq.parse_circuit(text, n)
        eng.apply_gates_noalloc(gates, ng)
        got = eng.get_state()
        eng.apply_gates_noalloc(gates, 0)
        assert np.array_equal(eng.get_state(), got)
    finally:
        eng.close()
    assert np.max(np.abs(got - oracle.run(s0, n, text))) <= 1e-12


def test_noalloc_rejects_a_program_beyond_the_array():
    # 40000 one-qubit gates on register qubits: about 5 MB of complex128 op
    # records, more than the 4 MiB array; nothing runs, the state is untouched
    n = 14
    text = "".join(f"RX {11 + (i % 3)} {0.001 * (i + 1)!r}\n" for i in range(40000))
    s0 = random_state(n, 3)
    eng = sq.Engine(n)
    try:
        eng.set_state(s0)
        gates, _, ng, _ = sq.parse_circuit(text, n)
        with pytest.raises(sq.SqError) as ex:
            eng.apply_gates_noalloc(gates, ng)
        assert ex.value.code == sq.native.SQ_ERR_ARG and "scratch" in str(ex.value)
        assert np.array_equal(eng.get_state(), s0)
    finally:
        eng.close()


def test_noalloc_rejects_the_out_of_core_runner():
    eng = sq.Engine(12, resident_qbit=9)
    try:
        gates, _, ng, _ = sq.parse_circuit("H 0\n", 12)
        with pytest.raises(sq.SqError) as ex:
            eng.apply_gates_noalloc(gates, ng)
        assert ex.value.code == sq.native.SQ_ERR_UNSUPPORTED
    finally:
        eng.close()
