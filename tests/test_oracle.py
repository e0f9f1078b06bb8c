"""Pins the FP64 CPU oracle (oracle/sq_oracle.c).

Each check is independent of the oracle's loop machinery:
  * closed forms of the reference's own fixture circuits from |0...0>
    (src/correctness/circuit/{ori,sub}/21/*.txt, copied as data);
  * dense 2^n x 2^n textbook matrices, gate by gate, every gate on every
    target and every loop variant (chunk sizes), like gate_check_MEM.py:26-301;
  * the VSWAP_2_2 identity of gate_check_MEM.py:339-341;
  * the QFT closed form U = R F R (cir_qft.py:27-35 order);
  * the reference's own compiled gate code (oracle/_ref), when built.
"""
import math
import os

import numpy as np
import pytest

from helpers import (GOLDEN, H_LITERAL, RefHarness, qft_closed_form, random_circuit, random_state,
                     ref_available, textbook_apply)

FIX = os.path.join(GOLDEN, "ref_circuits")
C = H_LITERAL


def read_fix(name):
    return open(os.path.join(FIX, name)).read()


def zero_state(n):
    s = np.zeros(1 << n, dtype=complex)
    s[0] = 1
    return s


# ---------------------------------------------------------------------------
# closed forms on the reference's fixture circuits (N = 21, chunk 12)
# ---------------------------------------------------------------------------

N21 = 21


def closed_form(circ, n):
    d = 1 << n
    if circ in ("h", "qft"):
        return np.full(d, C ** n, dtype=complex)
    if circ == "rx":
        s = np.zeros(d, dtype=complex)
        s[d - 1] = (-1j) ** n
        return s
    if circ == "rzz":
        s = np.zeros(d, dtype=complex)
        s[0] = (-1j) ** (n - 1)
        return s
    if circ == "qaoa":
        p = n * (n - 1) // 2
        w = np.array([bin(i).count("1") for i in range(d)])
        return C ** n * (-1j) ** (5 * (p + n)) * (-1.0) ** (w * (n - w))
    raise KeyError(circ)


FIXTURE_CASES = [(m, c) for m in ("ori", "sub") for c in ("h", "rx", "rzz", "qft", "qaoa")
                 if (m, c) != ("sub", "qaoa")]


@pytest.mark.parametrize("mode,circ", FIXTURE_CASES)
def test_reference_fixture_closed_forms(oracle, circ, mode):
    text = read_fix(f"{mode}_{circ}21.txt")
    out = oracle.run(zero_state(N21), N21, text, chunk=12, sub=(mode == "sub"))
    np.testing.assert_allclose(out, closed_form(circ, N21), rtol=0, atol=1e-12)


def test_reference_sub_qaoa_fixture_is_not_the_plain_circuit(oracle):
    # The finder's sub/21/qaoa21.txt reorders non-commuting RZZ/RX gates: the
    # state it produces (identical under the reference's own gate code, see
    # test_oracle_matches_reference_on_sub_qaoa) differs in sign from the plain
    # circuit's on half of the amplitudes.  Magnitudes still agree.
    out = oracle.run(zero_state(N21), N21, read_fix("sub_qaoa21.txt"), chunk=12, sub=True)
    want = closed_form("qaoa", N21)
    np.testing.assert_allclose(np.abs(out), np.abs(want), rtol=0, atol=1e-12)
    ratio = np.round((out / want).real)
    assert set(np.unique(ratio)) == {-1.0, 1.0}
    assert int((ratio < 0).sum()) == (1 << N21) // 2


# ---------------------------------------------------------------------------
# dense textbook matrices: every gate, every target, every loop variant
# ---------------------------------------------------------------------------

ONE_Q = ["H", "X", "Y", "Z", "P", "RX", "RY", "RZ", "U1"]
TWO_Q = ["U2", "SWAP", "CP", "RZZ"]


@pytest.mark.parametrize("kind", ONE_Q)
@pytest.mark.parametrize("chunk", [1, 2, 3])
def test_one_qubit_gates_every_target(oracle, kind, chunk):
    # gate_check_MEM.py:16-24 uses N=3, chunk=1; chunk 2, 3 exercise the chunk variant too.
    from helpers import gate_line
    import random
    n = 4
    rng = random.Random(7)
    nrng = np.random.default_rng(7)
    for t in range(n):
        line = gate_line(kind, [t], rng, nrng) + "\n"
        s0 = random_state(n, 100 + t)
        got = oracle.run(s0, n, line, chunk=chunk)
        want = textbook_apply(s0, n, line)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13, err_msg=line)


@pytest.mark.parametrize("kind", TWO_Q)
@pytest.mark.parametrize("chunk", [1, 2, 4, 6])
def test_two_qubit_gates_every_pair(oracle, kind, chunk):
    # gate_check_MEM.py:224-301: N=6, chunk=2, all ordered pairs.
    from helpers import gate_line
    import random
    n = 6
    rng = random.Random(11)
    nrng = np.random.default_rng(11)
    for a in range(n):
        for b in range(n):
            if a == b or (kind == "U2" and a > b):
                continue
            line = gate_line(kind, [a, b], rng, nrng) + "\n"
            s0 = random_state(n, 1000 + 8 * a + b)
            got = oracle.run(s0, n, line, chunk=chunk)
            want = textbook_apply(s0, n, line)
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-13, err_msg=line)


def test_u2_layout_matches_qiskit_order(oracle):
    # test_util.py:194-204: the U2 matrix acts with the first-listed qubit as MSB.
    # A CNOT with control 0 and target 3 in that layout flips qubit 3 when qubit 0 is set.
    n = 4
    cnot = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0]
    line = "U2 0 3 " + " ".join(f"{v} 0" for v in cnot) + "\n"
    for i in range(1 << n):
        s = np.zeros(1 << n, dtype=complex)
        s[i] = 1
        out = oracle.run(s, n, line, chunk=2)
        j = i ^ 8 if i & 1 else i
        assert abs(out[j] - 1) < 1e-15


def test_vswap_2_2_identity(oracle):
    # gate_check_MEM.py:307-348: N=7, chunk=3, VSWAP_2_2(q0,q1,q2,q3) == SWAP(q0,q2) SWAP(q1,q3)
    n, chunk = 7, 3
    s0 = random_state(n, 5)
    for q in [(0, 1, 3, 4), (0, 2, 3, 6), (1, 2, 4, 5), (0, 1, 5, 6)]:
        got = oracle.run(s0, n, "1\nVSWAP_2_2 %d %d %d %d\n" % q, chunk=chunk, sub=True)
        want = oracle.run(s0, n, f"SWAP {q[0]} {q[2]}\nSWAP {q[1]} {q[3]}\n", chunk=chunk)
        assert np.array_equal(got, want)
        dense = textbook_apply(s0, n, "VSWAP_2_2 %d %d %d %d\n" % q)
        assert np.array_equal(got, dense)


def test_vswap_outside_contract_rejected(oracle):
    s0 = random_state(6, 1)
    assert oracle.rc(s0, 6, "1\nVSWAP_2_2 0 4 2 5\n", chunk=3, sub=True) == -4


def test_unknown_gate_and_bad_qubit(oracle):
    s0 = random_state(4, 1)
    assert oracle.rc(s0, 4, "CZ 0 1\n") == -1        # finder accepts CZ, simulator does not
    assert oracle.rc(s0, 4, "H 4\n") == -2            # check_targ_in_N
    assert oracle.rc(s0, 4, "U2 2 1 " + "0 0 " * 16 + "\n") == -4   # circuit.cpp:2809


def test_bit_string_is_bijection(oracle):
    # MEM_Runner.cpp:90-99: task -> index with 0 at every (ascending) target.
    n = 8
    for targ in ([0], [3], [7], [0, 1], [2, 5], [1, 4, 6, 7]):
        seen = set()
        for task in range(1 << (n - len(targ))):
            idx = oracle.bit_string(task, targ)
            assert all(((idx >> t) & 1) == 0 for t in targ)
            seen.add(idx)
        assert len(seen) == 1 << (n - len(targ))


@pytest.mark.parametrize("chunk", [1, 3, 5, 8])
def test_random_circuit_vs_dense(oracle, chunk):
    n = 8
    text = random_circuit(n, 60, seed=chunk)
    s0 = random_state(n, 3)
    np.testing.assert_allclose(oracle.run(s0, n, text, chunk=chunk), textbook_apply(s0, n, text),
                               rtol=0, atol=1e-12)


@pytest.mark.parametrize("n", [5, 8, 10])
def test_qft_closed_form(oracle, n):
    import subquokka_dev_amd as sq
    text = sq.gen_circuit("qft", n)
    s0 = random_state(n, 42 + n)
    np.testing.assert_allclose(oracle.run(s0, n, text, chunk=min(n, 3)), qft_closed_form(s0, n),
                               rtol=0, atol=1e-12)


def test_sub_equals_plain_with_relabelling(oracle):
    # Blocked output = plain output with index bits permuted by the composed
    # VSWAP relabelling (the finder never restores the order, finder.cpp:304-330).
    # Built here from a plain circuit: relabel, insert VSWAPs, compare.
    n, chunk = 8, 4
    rng = np.random.default_rng(3)
    s0 = random_state(n, 9)
    plain = random_circuit(n, 40, seed=21, kinds=["H", "RX", "RZ", "U1", "CP", "RZZ", "U2"])
    gates = plain.strip().splitlines()
    # physical position of each logical qubit
    pos = list(range(n))
    out_lines = []
    block = []
    for line in gates:
        tok = line.split()
        nq = 2 if tok[0] in ("U2", "SWAP", "CP", "RZZ") else 1
        qs = [int(x) for x in tok[1:1 + nq]]
        phys = [pos[q] for q in qs]
        if any(p >= chunk for p in phys):
            if block:
                out_lines += [str(len(block))] + block
                block = []
            # bring the offending qubits into the chunk with VSWAP_1_1 (chunk qubit <-> high qubit)
            used = set(phys)
            for q in qs:
                if pos[q] >= chunk:
                    lo = next(p for p in range(chunk) if p not in used)
                    used.add(lo)
                    hi = pos[q]
                    other = pos.index(lo)
                    out_lines += ["1", f"VSWAP_1_1 {lo} {hi}"]
                    pos[q], pos[other] = lo, hi
            phys = [pos[q] for q in qs]
        if tok[0] == "U2" and phys[0] > phys[1]:
            # keep the U2 contract (ascending targets) by reordering the matrix basis
            m = np.array([float(x) for x in tok[3:35]]).reshape(16, 2)
            u = (m[:, 0] + 1j * m[:, 1]).reshape(4, 4)
            perm = [0, 2, 1, 3]
            u = u[np.ix_(perm, perm)]
            tok = ["U2", str(phys[1]), str(phys[0])] + [f"{float(v)!r}" for z in u.flatten() for v in (z.real, z.imag)]
            block.append(" ".join(tok))
            continue
        rest = tok[1 + nq:]
        if nq == 2:
            phys = sorted(phys) if tok[0] != "U2" else phys
        block.append(" ".join([tok[0]] + [str(p) for p in phys] + rest))
    if block:
        out_lines += [str(len(block))] + block
    sub_text = "\n".join(out_lines) + "\n"
    got = oracle.run(s0, n, sub_text, chunk=chunk, sub=True)
    want_logical = oracle.run(s0, n, plain, chunk=chunk)
    # logical index -> physical index
    idx = np.arange(1 << n)
    phys_idx = np.zeros_like(idx)
    for q in range(n):
        phys_idx |= ((idx >> q) & 1) << pos[q]
    # initial state was laid out logically == physically (pos starts as identity)
    np.testing.assert_allclose(got[phys_idx], want_logical, rtol=0, atol=1e-12)


def test_expect_zz_ring_basis_and_dense(oracle):
    n = 6
    for i in [0, 1, 5, 21, 63]:
        s = np.zeros(1 << n, dtype=complex)
        s[i] = 1
        bits = [(i >> q) & 1 for q in range(n)]
        want = sum(1 if bits[q] == bits[(q + 1) % n] else -1 for q in range(n))
        assert oracle.expect_zz_ring(s, n) == want
    s = random_state(n, 4)
    z = np.diag([1, -1])
    h = np.zeros((1 << n, 1 << n))
    for q in range(n):
        from helpers import full_1q
        h = h + (full_1q(n, q, z) @ full_1q(n, (q + 1) % n, z)).real
    assert abs(oracle.expect_zz_ring(s, n) - np.vdot(s, h @ s).real) < 1e-13


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("chunk", [1, 2, 4, 9])
def test_oracle_matches_reference_gate_code(oracle, chunk):
    ref = RefHarness()
    n = 9
    text = random_circuit(n, 150, seed=100 + chunk)
    s0 = random_state(n, 8)
    np.testing.assert_allclose(oracle.run(s0, n, text, chunk=chunk), ref.run(s0, n, text, chunk=chunk),
                               rtol=0, atol=1e-13)


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("circ", ["qft", "qaoa"])
def test_oracle_matches_reference_on_sub_fixture(oracle, circ):
    ref = RefHarness()
    text = read_fix(f"sub_{circ}21.txt")
    s0 = random_state(N21, 77)
    np.testing.assert_allclose(oracle.run(s0, N21, text, chunk=12, sub=True),
                               ref.run(s0, N21, text, chunk=12, sub=True), rtol=0, atol=1e-13)
