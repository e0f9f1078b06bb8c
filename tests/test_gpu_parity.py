"""GPU parity of the plain and blocked sweeps against the FP64 oracle.

Bars (stated here, asserted below):
  complex128: max |psi_gpu - psi_oracle| <= 1e-12 (normalised states)
  complex64 : max |psi_gpu - psi_oracle| <= 2e-6 and infidelity <= 1e-6
"""
import math
import os
import sys

import numpy as np
import pytest

import subquokka_dev_amd as sq
from helpers import GOLDEN, gate_line, infidelity, random_circuit, random_state

pytestmark = pytest.mark.gpu

TOL128 = 1e-12
TOL64 = 2e-6
INFID64 = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_gpu(n, text, state, dtype="c128", sub=False, chunk=0, sweep_mode=0):
    # blocked sweeps: sweep_mode 0 runs blocks of one-qubit dense and diagonal
    # gates as register passes, sweep_mode 1 runs every gate block as an LDS tile
    eng = sq.Engine(n, dtype=dtype, chunk=chunk, sweep_mode=sweep_mode)
    try:
        eng.set_state(state.astype(np.complex128 if dtype == "c128" else np.complex64))
        gates, blocks, ng, nb = sq.parse_circuit(text, n, is_subcircuit=sub)
        if sub:
            eng.apply_blocks(gates, blocks, nb)
        else:
            eng.apply_gates(gates, ng)
        return eng.get_state(np.complex128)
    finally:
        eng.close()


def check(got, want, dtype, what=""):
    err = float(np.max(np.abs(got - want)))
    if dtype == "c128":
        assert err <= TOL128, f"{what}: max abs err {err:.3e}"
    else:
        assert err <= TOL64, f"{what}: max abs err {err:.3e}"
        assert infidelity(got, want) <= INFID64, what


ONE_Q = ["H", "X", "Y", "Z", "P", "RX", "RY", "RZ", "U1"]
TWO_Q = ["U2", "SWAP", "CP", "RZZ"]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [4, 9, 13])
def test_one_qubit_every_target(oracle, dtype, n):
    # n=4: scalar fallback; 9 and 13: vector / lane-shuffle / register paths.
    import random
    rng = random.Random(n)
    nrng = np.random.default_rng(n)
    lines = [gate_line(k, [t], rng, nrng) for t in range(n) for k in ONE_Q]
    s0 = random_state(n, 11 * n)
    say(f"1q n={n} {dtype} gates={len(lines)}")
    for line in lines:
        want = oracle.run(s0, n, line + "\n")
        got = run_gpu(n, line + "\n", s0, dtype)
        check(got, want, dtype, line)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [4, 9, 12])
def test_two_qubit_every_pair(oracle, dtype, n):
    import random
    rng = random.Random(100 + n)
    nrng = np.random.default_rng(100 + n)
    s0 = random_state(n, 7 * n)
    say(f"2q n={n} {dtype}")
    text_lines = []
    for a in range(n):
        for b in range(a + 1, n):
            for k in TWO_Q:
                text_lines.append(gate_line(k, [a, b], rng, nrng))
    # one launch per gate through the ABI, checked after every 16 gates
    for i in range(0, len(text_lines), 16):
        chunk = "\n".join(text_lines[i:i + 16]) + "\n"
        want = oracle.run(s0, n, chunk)
        got = run_gpu(n, chunk, s0, dtype)
        check(got, want, dtype, chunk.splitlines()[0])


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_unsorted_symmetric_targets(oracle, dtype):
    n = 10
    s0 = random_state(n, 3)
    text = "CP 7 2 0.3\nRZZ 9 0 1.1\nSWAP 8 1\n"
    want = oracle.run(s0, n, text)     # the oracle sorts symmetric targets
    check(run_gpu(n, text, s0, dtype), want, dtype, text)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("q", [(0, 1, 6, 9), (2, 5, 7, 11), (0, 3, 10, 11), (6, 8, 9, 12)])
def test_vswap_2_2(oracle, dtype, q):
    n = 13
    s0 = random_state(n, sum(q))
    text = "1\nVSWAP_2_2 %d %d %d %d\n" % q
    chunk = q[1] + 1
    want = oracle.run(s0, n, text, chunk=chunk, sub=True)
    check(run_gpu(n, text, s0, dtype, sub=True, chunk=chunk), want, dtype, text)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [7, 14, 17])
def test_random_circuit(oracle, dtype, n):
    text = random_circuit(n, 300, seed=n)
    s0 = random_state(n, n + 1)
    say(f"random n={n} {dtype}")
    check(run_gpu(n, text, s0, dtype), oracle.run(s0, n, text), dtype, f"random n={n}")


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("kh", [1, 2, 3, 4, 5])
def test_register_passes_match_one_launch_per_gate(oracle, dtype, kh):
    # the plain sweep as register passes (kh register qubits) against the same
    # gates one launch each, and both against the oracle
    n = 16
    text = sq.gen_circuit("random", n, 6, seed=kh) + random_circuit(n, 200, seed=10 + kh)
    s0 = random_state(n, kh)
    say(f"register passes n={n} kh={kh} {dtype}")
    host = s0.astype(np.complex128 if dtype == "c128" else np.complex64)
    a = sq.Engine(n, dtype=dtype, pass_qbits=kh)
    b = sq.Engine(n, dtype=dtype, sweep_mode=1)
    try:
        a.set_state(host)
        b.set_state(host)
        a.apply_text(text)
        b.apply_text(text)
        assert a.max_abs_diff(b) <= (1e-14 if dtype == "c128" else 1e-6)
        check(a.get_state(np.complex128), oracle.run(s0, n, text), dtype, f"passes kh={kh}")
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("kh", [2, 5])
def test_register_passes_with_lane_moves(oracle, dtype, kh):
    # dense gates on more qubits than the register slots hold, with the top
    # lane qubits idle in between: the planner hands those lane positions to
    # other qubits (the pass swaps lane and address bits), and the gates on
    # the displaced lane qubits run as register or tile-bit work
    n = 18
    lines = []
    for layer in range(12):
        qs = [3, 4, 5] if layer % 3 == 0 else list(range(6 + layer % 4, n, 2))
        for q in qs:
            lines.append(f"U1 {q} 0.6 0.0 0.0 0.8 0.8 0.0 0.0 -0.6")
        lines.append(f"CP {layer % n} {(layer * 7 + 5) % n} 0.{layer + 1}")
        lines.append(f"RZ {(layer * 5) % n} 0.3")
    text = "\n".join(lines) + "\n" + random_circuit(n, 120, seed=40 + kh)
    s0 = random_state(n, 40 + kh)
    say(f"register passes lane moves n={n} kh={kh} {dtype}")
    host = s0.astype(np.complex128 if dtype == "c128" else np.complex64)
    a = sq.Engine(n, dtype=dtype, pass_qbits=kh)
    try:
        a.set_state(host)
        a.apply_text(text)
        check(a.get_state(np.complex128), oracle.run(s0, n, text), dtype, f"lane moves kh={kh}")
    finally:
        a.close()


@pytest.mark.parametrize("kh", [3, 5])
def test_register_passes_fold_controlled_phases(oracle, kh):
    # complex128 passes fold a controlled phase into the next dense gate on
    # either of its qubits: targets on register slots and on lane bits 0-5,
    # controls in register slots, lane bits and tile bits
    import random as _random
    from helpers import gate_line
    n = 16
    rng = _random.Random(100 + kh)
    nrng = np.random.default_rng(kh)
    lines = []
    for _ in range(160):
        a, b = rng.sample(range(n), 2)
        lines.append(f"CP {min(a, b)} {max(a, b)} {rng.uniform(0, 6.28)!r}")
        lines.append(gate_line(rng.choice(["U3", "U1", "H", "RX"]), [a], rng, nrng))
    text = "\n".join(lines) + "\n"
    s0 = random_state(n, 100 + kh)
    say(f"register passes folded phases n={n} kh={kh}")
    a_ = sq.Engine(n, dtype="c128", pass_qbits=kh)
    b_ = sq.Engine(n, dtype="c128", sweep_mode=1)
    try:
        a_.set_state(s0)
        b_.set_state(s0)
        a_.apply_text(text)
        b_.apply_text(text)
        assert a_.max_abs_diff(b_) <= 1e-14
        check(a_.get_state(np.complex128), oracle.run(s0, n, text), "c128", f"folded phases kh={kh}")
    finally:
        a_.close()
        b_.close()


def cnot_line(c, t):
    # CNOT c -> t in the reference's U2 layout (basis |first second>, first the
    # most significant; targets ascending)
    ct = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0]   # control is the first (smaller) qubit
    tc = [1, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0]   # control is the second
    m = ct if c < t else tc
    return f"U2 {min(c, t)} {max(c, t)} " + " ".join(f"{v} 0" for v in m)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_register_passes_cnot_and_diagonal_u2(oracle, dtype):
    # CNOTs written as U2 run inside register passes (H CZ H), diagonal U2
    # (CZ and controlled phases as 4x4 matrices) as diagonal gates: equal to
    # one launch per gate and to the oracle, and no U2 left as a launch of
    # its own
    import random as _random
    from helpers import gate_line
    n = 15
    rng = _random.Random(21)
    nrng = np.random.default_rng(21)
    lines = []
    for _ in range(300):
        r = rng.random()
        a, b = rng.sample(range(n), 2)
        if r < 0.35:
            lines.append(cnot_line(a, b))
        elif r < 0.45:
            ph = rng.uniform(0, 6.28)
            d = [1, 1, 1, complex(np.cos(ph), np.sin(ph))]
            m = []
            for i in range(4):
                for j in range(4):
                    v = d[i] if i == j else 0
                    m += [complex(v).real, complex(v).imag]
            lines.append(f"U2 {min(a, b)} {max(a, b)} " + " ".join(repr(float(x)) for x in m))
        else:
            lines.append(gate_line(rng.choice(["U1", "H", "RX", "RZ"]), [a], rng, nrng))
    text = "\n".join(lines) + "\n"
    s0 = random_state(n, 22)
    say(f"register passes: CNOT and diagonal U2 {dtype}")
    a_ = sq.Engine(n, dtype=dtype)
    b_ = sq.Engine(n, dtype=dtype, sweep_mode=1)
    try:
        a_.set_state(s0.astype(np.complex128 if dtype == "c128" else np.complex64))
        b_.set_state(s0.astype(np.complex128 if dtype == "c128" else np.complex64))
        gates, _, ng, _ = sq.parse_circuit(text, n)
        a_.stats_reset()
        a_.stats_enable(True)
        a_.apply_gates(gates, ng)
        st = a_.stats()
        a_.stats_enable(False)
        b_.apply_gates(gates, ng)
        assert not any(k.startswith("dense2") for k in st), st
        assert a_.max_abs_diff(b_) <= (1e-14 if dtype == "c128" else 2e-6)
        check(a_.get_state(np.complex128), oracle.run(s0, n, text), dtype, f"cnot u2 {dtype}")
    finally:
        a_.close()
        b_.close()


def test_register_passes_every_fold_form(oracle):
    # every folded form present in the program (census) and the result equal to
    # one launch per gate and to the oracle
    from helpers import fold_pairs_circuit
    n = 14
    text = fold_pairs_circuit(n, 400, seed=0)
    s0 = random_state(n, 7)
    say("register passes: every fold form")
    a_ = sq.Engine(n, dtype="c128")
    b_ = sq.Engine(n, dtype="c128", sweep_mode=1)
    try:
        a_.set_state(s0)
        b_.set_state(s0)
        a_.apply_text(text)
        b_.apply_text(text)
        census = a_.pass_census()
        fold = census["fold"]
        assert (fold > 0).all(), fold
        assert census["phases"] == 0   # real gates only: nothing is split
        assert a_.max_abs_diff(b_) <= 1e-14
        check(a_.get_state(np.complex128), oracle.run(s0, n, text), "c128", "every fold form")
    finally:
        a_.close()
        b_.close()


def split_edge_circuit(n, seed):
    # Dense gates written out as matrices on every qubit (vector, lane bits 0-3
    # and 4-5, register slots, tile bits): U3-form rotations, a phase-only gate
    # (its rotation is the identity), an anti-diagonal one (cos 0), Y, a matrix
    # that is not unitary, half a unitary (its rotation part has c^2 + s^2 =
    # 1/4), with CP(pi), P and RZ between them.
    import math
    import random as _random
    from helpers import gate_line
    rng = _random.Random(seed)
    nrng = np.random.default_rng(seed)

    def mat(q, u):
        return f"U1 {q} " + " ".join(f"{complex(z).real!r} {complex(z).imag!r}" for z in u)

    lines = []
    for rep in range(3):
        for q in range(n):
            a, b = 0.3 + 0.2 * q + rep, -1.1 + 0.1 * q
            lines.append(gate_line("U3", [q], rng, nrng))
            lines.append(mat(q, [1, 0, 0, complex(math.cos(a), math.sin(a))]))
            lines.append(mat(q, [0, complex(math.cos(b), math.sin(b)), complex(math.cos(a), math.sin(a)), 0]))
            lines.append(f"Y {q}")
            t = 0.4 + 0.05 * q
            lines.append(mat(q, [0.5 * math.cos(t), -0.5 * complex(math.cos(b), math.sin(b)) * math.sin(t),
                                 0.5 * complex(math.cos(a), math.sin(a)) * math.sin(t),
                                 0.5 * complex(math.cos(a + b), math.sin(a + b)) * math.cos(t)]))
            lines.append(mat(q, [complex(1.9, 0.1), complex(0.2, 0.3), complex(0.4, 0.5), complex(1.7, 0.7)]))
            x, y = rng.sample(range(n), 2)
            lines.append(f"CP {min(x, y)} {max(x, y)} 3.141592653589793")
            lines.append(f"P {rng.randrange(n)} {rng.uniform(0, 6.28)!r}")
            lines.append(f"RZ {rng.randrange(n)} {rng.uniform(0, 6.28)!r}")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_register_passes_split_edge_cases(oracle, dtype):
    # every form the split takes (shears, both scaled lane forms, sign flips,
    # identity rotations skipped, kept-whole matrices) against the oracle and
    # one launch per gate; the non-unitary gates scale the state, so the
    # tolerance is relative to the largest amplitude
    n = 14
    text = split_edge_circuit(n, 5)
    s0 = random_state(n, 21)
    say(f"register passes: split edge cases {dtype}")
    want = oracle.run(s0, n, text)
    scale = float(np.max(np.abs(want)))
    host = s0.astype(np.complex128 if dtype == "c128" else np.complex64)
    a = sq.Engine(n, dtype=dtype)
    b = sq.Engine(n, dtype=dtype, sweep_mode=1)
    try:
        a.set_state(host)
        b.set_state(host)
        a.apply_text(text)
        b.apply_text(text)
        got = a.get_state(np.complex128)
        assert a.pass_census()["phases"] > 0
        tol = (1e-12 if dtype == "c128" else 2e-6) * max(1.0, scale)
        assert float(np.max(np.abs(got - want))) <= tol, (float(np.max(np.abs(got - want))), scale)
        assert float(np.max(np.abs(got - b.get_state(np.complex128)))) <= tol
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [10, 11, 12])
def test_register_passes_smallest_states(oracle, dtype, n):
    # n = lane bits + register qubits exactly (one wave tile) and just above
    text = random_circuit(n, 250, seed=3 * n) + sq.gen_circuit("random", n, 3, seed=n)
    s0 = random_state(n, n)
    say(f"register passes small n={n} {dtype}")
    check(run_gpu(n, text, s0, dtype), oracle.run(s0, n, text), dtype, f"small passes n={n}")


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_config1_qft20_seeded_state(oracle, dtype):
    # 20-qubit QFT in cir_qft.py order on the seeded Gaussian state, one MEM chunk.
    n = 20
    text = sq.gen_circuit("qft", n)
    s0 = sq.gen_state(n, 1)
    say(f"qft20 {dtype}")
    want = oracle.run(s0, n, text, chunk=20)
    check(run_gpu(n, text, s0, dtype), want, dtype, "qft20")


@pytest.mark.parametrize("sweep_mode", [0, 1])
@pytest.mark.parametrize("circ", ["qft", "qaoa", "h", "rx", "rzz"])
def test_reference_sub_fixtures_blocked(oracle, circ, sweep_mode):
    n = 21
    text = open(os.path.join(GOLDEN, "ref_circuits", f"sub_{circ}21.txt")).read()
    s0 = random_state(n, 21)
    say(f"sub fixture {circ}21")
    want = oracle.run(s0, n, text, chunk=12, sub=True)
    check(run_gpu(n, text, s0, "c128", sub=True, chunk=12, sweep_mode=sweep_mode), want, "c128", circ)


def test_reductions(oracle):
    n = 16
    s0 = random_state(n, 2)
    eng = sq.Engine(n)
    try:
        eng.set_state(s0)
        assert abs(eng.norm2() - 1.0) < 1e-13
        assert abs(eng.expect_zz_ring() - oracle.expect_zz_ring(s0, n)) < 1e-12
        other = sq.Engine(n, dtype="c64")
        other.set_state(s0.astype(np.complex64))
        ip = eng.inner(other)
        want = np.vdot(s0, s0.astype(np.complex64).astype(np.complex128))
        assert abs(ip - want) < 1e-12
        assert abs(eng.max_abs_diff(other) - np.max(np.abs(s0 - s0.astype(np.complex64)))) < 1e-15
        other.close()
    finally:
        eng.close()


def test_error_paths_leave_state_untouched():
    n = 8
    s0 = random_state(n, 4)
    eng = sq.Engine(n)
    try:
        eng.set_state(s0)
        gates, _, ng, _ = sq.parse_circuit("H 0\nH 1\n", n)
        gates[1].q[0] = 9          # out of range after parsing
        with pytest.raises(sq.SqError) as ei:
            eng.apply_gates(gates, ng)
        assert ei.value.code == -3
        assert np.array_equal(eng.get_state(), s0)
        with pytest.raises(sq.SqError):
            eng.set_state(np.zeros(10, dtype=np.complex128))
    finally:
        eng.close()


def test_large_state_round_trip_property():
    # 30 qubits complex128 (16 GiB): a 2-layer random circuit followed by its
    # inverse returns |0...0>; the norm stays 1.  Size-independent checks.
    n = 30
    text = sq.gen_circuit("random", n, 2, seed=9)
    inv = []
    for line in reversed(text.strip().splitlines()):
        tok = line.split()
        if tok[0] == "U1":
            v = [float(x) for x in tok[2:10]]
            u = np.array([complex(v[2 * i], v[2 * i + 1]) for i in range(4)]).reshape(2, 2).conj().T
            inv.append(f"U1 {tok[1]} " + " ".join(f"{float(z.real)!r} {float(z.imag)!r}" for z in u.flatten()))
        else:
            inv.append(f"CP {tok[1]} {tok[2]} {-float(tok[3])!r}")
    say("30q round trip")
    eng = sq.Engine(n)
    try:
        eng.apply_text(text)
        assert abs(eng.norm2() - 1) < 1e-12
        eng.apply_text("\n".join(inv) + "\n")
        ref = sq.Engine(n)
        try:
            assert eng.max_abs_diff(ref) < 1e-12
        finally:
            ref.close()
    finally:
        eng.close()


def test_one_launch_per_gate_32q_lane_targets():
    # 32 qubits complex128 (64 GiB per state): a one-qubit gate on a lane bit
    # asks for 2^24 workgroups of 256 threads, 2^32 threads, past what one
    # launch holds; the grid is capped and the kernels' grid-stride loops
    # cover the rest.  One launch per gate against the register passes, both
    # on the device; the norm is that of six H gates with the reference's
    # literal 0.70710678118 (2 h^2 per gate, not exactly 1).
    from helpers import H_LITERAL
    n = 32
    text = "".join(f"H {q}\n" for q in range(6)) + "RX 3 0.7\nRY 31 1.3\nCP 2 30 3.141592653589793\n" + \
        "U1 0 0.6 0.0 0.0 0.8 0.8 0.0 0.0 -0.6\nP 5 0.4\n"
    say("32q one launch per gate")
    a = sq.Engine(n, dtype="c128", sweep_mode=1)
    try:
        a.apply_text(text)
        assert abs(a.norm2() - (2 * H_LITERAL ** 2) ** 6) < 1e-13
        b = sq.Engine(n, dtype="c128")
        try:
            b.apply_text(text)
            assert a.max_abs_diff(b) < 1e-14
        finally:
            b.close()
    finally:
        a.close()


def test_full_circuit_30q_c64_against_c128():
    # the whole 900-gate random-layer circuit at 30 qubits as complex64
    # register passes against complex128 (size-independent: fidelity and the
    # norms), and the complex128 result's norm and <sum ZZ> unchanged when the
    # same circuit runs one launch per gate
    n = 30
    text = sq.gen_circuit("random", n, 20, seed=2)
    gates, _, ng, _ = sq.parse_circuit(text, n)
    say("30q full circuit c64 vs c128")
    a = sq.Engine(n, dtype="c128")
    try:
        a.apply_gates(gates, ng)
        assert abs(a.norm2() - 1) < 1e-12
        b = sq.Engine(n, dtype="c64")
        try:
            b.apply_gates(gates, ng)
            inf = 1.0 - abs(a.inner(b)) ** 2 / (a.norm2() * b.norm2())
            err = a.max_abs_diff(b)
            say(f"30q c64 vs c128: infidelity {inf:.3e}, max-abs {err:.3e}")
            assert inf <= 1e-6
            assert err <= 1e-7
        finally:
            b.close()
    finally:
        a.close()


@pytest.mark.parametrize("sweep_mode", [0, 1])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,k", [(14, 5), (18, 4), (20, 6)])
def test_planned_blocks_in_lds_tiles(oracle, dtype, n, k, sweep_mode):
    # planner blocks (<= k qubits each) run as LDS tile passes; same state as plain
    text = random_circuit(n, 400, seed=n + k)
    plan = sq.plan(text, n, 1, k)
    s0 = random_state(n, n)
    say(f"tiles n={n} k={k} {dtype}")
    check(run_gpu(n, plan, s0, dtype, sub=True, sweep_mode=sweep_mode), oracle.run(s0, n, text), dtype, f"tiles n={n}")


@pytest.mark.parametrize("sweep_mode", [0, 1])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_wide_block_is_split_into_tiles(oracle, dtype, sweep_mode):
    # one block touching all 16 qubits: wider than any tile, split by the planner
    n = 16
    text = random_circuit(n, 300, seed=5)
    sub = f"{len(text.strip().splitlines())}\n" + text
    s0 = random_state(n, 6)
    check(run_gpu(n, sub, s0, dtype, sub=True, sweep_mode=sweep_mode), oracle.run(s0, n, text), dtype, "wide block")


@pytest.mark.parametrize("sweep_mode", [0, 1])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,chunk", [(16, 10), (20, 12)])
def test_finder_plan_without_restore(oracle, n, chunk, dtype, sweep_mode):
    # a plan that ends with its qubits permuted (no restore swaps): the final
    # layout is the one the VSWAPs leave.  sweep_mode 0 relabels the gates and
    # swaps qubits once at the end; sweep_mode 1 moves the data at every VSWAP.
    # The oracle applies the file VSWAP by VSWAP.  Two-qubit gates (U2, SWAP,
    # CP, RZZ) meet relabellings that reverse their targets.
    text = random_circuit(n, 300, seed=chunk + 7)
    plan = sq.plan(text, n, 0, chunk)
    assert "VSWAP" in plan
    s0 = random_state(n, 4)
    say(f"plan without restore n={n} chunk={chunk} {dtype} sweep_mode={sweep_mode}")
    want = oracle.run(s0, n, plan, chunk=chunk, sub=True)
    check(run_gpu(n, plan, s0, dtype, sub=True, chunk=chunk, sweep_mode=sweep_mode), want, dtype, "plan without restore")


@pytest.mark.parametrize("sweep_mode", [0, 1])
@pytest.mark.parametrize("n,chunk", [(16, 10), (20, 12), (24, 12)])
def test_finder_plan_vswap_blocks(oracle, n, chunk, sweep_mode):
    # planner in the finder's format (blocks + VSWAP passes, order restored)
    text = random_circuit(n, 250, seed=chunk)
    plan = sq.plan(text, n, 3, chunk)
    s0 = random_state(n, 2)
    say(f"finder plan n={n} chunk={chunk}")
    check(run_gpu(n, plan, s0, "c128", sub=True, chunk=chunk, sweep_mode=sweep_mode), oracle.run(s0, n, text), "c128",
          "finder plan")


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_schedule_reused_for_the_same_structure(oracle, dtype):
    # the register-pass schedule is cached by the gates' qubits and classes:
    # the same circuit with other angles reuses it, with its own coefficients
    n = 16
    a = sq.gen_circuit("random", n, 6, seed=11)
    b = sq.gen_circuit("random", n, 6, seed=12)   # another matching and other angles
    a2 = "".join(line.split()[0] + " " + " ".join(line.split()[1:2]) + " " +
                 " ".join(repr(float(x) * 0.5) for x in line.split()[2:]) + "\n"
                 if line.startswith("U1") else line + "\n" for line in a.strip().splitlines())
    s0 = random_state(n, 4)
    tol = 1e-12 if dtype == "c128" else 2e-6
    eng = sq.Engine(n, dtype=dtype)
    try:
        state = s0
        for text in (a, a2, b, a):
            eng.set_state(state.astype(np.complex128 if dtype == "c128" else np.complex64))
            eng.apply_text(text)
            want = oracle.run(state, n, text)
            assert np.max(np.abs(eng.get_state() - want)) <= tol
            state = want
    finally:
        eng.close()
