"""Fused passes (sq_apply_gates_fused): a plain gate list on a complex64 state
cut into passes over LDS-resident tiles, each pass's gates grouped into
four-qubit blocks composed in FP64 and applied on MFMA with split bf16
products (FP32 accumulation), checked against the FP64 oracle applying the
same gates one by one.

Bar: infidelity <= 1e-6 (the fused-form target) and max-abs <= 1e-4 of the
largest amplitude (the split products carry about 1e-5 relative error per
block; a dropped term or a transposed operand gives O(1) errors).
"""
import math
import random
import sys

import numpy as np
import pytest

import subquokka_dev_amd as sq
from helpers import gate_line, infidelity, random_circuit, random_state

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_fp(n, text, state, tile=0):
    gates, _, ng, _ = sq.parse_circuit(text, n)
    eng = sq.Engine(n, dtype="c64")
    try:
        eng.set_state(state.astype(np.complex64))
        eng.apply_gates_fused(gates, ng, tile)
        return eng.get_state(np.complex128)
    finally:
        eng.close()


def check(got, want, what):
    err = float(np.max(np.abs(got - want)))
    scale = float(np.max(np.abs(want)))
    inf = infidelity(got, want)
    say(f"{what}: infidelity {inf:.3e}, max-abs {err:.3e} (max |amp| {scale:.3e})")
    assert inf <= 1e-6, what
    assert err <= 1e-4 * scale, what


@pytest.mark.parametrize("n,tile", [(9, 0), (10, 9), (12, 0), (13, 12), (14, 13), (16, 11), (17, 12)])
def test_fused_passes_match_oracle(oracle, n, tile):
    # every gate kind on random qubits from a seeded Gaussian state
    say(f"fused passes n={n} tile={tile}")
    text = random_circuit(n, 400, seed=31 * n + tile)
    state = random_state(n, n + 100)
    want = oracle.run(state, n, text)
    check(run_fp(n, text, state, tile), want, f"n={n} tile={tile}")


def test_fused_passes_random_layers_18q(oracle):
    say("fused passes: random layers 18q")
    n = 18
    text = sq.gen_circuit("random", n, 20, seed=2)
    state = np.zeros(1 << n, dtype=np.complex128)
    state[0] = 1
    want = oracle.run(state, n, text)
    check(run_fp(n, text, state), want, "random layers 18q")


def test_fused_passes_diagonals_outside_the_tile(oracle):
    # dense gates on qubits 0-6 only, diagonal gates on pairs that reach the
    # high qubits: with a 9-qubit tile most passes hold diagonals whose
    # qubits lie partly (variants) or wholly (tile scalars) outside the tile
    say("fused passes: diagonals outside the tile")
    n = 16
    rng = random.Random(7)
    nrng = np.random.default_rng(7)
    lines = []
    for _ in range(500):
        k = rng.random()
        if k < 0.4:
            lines.append(gate_line(rng.choice(["H", "RX", "RY", "U1"]), [rng.randrange(7)], rng, nrng))
        else:
            kind = rng.choice(["CP", "RZZ", "P", "RZ", "Z"])
            if kind in ("CP", "RZZ"):
                a, b = sorted(rng.sample(range(n), 2))
                lines.append(f"{kind} {a} {b} {rng.uniform(0, 2 * math.pi)!r}")
            elif kind == "Z":
                lines.append(f"Z {rng.randrange(n)}")
            else:
                lines.append(f"{kind} {rng.randrange(n)} {rng.uniform(0, 2 * math.pi)!r}")
    text = "\n".join(lines) + "\n"
    state = random_state(n, 5)
    want = oracle.run(state, n, text)
    for tile in (9, 12):
        check(run_fp(n, text, state, tile), want, f"diagonals outside tile={tile}")


def test_fused_passes_same_structure_new_angles(oracle):
    # the second call reuses the first call's plan (same qubits and gate
    # classes) with new coefficients
    say("fused passes: same structure, new angles")
    n = 14
    for seed in (3, 4):
        text = sq.gen_circuit("random", n, 12, seed=2)
        if seed == 4:   # same lines, other angles
            rng = random.Random(seed)
            out = []
            for line in text.splitlines():
                t = line.split()
                if t and t[0] == "CP":
                    t[3] = repr(rng.uniform(0, 2 * math.pi))
                out.append(" ".join(t))
            text = "\n".join(out) + "\n"
        state = random_state(n, seed)
        want = oracle.run(state, n, text)
        check(run_fp(n, text, state), want, f"angles {seed}")


def test_fused_passes_equal_plain_c128_20q():
    # against the plain complex128 sweep (register passes)
    say("fused passes vs plain c128, 20q")
    n = 20
    text = sq.gen_circuit("random", n, 20, seed=4)
    gates, _, ng, _ = sq.parse_circuit(text, n)
    a, b = sq.Engine(n, dtype="c64"), sq.Engine(n, dtype="c128")
    try:
        a.init_zero()
        a.apply_gates_fused(gates, ng)
        b.init_zero()
        b.apply_gates(gates, ng)
        got, want = a.get_state(np.complex128), b.get_state(np.complex128)
    finally:
        a.close()
        b.close()
    check(got, want, "20q vs plain c128")


def test_fused_passes_rejects_and_empty():
    say("fused passes: rejected calls, empty list")
    text = random_circuit(9, 10, seed=1)
    gates, _, ng, _ = sq.parse_circuit(text, 9)
    e128 = sq.Engine(9, dtype="c128")
    try:
        with pytest.raises(sq.SqError):
            e128.apply_gates_fused(gates, ng)
    finally:
        e128.close()
    e = sq.Engine(9, dtype="c64")
    try:
        with pytest.raises(sq.SqError):
            e.apply_gates_fused(gates, ng, 14)
        s = random_state(9, 2).astype(np.complex64)
        e.set_state(s)
        e.apply_gates_fused(gates, 0)
        assert np.array_equal(e.get_state(np.complex64), s)
    finally:
        e.close()
    t8 = random_circuit(8, 10, seed=1)
    g8, _, n8, _ = sq.parse_circuit(t8, 8)
    e8 = sq.Engine(8, dtype="c64")
    try:
        with pytest.raises(sq.SqError):
            e8.apply_gates_fused(g8, n8)
    finally:
        e8.close()


def test_fused_passes_30q_vs_plain_c128():
    # full size: the 900-gate random-layer circuit, fused complex64 against the
    # plain complex128 register passes
    say("fused passes 30q vs plain c128")
    n = 30
    text = sq.gen_circuit("random", n, 20, seed=2)
    gates, _, ng, _ = sq.parse_circuit(text, n)
    ref = sq.Engine(n, dtype="c128")
    try:
        ref.init_zero()
        ref.apply_gates(gates, ng)
        ref.sync()
        f = sq.Engine(n, dtype="c64")
        try:
            f.init_zero()
            f.apply_gates_fused(gates, ng)
            f.sync()
            ov = ref.inner(f)
            inf = 1.0 - abs(ov) ** 2 / (ref.norm2() * f.norm2())
            err = ref.max_abs_diff(f)
            say(f"30q: infidelity {inf:.3e}, max-abs {err:.3e}")
            assert inf <= 1e-6
            # the largest amplitude of this state is about 1.4e-4 (2^30
            # Porter-Thomas amplitudes): 2e-7 is 1.4e-3 of it
            assert err <= 2e-7
        finally:
            f.close()
    finally:
        ref.close()
