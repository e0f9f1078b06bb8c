"""Fused five-qubit blocks: FP64 composition on the GPU and application on
MFMA, checked against the FP64 oracle applying the same gates one by one.

Bars: complex128 (FP64 MFMA) max-abs <= 1e-12; complex64 (FP32 MFMA, FP32
accumulation) max-abs <= 2e-6 and infidelity <= 1e-6.
"""
import sys

import numpy as np
import pytest

import subquokka_dev_amd as sq
from helpers import infidelity, random_circuit, random_state, textbook_apply

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fused_run(n, plain_text, k, state, dtype):
    plan = sq.plan(plain_text, n, 1, k)
    gates, blocks, ng, nb = sq.parse_circuit(plan, n, is_subcircuit=True)
    eng = sq.Engine(n, dtype=dtype)
    try:
        eng.set_state(state.astype(np.complex128 if dtype == "c128" else np.complex64))
        eng.run_fused(gates, blocks, nb)
        return eng.get_state(np.complex128), nb
    finally:
        eng.close()


def check(got, want, dtype, what):
    err = float(np.max(np.abs(got - want)))
    if dtype == "c128":
        assert err <= 1e-12, f"{what}: {err:.3e}"
    else:
        assert err <= 2e-6, f"{what}: {err:.3e}"
        assert infidelity(got, want) <= 1e-6, what


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_compose_block_matches_dense_product(k):
    # compose (FP64, GPU) vs the product of textbook dense matrices, in the
    # caller's qubit order (local bit j <-> q[j])
    n = 8
    rng = np.random.default_rng(k)
    q = [int(x) for x in rng.permutation(n)[:k]]
    lines = []
    import random
    from helpers import gate_line
    pr = random.Random(k)
    for _ in range(12):
        kinds = ["H", "RX", "RZ", "U1", "P", "Y"] + (["U2", "CP", "RZZ", "SWAP"] if k >= 2 else [])
        kind = pr.choice(kinds)
        qs = sorted(pr.sample(q, 2)) if kind in ("U2", "CP", "RZZ", "SWAP") else [pr.choice(q)]
        lines.append(gate_line(kind, qs, pr, rng))
    text = "\n".join(lines) + "\n"
    gates, _, ng, _ = sq.parse_circuit(text, n)
    eng = sq.Engine(n)
    try:
        u = eng.compose(gates, ng, q)
    finally:
        eng.close()
    # reference: apply the gates to every basis state of the k-qubit subspace
    d = 1 << k
    want = np.zeros((d, d), dtype=complex)
    for c in range(d):
        s = np.zeros(1 << n, dtype=complex)
        s[sum(((c >> j) & 1) << q[j] for j in range(k))] = 1
        out = textbook_apply(s, n, text)
        for r in range(d):
            want[r, c] = out[sum(((r >> j) & 1) << q[j] for j in range(k))]
    np.testing.assert_allclose(u, want, rtol=0, atol=1e-13)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [7, 9, 12, 16])
def test_run_fused_matches_oracle(oracle, dtype, n):
    text = random_circuit(n, 300, seed=n + 40)
    s0 = random_state(n, n + 5)
    say(f"fused n={n} {dtype}")
    got, nb = fused_run(n, text, 5, s0, dtype)
    assert nb < 300
    check(got, oracle.run(s0, n, text), dtype, f"fused n={n}")


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_run_fused_random_layers_18q(oracle, dtype):
    n = 18
    text = sq.gen_circuit("random", n, 6, seed=3)
    s0 = sq.gen_state(n, 4)
    got, _ = fused_run(n, text, 5, s0, dtype)
    check(got, oracle.run(s0, n, text), dtype, "random layers 18q")


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_apply_fused_host_matrices(oracle, dtype):
    n = 11
    rng = np.random.default_rng(9)
    from helpers import random_unitary
    blocks = []
    text_equiv = []
    for k, q in [(2, [3, 9]), (3, [10, 0, 5]), (5, [1, 2, 4, 7, 8]), (1, [6])]:
        u = random_unitary(1 << k, rng)
        blocks.append((q, u))
    s0 = random_state(n, 2)
    eng = sq.Engine(n, dtype=dtype)
    try:
        eng.set_state(s0.astype(np.complex128 if dtype == "c128" else np.complex64))
        eng.apply_fused(blocks)
        got = eng.get_state()
    finally:
        eng.close()
    # dense reference: each block as a full operator (local bit j <-> q[j])
    want = s0.copy()
    for q, u in blocks:
        k = len(q)
        rest = [b for b in range(n) if b not in q]
        out = np.zeros_like(want)
        for idx in range(1 << n):
            loc = sum(((idx >> q[j]) & 1) << j for j in range(k))
            base = idx & ~sum(1 << b for b in q)
            for r in range(1 << k):
                out[base | sum(((r >> j) & 1) << q[j] for j in range(k))] += u[r, loc] * want[idx]
        want = out
    check(got, want, dtype, "apply_fused")


def test_fused_c64_vs_plain_c128_30q_infidelity():
    # The headline pair at full size: fused complex64 (FP32 MFMA) against the
    # plain complex128 gate-by-gate sweep of the same 30-qubit circuit (4 layers).
    n = 30
    text = sq.gen_circuit("random", n, 4, seed=2)
    plan = sq.plan(text, n, 1, 5)
    say("30q fused c64 vs plain c128")
    a = sq.Engine(n, dtype="c128")
    b = sq.Engine(n, dtype="c64")
    try:
        a.apply_text(text)
        gates, blocks, ng, nb = sq.parse_circuit(plan, n, is_subcircuit=True)
        b.run_fused(gates, blocks, nb)
        ov = a.inner(b)
        inf = 1 - abs(ov) ** 2 / (a.norm2() * b.norm2())
        assert inf <= 1e-6, inf
        assert a.max_abs_diff(b) <= 1e-6
    finally:
        a.close()
        b.close()
