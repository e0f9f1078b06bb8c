"""The planner (finder role, subcircuitFinder/finder.cpp:69-877), checked on the
CPU with the oracle: a plan must compute the same state as the plain circuit."""
import numpy as np
import pytest

import subquokka_dev_amd as sq
from helpers import random_circuit, random_state


def blocks_of(text):
    lines = [l for l in text.splitlines() if l.strip()]
    out, i = [], 0
    while i < len(lines):
        c = int(lines[i])
        out.append(lines[i + 1:i + 1 + c])
        i += 1 + c
    return out


def qubits_of(line):
    tok = line.split()
    nq = {"U2": 2, "SWAP": 2, "CP": 2, "RZZ": 2, "VSWAP_1_1": 2, "VSWAP_2_2": 4}.get(tok[0], 1)
    return [int(x) for x in tok[1:1 + nq]]


@pytest.mark.parametrize("n,chunk,seed", [(8, 4, 1), (10, 5, 2), (12, 6, 3)])
def test_subcircuit_plan_restored_equals_plain(oracle, n, chunk, seed):
    text = random_circuit(n, 150, seed=seed)
    plan = sq.plan(text, n, 3, chunk)
    # contract of MEM_Runner.cpp:173-182: every non-VSWAP block is chunk-local
    for blk in blocks_of(plan):
        if not blk[0].startswith("VSWAP"):
            assert all(q < chunk for l in blk for q in qubits_of(l))
        else:
            assert len(blk) == 1
    s0 = random_state(n, seed)
    np.testing.assert_allclose(oracle.run(s0, n, plan, chunk=chunk, sub=True), oracle.run(s0, n, text, chunk=chunk),
                               rtol=0, atol=1e-12)


def test_subcircuit_plan_unrestored_is_a_permutation(oracle):
    # Without the final swaps (the reference finder's behaviour) the output is
    # the plain result with index bits permuted.
    n, chunk = 10, 4
    text = random_circuit(n, 120, seed=9)
    plan = sq.plan(text, n, 0, chunk)
    s0 = random_state(n, 1)
    got = oracle.run(s0, n, plan, chunk=chunk, sub=True)
    want = oracle.run(s0, n, text, chunk=chunk)
    # recover the permutation by replaying the VSWAPs
    at = list(range(n))
    for blk in blocks_of(plan):
        if blk[0].startswith("VSWAP"):
            q = qubits_of(blk[0])
            pairs = [(q[0], q[1])] if len(q) == 2 else [(q[0], q[2]), (q[1], q[3])]
            for a, b in pairs:
                at[a], at[b] = at[b], at[a]
    pos = [0] * n
    for p, q in enumerate(at):
        pos[q] = p
    idx = np.arange(1 << n)
    phys = np.zeros_like(idx)
    for q in range(n):
        phys |= ((idx >> q) & 1) << pos[q]
    np.testing.assert_allclose(got[phys], want, rtol=0, atol=1e-12)
    assert at != list(range(n))   # the plan did move qubits


@pytest.mark.parametrize("k", [2, 3, 5])
def test_fused_plan_blocks_are_k_local_and_equivalent(oracle, k):
    n = 9
    text = random_circuit(n, 200, seed=k)
    plan = sq.plan(text, n, 1, k)
    blocks = blocks_of(plan)
    assert sum(len(b) for b in blocks) == 200
    for b in blocks:
        assert len({q for l in b for q in qubits_of(l)}) <= k
    flat = "\n".join(l for b in blocks for l in b) + "\n"
    s0 = random_state(n, 4)
    np.testing.assert_allclose(oracle.run(s0, n, flat), oracle.run(s0, n, text), rtol=0, atol=1e-12)


def test_fused_plan_on_random_layers_reduces_passes():
    n = 30
    text = sq.gen_circuit("random", n, 20, seed=2)
    blocks = blocks_of(sq.plan(text, n, 1, 5))
    assert sum(len(b) for b in blocks) == 900
    assert len(blocks) < 900 / 3


def test_plan_rejects_bad_params():
    with pytest.raises(sq.SqError):
        sq.plan("H 0\n", 4, 1, 9)


DIAG = {"Z", "P", "RZ", "CP", "RZZ"}
DENSE1 = {"H", "X", "Y", "U1", "RX", "RY"}


@pytest.mark.parametrize("n,kh,seed", [(10, 1, 1), (11, 2, 2), (13, 4, 3), (16, 4, 4), (16, 3, 5), (16, 5, 6), (12, 5, 7)])
def test_register_pass_schedule(oracle, n, kh, seed):
    # the plain sweep's schedule (sq_plan mode 2, complex128: lane qubits 0-2
    # fixed, three lane positions and kh register slots for chosen qubits):
    #   * a step of several gates holds only one-qubit dense and diagonal gates,
    #     with its dense targets in qubits 0-2 plus at most kh + 3 other qubits;
    #   * every qubit sees its gates in the input order;
    #   * so the oracle computes the same state from the schedule.
    text = sq.gen_circuit("random", n, 4, seed=seed) + random_circuit(n, 200, seed=seed)
    plan = sq.plan(text, n, 2, kh)
    blocks = blocks_of(plan)
    flat = [l for b in blocks for l in b]
    src = [l for l in text.splitlines() if l.strip()]

    def norm(line):   # the plan prints reals with %.17g: compare values, not text
        tok = line.split()
        nq = len(qubits_of(line))
        return (tok[0], tuple(int(x) for x in tok[1:1 + nq]), tuple(float(x) for x in tok[1 + nq:]))

    flat_n, src_n = [norm(l) for l in flat], [norm(l) for l in src]
    assert sorted(flat_n) == sorted(src_n)
    merged = 0
    for blk in blocks:
        if len(blk) == 1:
            continue
        merged += 1
        kinds = {l.split()[0] for l in blk}
        assert kinds <= DIAG | DENSE1, kinds
        chosen = {qubits_of(l)[0] for l in blk if l.split()[0] in DENSE1} - set(range(3))
        assert len(chosen) <= kh + 3, (chosen, kh)
    assert merged > 0
    for q in range(n):
        assert [g for g in flat_n if q in g[1]] == [g for g in src_n if q in g[1]]
    s0 = random_state(n, seed)
    np.testing.assert_allclose(oracle.run(s0, n, "\n".join(flat) + "\n"), oracle.run(s0, n, text), rtol=0, atol=1e-12)


def test_register_pass_schedule_lane_moves_cut_passes():
    # the engine's schedule for the 30-qubit, 900-gate random-layer circuit:
    # choosing the top three lane qubits per pass cuts 36 passes to 25
    text = sq.gen_circuit("random", 30, 20, seed=2)
    blocks = blocks_of(sq.plan(text, 30, 2, 5))
    assert sum(len(b) for b in blocks) == 900
    assert len(blocks) <= 26, len(blocks)


def test_register_pass_schedule_cuts_passes():
    # 30-qubit random layers: about 6 register passes per layer in place of
    # 30 dense and 15 diagonal one-gate sweeps
    text = sq.gen_circuit("random", 30, 20, seed=2)
    blocks = blocks_of(sq.plan(text, 30, 2, 4))
    steps = len(blocks)
    assert sum(len(b) for b in blocks) == 900
    assert steps <= 20 * 8, steps
