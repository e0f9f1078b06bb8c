"""Host-side (no GPU) checks of the product library: ABI surface, the circuit
grammar of setGate_MEM (simulator.cpp:542-597) and the seeded generators."""
import hashlib
import math
import os
import re
import subprocess

import numpy as np
import pytest

import subquokka_dev_amd as sq
from subquokka_dev_amd import native
from helpers import GOLDEN, ROOT, qft_closed_form, random_circuit, random_state, textbook_apply

FIX = os.path.join(GOLDEN, "ref_circuits")


def test_library_exports_every_declared_symbol():
    lib_path = native.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (sq_[a-z0-9_]+)", out))
    declared = set(native.declared_symbols())
    assert declared, "no declarations found"
    assert declared <= exported, declared - exported
    sq.lib()   # loads without a GPU


def test_parse_every_gate_kind():
    n = 6
    text = random_circuit(n, 200, seed=4) + "VSWAP_1_1 1 4\nVSWAP_2_2 0 1 4 5\n"
    gates, _, ng, _ = sq.parse_circuit(text, n)
    lines = text.strip().splitlines()
    assert ng == len(lines)
    for g, line in zip(gates, lines):
        tok = line.split()
        assert native.GATE_KINDS[g.kind - 1] == tok[0]
        nq = g.nq
        assert [g.q[i] for i in range(nq)] == [int(x) for x in tok[1:1 + nq]]
        rest = [float(x) for x in tok[1 + nq:]]
        if tok[0] in ("P", "RX", "RY", "RZ", "CP", "RZZ"):
            assert g.param == rest[0]           # bit-identical double parse
        if tok[0] in ("U1", "U2"):
            assert list(g.mat)[:len(rest)] == rest


@pytest.mark.parametrize("line,code", [
    ("CZ 0 1", native.SQ_ERR_UNKNOWN_GATE),       # finder accepts CZ, setGate_MEM does not
    ("FOO 1", native.SQ_ERR_UNKNOWN_GATE),
    ("H 6", native.SQ_ERR_QUBIT_INDEX),           # check_targ_in_N
    ("H -1", native.SQ_ERR_QUBIT_INDEX),
    ("U2 3 1 " + "0 0 " * 16, native.SQ_ERR_CONTRACT),   # assert targ[0] < targ[1]
    ("CP 2 2 0.5", native.SQ_ERR_CONTRACT),
    ("VSWAP_3_3 0 1 2 3 4 5", native.SQ_ERR_UNSUPPORTED),  # broken in the reference
    ("U3 0 1 2", native.SQ_ERR_UNSUPPORTED),             # parse disabled, simulator.cpp:577
    ("RX 1", native.SQ_ERR_PARSE),
    ("U1 0 1 0 0", native.SQ_ERR_PARSE),
])
def test_parse_errors(line, code):
    g = native.SqGate()
    assert sq.lib().sq_parse_gate(line.encode(), 6, g) == code


def test_parse_reference_fixtures_plain_and_sub():
    for name, sub, want_blocks in [("ori_qft21.txt", False, None), ("sub_qft21.txt", True, True),
                                   ("sub_qaoa21.txt", True, True)]:
        text = open(os.path.join(FIX, name)).read()
        gates, blocks, ng, nb = sq.parse_circuit(text, 21, is_subcircuit=sub)
        if sub:
            lines = [l for l in text.splitlines() if l.strip()]
            counts, i = [], 0
            while i < len(lines):
                c = int(lines[i])
                counts.append(c)
                i += 1 + c
            assert [blocks[i] for i in range(nb)] == counts
            assert ng == sum(counts)
        else:
            assert ng == 21 + 210 + 10


def test_symmetric_two_qubit_gates_accept_either_order():
    a = native.SqGate()
    b = native.SqGate()
    assert sq.lib().sq_parse_gate(b"CP 4 1 0.25", 6, a) == 0
    assert sq.lib().sq_parse_gate(b"CP 1 4 0.25", 6, b) == 0


def test_gen_circuit_qft_matches_reference_order():
    # python_script/cir_qft.py:27-35 with %.17g angles parses to the same doubles
    # as the reference fixture's Python repr angles.
    ref = open(os.path.join(FIX, "ori_qft21.txt")).read().split("\n")
    mine = sq.gen_circuit("qft", 21).split("\n")
    norm = lambda l: [float(x) if "." in x else x for x in l.split()]
    assert [norm(l) for l in ref if l.strip()] == [norm(l) for l in mine if l.strip()]


def test_gen_random_layers_shape_and_determinism():
    n, layers = 30, 20
    t1 = sq.gen_circuit("random", n, layers, seed=2)
    t2 = sq.gen_circuit("random", n, layers, seed=2)
    assert t1 == t2
    assert sq.gen_circuit("random", n, layers, seed=3) != t1
    lines = t1.strip().splitlines()
    assert len(lines) == layers * (n + n // 2) == 900
    per_layer = n + n // 2
    for l in range(layers):
        blk = lines[l * per_layer:(l + 1) * per_layer]
        assert [x.split()[0] for x in blk[:n]] == ["U1"] * n
        cz = [x.split() for x in blk[n:]]
        assert all(x[0] == "CP" and float(x[3]) == math.pi and int(x[1]) < int(x[2]) for x in cz)
        qs = sorted(int(q) for x in cz for q in x[1:3])
        assert qs == list(range(n))          # a perfect matching
        for x in blk[:n]:                     # U3 matrices are unitary
            v = [float(y) for y in x.split()[2:]]
            u = np.array([complex(v[2 * i], v[2 * i + 1]) for i in range(4)]).reshape(2, 2)
            assert np.allclose(u @ u.conj().T, np.eye(2), atol=1e-15)
    # frozen input: any change to the generator changes this digest
    assert hashlib.sha256(t1.encode()).hexdigest()[:16] == open(
        os.path.join(GOLDEN, "random30_l20_seed2.sha256")).read().strip()[:16]


def test_gen_hea_structure():
    n, layers = 24, 8
    rng = np.random.default_rng(0)
    theta = rng.uniform(-math.pi, math.pi, 2 * n * layers)
    text = sq.gen_circuit("hea", n, layers, theta=theta)
    lines = text.strip().splitlines()
    assert len(lines) == layers * 3 * n
    ry = [l for l in lines if l.startswith("RY")]
    assert len(ry) == n * layers
    assert float(lines[0].split()[2]) == theta[0]


def test_hea_theta_length_is_checked_before_any_call():
    # a theta of the wrong length must never reach the C side (which reads
    # 2 * n * layers angles and writes as many gradients)
    n, layers = 6, 2
    with pytest.raises(ValueError):
        sq.gen_circuit("hea", n, layers, theta=np.zeros(2 * n * layers - 1))
    with pytest.raises(ValueError):
        sq.gen_circuit("random", n, layers, theta=np.zeros(2 * n * layers))
    eng = sq.Engine.__new__(sq.Engine)   # no device: the checks come first
    eng.h, eng.n, eng.dtype = None, n, "c128"
    for bad in (np.zeros(2 * n * layers - 1), np.zeros(2 * n * layers + 1), np.zeros((2, n * layers + 1))):
        with pytest.raises(ValueError):
            eng.hea_energy(layers, bad)
        with pytest.raises(ValueError):
            eng.adjoint_grad_hea(layers, bad)
    torch = pytest.importorskip("torch")
    from subquokka_dev_amd.autograd import hea_energy
    with pytest.raises(ValueError):
        hea_energy(torch.zeros(2 * n * layers - 2, dtype=torch.float64), eng, layers)


def test_gen_state_normalised_and_seeded():
    s1 = sq.gen_state(12, 1)
    s2 = sq.gen_state(12, 1)
    assert np.array_equal(s1, s2)
    assert abs(np.vdot(s1, s1).real - 1) < 1e-14
    assert not np.array_equal(s1, sq.gen_state(12, 2))
    # Gaussian moments (re and im i.i.d.): E|z|^2 equal split, mean ~ 0
    s = sq.gen_state(18, 5) * math.sqrt(1 << 18)
    assert abs(np.mean(s.real ** 2) - 0.5) < 0.01 and abs(np.mean(s.imag ** 2) - 0.5) < 0.01
    assert abs(np.mean(s.real)) < 0.01 and abs(np.mean(s.imag)) < 0.01
