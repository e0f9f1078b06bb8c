"""Adjoint gradient of the hardware-efficient ansatz (RY, RZ on every qubit,
CNOT ring; H = sum Z_i Z_{i+1}), checked against
  * the FP64 oracle's energy (forward agreement),
  * FP64 central finite differences (h = 1e-5) of the oracle energy,
  * parameter shift [E(theta + pi/2) - E(theta - pi/2)] / 2 on the GPU (complex128),
for complex128 and complex64 engines.  Bars: complex128 gradient within 1e-7
of finite differences (their truncation error) and 1e-10 of parameter shift;
complex64 within 1e-4 absolute.
"""
import math
import sys

import numpy as np
import pytest

import subquokka_dev_amd as sq

pytestmark = pytest.mark.gpu


def oracle_energy(oracle, n, layers, theta):
    text = sq.gen_circuit("hea", n, layers, theta=theta)
    s = np.zeros(1 << n, dtype=complex)
    s[0] = 1
    return oracle.expect_zz_ring(oracle.run(s, n, text), n)


@pytest.mark.parametrize("n,layers", [(4, 2), (7, 3)])
def test_forward_energy_matches_oracle(oracle, n, layers):
    rng = np.random.default_rng(n)
    theta = rng.uniform(-math.pi, math.pi, 2 * n * layers)
    want = oracle_energy(oracle, n, layers, theta)
    for dtype, tol in (("c128", 1e-12), ("c64", 1e-5)):
        eng = sq.Engine(n, dtype=dtype)
        try:
            assert abs(eng.hea_energy(layers, theta) - want) < tol
        finally:
            eng.close()


@pytest.mark.parametrize("n,layers,ckpt", [(5, 2, 0), (6, 4, 1), (8, 3, 2)])
def test_gradient_vs_finite_differences(oracle, n, layers, ckpt):
    rng = np.random.default_rng(10 * n + layers)
    theta = rng.uniform(-math.pi, math.pi, 2 * n * layers)
    h = 1e-5
    fd = np.empty_like(theta)
    for p in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (oracle_energy(oracle, n, layers, tp) - oracle_energy(oracle, n, layers, tm)) / (2 * h)
    for dtype, tol in (("c128", 1e-7), ("c64", 1e-4)):
        eng = sq.Engine(n, dtype=dtype)
        try:
            e, g = eng.adjoint_grad_hea(layers, theta, ckpt)
        finally:
            eng.close()
        assert abs(e - oracle_energy(oracle, n, layers, theta)) < (1e-12 if dtype == "c128" else 1e-5)
        assert np.max(np.abs(g - fd)) < tol, (dtype, np.max(np.abs(g - fd)))


def test_gradient_vs_parameter_shift_and_checkpoints():
    n, layers = 10, 4
    rng = np.random.default_rng(3)
    theta = rng.uniform(-math.pi, math.pi, 2 * n * layers)
    e128 = sq.Engine(n, dtype="c128")
    try:
        _, g = e128.adjoint_grad_hea(layers, theta, 0)
        ps = np.empty_like(theta)
        for p in range(theta.size):
            tp, tm = theta.copy(), theta.copy()
            tp[p] += math.pi / 2
            tm[p] -= math.pi / 2
            ps[p] = (e128.hea_energy(layers, tp) - e128.hea_energy(layers, tm)) / 2
        assert np.max(np.abs(g - ps)) < 1e-10
        _, g_ck = e128.adjoint_grad_hea(layers, theta, 1)   # re-anchoring must not change FP64 results
        assert np.max(np.abs(g_ck - g)) < 1e-12
    finally:
        e128.close()


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    from subquokka_dev_amd.autograd import hea_energy
    n, layers = 6, 2
    eng = sq.Engine(n, dtype="c128")
    try:
        theta = torch.tensor(np.random.default_rng(1).uniform(-3, 3, 2 * n * layers), dtype=torch.float64,
                             requires_grad=True)
        energy = hea_energy(theta, eng, layers)
        (2.0 * energy).backward()
        _, g = eng.adjoint_grad_hea(layers, theta.detach().numpy())
        assert np.allclose(theta.grad.numpy(), 2.0 * g, atol=1e-14, rtol=0)
    finally:
        eng.close()


def test_config4_24q_384_params_c64_vs_c128():
    # 24 qubits, 8 layers, 384 parameters: complex64 with FP64 checkpoints every
    # 3 layers against the complex128 adjoint; spot-check 16 parameters by
    # complex128 parameter shift.
    n, layers = 24, 8
    theta = np.random.default_rng(4).uniform(-math.pi, math.pi, 2 * n * layers)
    print("config 4 adjoint", flush=True)
    e64 = sq.Engine(n, dtype="c64")
    e128 = sq.Engine(n, dtype="c128")
    try:
        en64, g64 = e64.adjoint_grad_hea(layers, theta, 3)
        en128, g128 = e128.adjoint_grad_hea(layers, theta, 0)
        assert abs(en64 - en128) < 1e-4
        assert np.max(np.abs(g64 - g128)) < 1e-3
        for p in np.random.default_rng(5).choice(theta.size, 16, replace=False):
            tp, tm = theta.copy(), theta.copy()
            tp[p] += math.pi / 2
            tm[p] -= math.pi / 2
            ps = (e128.hea_energy(layers, tp) - e128.hea_energy(layers, tm)) / 2
            assert abs(ps - g128[p]) < 1e-9
    finally:
        e64.close()
        e128.close()


# ---------------------------------------------------------------------------
# general parameterised circuits (sq_adjoint_grad)
# ---------------------------------------------------------------------------

PARAM_KINDS = ["RX", "RY", "RZ", "P", "CP", "RZZ"]
FIXED_KINDS = ["H", "X", "Y", "Z", "U1", "U3", "U2", "SWAP", "CP", "RZ", "RY"]


def param_circuit(n, n_gates, n_theta, seed, share=False):
    """Random circuit lines and param indices: about half the gates take an
    angle from theta (each parameter once unless share), the others fixed."""
    import random as _random
    from helpers import gate_line
    rng = _random.Random(seed)
    nrng = np.random.default_rng(seed)
    lines, param = [], []
    nxt = 0
    for _ in range(n_gates):
        if rng.random() < 0.5 and (share or nxt < n_theta):
            k = rng.choice(PARAM_KINDS)
            qs = sorted(rng.sample(range(n), 2)) if k in ("CP", "RZZ") else [rng.randrange(n)]
            lines.append(f"{k} {' '.join(map(str, qs))} 0.0")
            if share:
                param.append(rng.randrange(n_theta))
            else:
                param.append(nxt)
                nxt += 1
        else:
            k = rng.choice(FIXED_KINDS)
            qs = sorted(rng.sample(range(n), 2)) if k in ("U2", "SWAP", "CP") else [rng.randrange(n)]
            lines.append(gate_line(k, qs, rng, nrng))
            param.append(-1)
    return lines, np.array(param, dtype=np.int32), (n_theta if share else nxt)


def bind(lines, param, theta):
    out = []
    for line, p in zip(lines, param):
        if p < 0:
            out.append(line)
        else:
            t = line.split()
            out.append(" ".join(t[:-1] + [repr(float(theta[p]))]))
    return "\n".join(out) + "\n"


def z_energy(state, n, terms):
    idx = np.arange(1 << n, dtype=np.uint64)
    h = np.zeros(1 << n)
    for coef, qs in terms:
        mask = np.uint64(sum(1 << q for q in qs))
        par = np.zeros(1 << n, dtype=np.uint64)
        x = idx & mask
        while np.any(x):
            par ^= x & np.uint64(1)
            x = x >> np.uint64(1)
        h += np.where(par == 1, -coef, coef)
    return float(np.sum(np.abs(state) ** 2 * h))


def random_terms(n, seed):
    rng = np.random.default_rng(seed)
    terms = [(0.3, [])]
    for _ in range(6):
        k = int(rng.integers(1, 4))
        terms.append((float(rng.normal()), sorted(rng.choice(n, size=k, replace=False).tolist())))
    return terms


def oracle_param_energy(oracle, n, lines, param, theta, terms, s0):
    return z_energy(oracle.run(s0, n, bind(lines, param, theta)), n, terms)


@pytest.mark.parametrize("n,ckpt", [(6, 0), (9, 3), (12, 5)])
def test_general_gradient_vs_fd_and_shift(oracle, n, ckpt):
    lines, param, n_theta = param_circuit(n, 60, 24, seed=n)
    rng = np.random.default_rng(n + 1)
    theta = rng.uniform(-math.pi, math.pi, n_theta)
    terms = random_terms(n, n)
    from helpers import random_state
    s0 = random_state(n, n + 2)
    h = 1e-5
    fd = np.empty(n_theta)
    shift = np.empty(n_theta)
    for p in range(n_theta):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (oracle_param_energy(oracle, n, lines, param, tp, terms, s0) -
                 oracle_param_energy(oracle, n, lines, param, tm, terms, s0)) / (2 * h)
        tp, tm = theta.copy(), theta.copy()
        tp[p] += math.pi / 2
        tm[p] -= math.pi / 2
        # every parameter drives one gate whose generator spans one unit: exact
        shift[p] = (oracle_param_energy(oracle, n, lines, param, tp, terms, s0) -
                    oracle_param_energy(oracle, n, lines, param, tm, terms, s0)) / 2
    want_e = oracle_param_energy(oracle, n, lines, param, theta, terms, s0)
    text = "\n".join(lines) + "\n"
    for dtype, tol_fd, tol_shift, tol_e in (("c128", 1e-7, 1e-9, 1e-12), ("c64", 1e-4, 1e-4, 1e-5)):
        eng = sq.Engine(n, dtype=dtype)
        try:
            eng.set_state(s0.astype(np.complex128 if dtype == "c128" else np.complex64))
            e, g = eng.adjoint_grad(text, param, theta, terms, ckpt)
            back = eng.get_state()
            e_only = eng.expect_z(terms)
        finally:
            eng.close()
        assert abs(e - want_e) < tol_e, (dtype, e, want_e)
        assert np.max(np.abs(g - fd)) < tol_fd, (dtype, np.max(np.abs(g - fd)))
        assert np.max(np.abs(g - shift)) < tol_shift, (dtype, np.max(np.abs(g - shift)))
        # the state is the one the call started from: exactly with checkpoints
        # (the first copy); without, up to the reverse sweep's rounding and the
        # reference's H literal (H^dag H = 2 * 0.70710678118^2 = 1 - 2.6e-12)
        assert np.max(np.abs(back - s0)) < ((1e-12 if ckpt else 1e-10) if dtype == "c128" else 2e-6)
        assert abs(e_only - z_energy(back, n, terms)) < tol_e


def test_general_gradient_with_shared_parameters(oracle):
    # a parameter driving several gates gets the sum of their terms
    n = 8
    lines, param, n_theta = param_circuit(n, 70, 5, seed=3, share=True)
    theta = np.random.default_rng(4).uniform(-math.pi, math.pi, n_theta)
    terms = random_terms(n, 5)
    s0 = np.zeros(1 << n, dtype=complex)
    s0[0] = 1
    h = 1e-5
    fd = np.empty(n_theta)
    for p in range(n_theta):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (oracle_param_energy(oracle, n, lines, param, tp, terms, s0) -
                 oracle_param_energy(oracle, n, lines, param, tm, terms, s0)) / (2 * h)
    eng = sq.Engine(n)
    try:
        eng.init_zero()
        _, g = eng.adjoint_grad("\n".join(lines) + "\n", param, theta, terms, 2)
    finally:
        eng.close()
    assert np.max(np.abs(g - fd)) < 1e-7


def test_general_path_equals_hea_entry():
    # the hardware-efficient ansatz through the general entry point (CNOT as U2)
    n, layers = 8, 3
    theta = np.random.default_rng(9).uniform(-math.pi, math.pi, 2 * n * layers)
    text = sq.gen_circuit("hea", n, layers, theta=theta)
    param = []
    k = 0
    for line in text.strip().splitlines():
        if line.startswith(("RY", "RZ")):
            param.append(k)
            k += 1
        else:
            param.append(-1)
    terms = [(1.0, [q, (q + 1) % n]) for q in range(n)]
    eng = sq.Engine(n)
    try:
        e_h, g_h = eng.adjoint_grad_hea(layers, theta, 1)
        eng.init_zero()
        e_g, g_g = eng.adjoint_grad(text, np.array(param), theta, terms, 4)
    finally:
        eng.close()
    assert abs(e_h - e_g) < 1e-12
    assert np.max(np.abs(g_h - g_g)) < 1e-11


def test_general_autograd_wrapper():
    torch = pytest.importorskip("torch")
    from subquokka_dev_amd.autograd import circuit_energy
    n = 6
    lines, param, n_theta = param_circuit(n, 40, 10, seed=21)
    terms = random_terms(n, 21)
    text = "\n".join(lines) + "\n"
    theta = torch.tensor(np.random.default_rng(2).uniform(-1, 1, n_theta), dtype=torch.float64, requires_grad=True)
    eng = sq.Engine(n)
    try:
        energy = circuit_energy(theta, eng, text, param, terms, 3)
        energy.backward()
        eng.init_zero()
        e2, g = eng.adjoint_grad(text, param, theta.detach().numpy(), terms, 3)
    finally:
        eng.close()
    assert abs(energy.item() - e2) < 1e-12
    assert np.max(np.abs(theta.grad.numpy() - g)) < 1e-12


def test_general_rejects_bad_parameters():
    n = 4
    eng = sq.Engine(n)
    try:
        with pytest.raises(ValueError):
            eng.adjoint_grad("RX 0 0.1\n", [3], np.zeros(2), [(1.0, [0])])
        with pytest.raises(sq.SqError):
            eng.adjoint_grad("H 0\n", [0], np.zeros(1), [(1.0, [0])])   # H takes no angle
    finally:
        eng.close()
