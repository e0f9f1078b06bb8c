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
