"""Measurement sampling on the GPU (sq_sample) against its CPU restatement."""
import numpy as np
import pytest

import subquokka_dev_amd as sq
from helpers import random_state, sample_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [4, 7, 10, 13, 16])
def test_sampler_matches_oracle_exactly(dtype, n):
    s0 = random_state(n, 40 + n)
    s0[1::5] = 0
    eng = sq.Engine(n, dtype=dtype)
    try:
        eng.set_state(s0.astype(np.complex128 if dtype == "c128" else np.complex64))
        held = eng.get_state()          # the amplitudes the GPU holds, exactly
        shots = eng.sample(3000, seed=n)
        again = eng.sample(3000, seed=n)
    finally:
        eng.close()
    assert np.array_equal(shots, again)
    assert np.array_equal(shots, sample_oracle(held, n, 3000))


def test_sampler_26q_marginals_and_determinism():
    n = 26
    eng = sq.Engine(n, dtype="c128")
    try:
        eng.apply_text(sq.gen_circuit("random", n, 3, seed=7))
        st = eng.get_state()
        shots = eng.sample(200000, seed=3)
        head = eng.sample(1000, seed=3)
    finally:
        eng.close()
    assert np.array_equal(head, shots[:1000])
    assert shots.max() < (1 << n)
    p = np.abs(st) ** 2
    p /= p.sum()
    idx = np.arange(p.size)
    for q in (0, 5, 13, 25):
        prob1 = p[(idx >> q) & 1 == 1].sum()
        freq1 = np.mean((shots.astype(np.int64) >> q) & 1)
        sd = np.sqrt(prob1 * (1 - prob1) / shots.size)
        assert abs(freq1 - prob1) < 5 * sd, (q, freq1, prob1)
    # spot-check 2000 shots against the exact rule
    np.testing.assert_array_equal(shots[:2000], sample_oracle(st, 3, 2000))
