"""The sampler's CPU restatement (tests/helpers.py sample_oracle), pinned by
the generator's published outputs and by what the rule fixes exactly."""
import numpy as np
import pytest

from helpers import random_state, sample_oracle, splitmix64


def test_splitmix64_published_outputs():
    # SplitMix64 with state 0: the first outputs quoted with the generator
    assert splitmix64(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_single_amplitude_state_always_gives_its_index():
    for n, k in [(4, 0), (6, 63), (11, 1500), (12, 4095)]:
        s = np.zeros(1 << n, dtype=complex)
        s[k] = 0.6 - 0.8j
        assert np.all(sample_oracle(s, 5, 200) == k)


def test_zero_amplitudes_are_never_drawn():
    n = 10
    s = random_state(n, 3)
    s[::3] = 0
    shots = sample_oracle(s, 9, 2000)
    assert np.all(shots % 3 != 0)


def test_uniform_state_is_the_uniform_inverse_cdf():
    # equal probabilities: shot = floor(u * 2^n) for u = (x >> 11) * 2^-53
    # (exact: sums of equal powers of two carry no rounding)
    n = 12
    s = np.full(1 << n, 2.0 ** (-n / 2), dtype=complex)
    shots = sample_oracle(s, 11, 500)
    want = [int(((x >> 11) * 2.0 ** -53) * (1 << n)) for x in splitmix64(11, 500)]
    assert list(shots) == want


@pytest.mark.parametrize("n", [5, 10, 13])
def test_frequencies_follow_the_probabilities(n):
    s = random_state(n, n)
    p = np.abs(s) ** 2
    shots = sample_oracle(s, 1, 20000)
    # marginal of the top qubit and of qubit 0: within 5 standard deviations
    for q in (0, n - 1):
        prob1 = p[(np.arange(p.size) >> q) & 1 == 1].sum()
        freq1 = np.mean((shots.astype(np.int64) >> q) & 1)
        sd = np.sqrt(prob1 * (1 - prob1) / shots.size)
        assert abs(freq1 - prob1) < 5 * sd + 1e-12
