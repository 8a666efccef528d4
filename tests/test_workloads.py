"""The synthetic ragged workload generator (CPU)."""
import numpy as np

from kfft_amd.workloads import _is_prime, ragged_batch


def test_ragged_distribution():
    n, k = ragged_batch(20000)
    assert n.min() >= 16 and n.max() <= 65536
    primes = np.array([_is_prime(int(v)) for v in n])
    pow2 = (n & (n - 1)) == 0
    assert 0.15 <= primes.mean() <= 0.30   # forced share plus naturally prime draws
    assert 0.29 <= pow2.mean() <= 0.32
    assert 0.48 <= k.mean() <= 0.52
    # log-uniform: about equal counts per octave among the unforced draws
    rest = n[~primes & ~pow2]
    counts = np.histogram(np.log2(rest), bins=np.arange(4, 17))[0]
    assert counts.min() > 0.6 * counts.mean()
    assert 7000 < n.mean() < 9000  # mean length of the log-uniform draw, about 7.9k


def test_ragged_is_seeded():
    a = ragged_batch(100, seed=3)
    b = ragged_batch(100, seed=3)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
