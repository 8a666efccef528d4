"""Seeded sweep over many lengths, both builds, complex and real input: every
plan shape the planner can pick (one LDS pass, shaped rows up to 8192,
two-pass, Rader and Bluestein convolution passes with one-pass or two-pass
inner lengths) against the oracle (numpy where the oracle restates the
reference's N = p^2 defect, tests/refs.py)."""
import math

import numpy as np
import pytest

import kfft_amd as K
from refs import cfft, rfft

pytestmark = pytest.mark.gpu


def _lengths(count, seed, hi):
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(math.log(2), math.log(hi), count)).round().astype(int)
    return sorted(set(int(v) for v in n))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("half", [False, True])
def test_complex_lengths_sweep(half):
    ct = np.complex64 if half else np.complex128
    bar = 5e-7 if half else 1e-12
    bad = []
    for n in _lengths(90, 11 if half else 12, 150000):
        rng = np.random.default_rng(n)
        x = (rng.uniform(-1, 1, (2, n)) + 1j * rng.uniform(-1, 1, (2, n))).astype(ct)
        ref = np.stack([cfft(r.astype(np.complex128)) for r in x])
        got = K.CpxPlan(n, half=half)(x)
        err = _rel(got, ref)
        back = K.CpxPlan(n, K.FLAG_INVERSE, half=half)(ref.astype(ct))
        err_inv = _rel(back, np.stack([cfft(r.astype(np.complex128), True)
                                       for r in ref.astype(ct)]))
        if max(err, err_inv) > bar * max(1.0, math.log2(n)):
            bad.append((n, err, err_inv, K.CpxPlan(n, half=half).describe()))
    assert not bad, bad


@pytest.mark.parametrize("half", [False, True])
def test_real_lengths_sweep(half):
    ft = np.float32 if half else np.float64
    bar = 5e-7 if half else 1e-12
    bad = []
    for n in _lengths(60, 21 if half else 22, 150000):
        rng = np.random.default_rng(n + 1)
        x = rng.uniform(-1, 1, (2, n)).astype(ft)
        ref = np.stack([rfft(r.astype(np.float64)) for r in x])
        got = K.ScalarPlan(n, half=half).forward(x)
        err = _rel(got, ref)
        if got.shape != ref.shape or err > bar * max(1.0, math.log2(n)):
            bad.append((n, got.shape, err))
    assert not bad, bad
