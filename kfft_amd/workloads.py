"""Synthetic workloads of the benchmark and the tests (inputs only, no arithmetic)."""
from __future__ import annotations

import math

import numpy as np


def _is_prime(n: int) -> bool:
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    d = 3
    while d * d <= n:
        if n % d == 0:
            return False
        d += 2
    return True


def _prime_near(n: int, lo: int, hi: int) -> int:
    p = max(n, lo)
    while p <= hi and not _is_prime(p):
        p += 1
    if p > hi:
        p = min(n, hi)
        while not _is_prime(p):
            p -= 1
    return p


def ragged_batch(count: int = 20000, seed: int = 20260104, lo: int = 16, hi: int = 65536,
                 prime_frac: float = 0.15, pow2_frac: float = 0.30, real_frac: float = 0.5):
    """Lengths drawn log-uniform over [lo, hi]; a `prime_frac` share forced to a
    prime (the next one up), a `pow2_frac` share to the nearest power of two;
    a `real_frac` share of the transforms takes real input, the rest complex.
    Returns (lengths uint32, kinds uint8: 0 complex, 1 real)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(math.log(lo), math.log(hi), count)
    n = np.exp(u).round().astype(np.int64)
    sel = rng.uniform(0, 1, count)
    for i in range(count):
        if sel[i] < prime_frac:
            n[i] = _prime_near(int(n[i]), lo, hi)
        elif sel[i] < prime_frac + pow2_frac:
            n[i] = 1 << int(round(math.log2(n[i])))
    n = np.clip(n, lo, hi)
    kinds = (rng.uniform(0, 1, count) < real_frac).astype(np.uint8)
    return n.astype(np.uint32), kinds
