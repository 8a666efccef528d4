"""CPU oracle of libkfft's public math and shift helpers (test infrastructure only).

Plain-Python / numpy restatements of the reference's loops, used by tests/ as the
checker of the library's host and GPU paths:

- ``shift``          kfft_shift.c:4-61 (element moves of the in-place loops)
- ``transpose_ip``   kfft_math.c:80-100 (the in-place cycle-walk transpose)
- ``sqrt_bisect``    math/kfft_custom_math.h:43-67 (the default build's square
                     root: bisection until the bracket is <= 0.001 wide)
- ``magnitude``      kfft_math.c:123-130 with that root
- ``average_serial`` kfft_math.c:140-148 run on one thread (a plain sum in the
                     build's precision, then divided by the size)

Pins: ``shift`` against numpy's fftshift / ifftshift, ``transpose_ip`` against
the matrix transpose, ``sqrt_bisect`` against the bracket bound |err| <= 5e-4
and exact squares (tests/test_oracle.py).
"""
from __future__ import annotations

import numpy as np


def shift(buf, inverse: bool):
    """kfft_shift_cpx / _scalar on a copy of the 1D array `buf`."""
    b = list(np.asarray(buf).copy())
    size = len(b)
    c = size // 2
    if size % 2 == 0:
        for k in range(c):
            b[k], b[k + c] = b[k + c], b[k]
    elif inverse:
        tmp = b[size - 1]
        for k in range(c - 1, -1, -1):
            b[c + k + 1] = b[k]
            b[k] = b[c + k]
        b[c] = tmp
    else:
        tmp = b[0]
        for k in range(c):
            b[k] = b[c + k + 1]
            b[c + k + 1] = b[k + 1]
        b[c] = tmp
    return np.array(b, dtype=np.asarray(buf).dtype)


def transpose_ip(buf, x: int, y: int):
    """kfft_math_transpose_ip_cpx / _scalar on a copy of the flat array `buf`
    holding y rows of x values; returns the flat result (x rows of y)."""
    f = list(np.asarray(buf).copy())
    r, c = y, x
    if r <= 1 or c <= 1:
        return np.array(f, dtype=np.asarray(buf).dtype)
    last = r * c - 2
    while c > 1:
        for i in range(r - 2, -1, -1):
            ind = i * c + (c - 1)
            v = f[ind]
            while ind < last:
                f[ind] = f[ind + 1]
                ind += 1
            f[last] = v
            last -= 1
        last -= 1
        c -= 1
    return np.array(f, dtype=np.asarray(buf).dtype)


def sqrt_bisect(number, dtype=np.float64):
    """The default build's square root, elementwise."""
    x = np.asarray(number, dtype=dtype)
    out = np.zeros_like(x)
    acc = dtype(0.001)
    for idx in np.ndindex(x.shape):
        n = x[idx]
        if n <= 0:
            continue
        lo = hi = dtype(1)
        if n < 1:
            lo = n
        else:
            hi = n
        while hi - lo > acc:
            g = (lo + hi) / dtype(2)
            if g * g > n:
                hi = g
            else:
                lo = g
        out[idx] = (lo + hi) / dtype(2)
    return out


def magnitude(z, dtype=np.float64):
    z = np.asarray(z)
    return sqrt_bisect(z.real.astype(dtype) ** 2 + z.imag.astype(dtype) ** 2, dtype)


def average_serial(x, dtype=np.float64):
    acc = dtype(0)
    for v in np.asarray(x, dtype=dtype):
        acc = dtype(acc + v)
    return dtype(acc / dtype(len(x)))
