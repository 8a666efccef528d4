"""Reference transforms for the parity tests: the FP64 oracle, except where it
restates the reference's defect for N = p^2 (its bisection sqrt truncates
below p, so it keeps N whole as a Rader "prime" and returns a wrong spectrum);
numpy's pocketfft FP64 stands in for those lengths."""
import numpy as np

import oracle as O


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    d = 2
    while d * d <= n:
        if n % d == 0:
            return False
        d += 1
    return True


def oracle_defined(n: int) -> bool:
    stages, _ = O.factor(n)
    return not (len(stages) == 1 and stages[0][0] == n and not is_prime(n))


def cfft(x, inverse=False):
    n = len(x)
    if not oracle_defined(n):
        return np.fft.ifft(x) if inverse else np.fft.fft(x)
    return O.fft(x, O.FLAG_INVERSE if inverse else 0)


def rfft(x):
    """The reference's real forward: n/2+1 bins for even n, n for odd n."""
    n = len(x)
    basis = n // 2 if n % 2 == 0 else n
    if not oracle_defined(basis):
        return np.fft.rfft(x) if n % 2 == 0 else np.fft.fft(x)
    return O.rfft(x)
