"""Shift and math helpers on device buffers (the GPU paths of kfft_shift_*,
kfft_shift2_*, kfft_math_transpose_ip_*, kfft_math_magnitude(_ip) and
kfft_math_average), checked against the oracle's restatements of the
reference's loops and against numpy's closed forms."""
import ctypes as C

import numpy as np
import pytest

import kfft_amd as K
import oracle.mathops as M

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _p(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("n", [1, 2, 7, 8, 1001, 4096, 100003])
def test_shift_device(half, n):
    lib = K.load(half)
    ct = np.complex64 if half else np.complex128
    ft = np.float32 if half else np.float64
    x = (np.arange(n) + 1j * np.arange(n)[::-1]).astype(ct)
    for inverse in (False, True):
        ref = np.fft.ifftshift(x) if inverse else np.fft.fftshift(x)
        if n <= 1001:
            assert np.array_equal(ref, M.shift(x, inverse))
        d = _dev(x)
        lib.kfft_shift_cpx(_p(d), n, inverse, None)
        assert np.array_equal(d.cpu().numpy(), ref)
        r = _dev(x.real.astype(ft))
        lib.kfft_shift_scalar(_p(r), n, inverse, None)
        assert np.array_equal(r.cpu().numpy(), ref.real)


@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("y,x", [(4, 6), (5, 6), (4, 7), (5, 7), (1, 9), (8, 1), (1024, 768),
                                 (1023, 769)])
@pytest.mark.parametrize("with_tmp", [False, True])
def test_shift2_device(half, y, x, with_tmp):
    """One 2D roll (rows then columns, ext/2d/kfft_cpx2.c:188-240) equals the
    host path and numpy's 2D fftshift / ifftshift."""
    import torch

    lib = K.load(half)
    ct = np.complex64 if half else np.complex128
    rng = np.random.default_rng(x * 7 + y)
    a = (rng.uniform(-1, 1, (y, x)) + 1j * rng.uniform(-1, 1, (y, x))).astype(ct)
    for inverse in (False, True):
        ref = np.fft.ifftshift(a) if inverse else np.fft.fftshift(a)
        d = _dev(a)
        tmp = torch.empty_like(d) if with_tmp else None
        lib.kfft_shift2_cpx(_p(d), _p(tmp) if with_tmp else None, x, y, inverse, None)
        assert np.array_equal(d.cpu().numpy(), ref)
        h = a.copy()
        lib.kfft_shift2_cpx(C.c_void_p(h.ctypes.data), None, x, y, inverse, None)
        assert np.array_equal(h, ref)
        r = _dev(a.real.copy())
        lib.kfft_shift2_scalar(_p(r), None, x, y, inverse, None)
        assert np.array_equal(r.cpu().numpy(), ref.real)


@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("y,x", [(5, 3), (64, 48), (1000, 37)])
def test_transpose_ip_device(half, y, x):
    lib = K.load(half)
    ct = np.complex64 if half else np.complex128
    rng = np.random.default_rng(y + x)
    a = (rng.uniform(-1, 1, x * y) + 1j * rng.uniform(-1, 1, x * y)).astype(ct)
    ref = a.reshape(y, x).T.reshape(-1)
    if x * y <= 2000:
        assert np.array_equal(M.transpose_ip(a, x, y), ref)
    d = _dev(a)
    lib.kfft_math_transpose_ip_cpx(_p(d), x, y)
    assert np.array_equal(d.cpu().numpy(), ref)
    r = _dev(a.real.copy())
    lib.kfft_math_transpose_ip_scalar(_p(r), x, y)
    assert np.array_equal(r.cpu().numpy(), ref.real)


@pytest.mark.parametrize("half", [False, True])
def test_magnitude_and_average_device(half):
    import torch

    lib = K.load(half)
    ct = np.complex64 if half else np.complex128
    ft = np.float32 if half else np.float64
    rng = np.random.default_rng(5)
    n = 300007
    z = (rng.uniform(-3, 3, n) + 1j * rng.uniform(-3, 3, n)).astype(ct)
    exact = np.abs(z.astype(np.complex128))
    d = _dev(z)
    m = torch.empty(n, dtype=torch.float32 if half else torch.float64, device="cuda")
    lib.kfft_math_magnitude(_p(d), _p(m), n)
    got = m.cpu().numpy()
    # sqrt(re^2 + im^2): two roundings before a rounded root, within 2 ulp of
    # the correctly rounded |z| in FP64 (the FP32 root is within about 1 ulp more)
    assert np.max(np.abs(got - exact) / exact) <= (4e-7 if half else 4.5e-16)
    k = 2000  # the reference's bisection root on a sample: within 5e-4 absolute
    assert np.max(np.abs(got[:k] - M.magnitude(z[:k], ft))) <= 5e-4 + 1e-6
    lib.kfft_math_magnitude_ip(_p(d), n)
    assert np.array_equal(d.view(torch.float32 if half else torch.float64)[:n].cpu().numpy(), got)
    s = rng.uniform(-1, 1, 1 << 20).astype(ft)
    avg = lib.kfft_math_average(_p(_dev(s)), 1 << 20)
    assert abs(avg - np.mean(s.astype(np.float64))) <= (1e-7 if half else 1e-15)
    assert lib.kfft_math_average(_p(_dev(s[:7])), 7) == pytest.approx(M.average_serial(s[:7]),
                                                                        abs=1e-7 if half else 1e-16)


def test_tiny_api_program():
    """tests/c/tiny_check.c (built by build()): the tiny API's one-handle plans
    of every 1D / 2D / convolution kind through the C headers, checked in C
    against brute-force DFTs and direct convolutions."""
    import os
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tests", "c", "build", "tiny_check")
    if not os.path.exists(exe):
        pytest.skip("tests/c programs not built (run __graft_entry__.build())")
    print("running", exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "tiny api: ok" in r.stdout
