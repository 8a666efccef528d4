"""CPU checks of the C-ABI libraries: loading, exported symbols, struct layout,
service functions, pool / lenmem semantics and host-side helpers.  No GPU is
needed; the plan-creation paths that need one are exercised in test_gpu_*.py.
"""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import kfft_amd as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

u8, u32, vp = C.c_uint8, C.c_uint32, C.c_void_p


class PlanRader(C.Structure):
    _fields_ = [("prime", u32), ("p", u32), ("q", u32), ("qidx", vp), ("pidx", vp),
                ("shuffle_twiddles", vp), ("plan", vp), ("plan_inv", vp)]


class PlanCpx(C.Structure):
    """Mirror of kfft_plan_cpx (include/kfft_cpx.h; reference incs/kfft_cpx.h:46-65)."""
    _fields_ = [("mmgr", vp), ("nfft", u32), ("level", u8), ("flags", u32), ("fac_count", u8),
                ("factors", u32 * 64), ("prm_count", u8), ("primes", PlanRader * 32),
                ("twiddles", vp), ("gpu", vp)]


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        for m in re.finditer(r"KFFT_API\s+[^;{]*?\b(kfft_\w+)\s*\(", text, re.S):
            names.add(m.group(1))
    return names


@pytest.mark.parametrize("half", [False, True])
def test_library_loads_and_exports_declared_symbols(half):
    lib = K.load(half)
    declared = declared_symbols()
    assert len(declared) >= 50
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(K.EXPORTS) <= declared


@pytest.mark.parametrize("half", [False, True])
def test_info_and_service(half):
    ver, flags = K.info(half)
    assert ver == (0, 8, 1)
    assert flags & (1 << 5)  # KFFT_INFO_RADER_ALGO
    assert bool(flags & (1 << 7)) == half  # KFFT_INFO_HALF_SCALAR
    lib = K.load(half)
    assert lib.kfft_strerr(0) == b"Success operation result"
    assert lib.kfft_strerr(4) == b"Use improper real function for this plan"

    def smooth(m):
        for q in (2, 3, 5):
            while m % q == 0:
                m //= q
        return m == 1

    for n in [1, 7, 11, 97, 1009, 4097, 65537]:
        assert lib.kfft_next_fast_size(n) == next(m for m in range(n, 4 * n) if smooth(m))


def test_plan_struct_layout_and_lenmem():
    lib = K.load(False)
    need = C.c_size_t(0)
    assert lib.kfft_config_cpx(0, 0, None, C.byref(need)) is None
    # test/unit/suites/api_cpx1.c:26-36: the size query for N=0 is sizeof(kfft_plan_cpx)
    assert need.value == C.sizeof(PlanCpx)
    need101 = C.c_size_t(0)
    lib.kfft_config_cpx(101, 0, None, C.byref(need101))
    assert need101.value == C.sizeof(PlanCpx) + 101 * 16  # plan plus its twiddle table
    # a pool given with too little room reports the need and returns NULL (incs/kfft_algo.h:27-35)
    pool = lib.kfft_pool_create(64)
    small = C.c_size_t(8)
    assert lib.kfft_config_cpx(101, 0, pool, C.byref(small)) is None
    assert small.value == need101.value
    lib.kfft_pool_free(pool)
    # the float build's complex type is 8 bytes
    lh = K.load(True)
    needh = C.c_size_t(0)
    lh.kfft_config_cpx(101, 0, None, C.byref(needh))
    assert needh.value - C.sizeof(PlanCpx) == 101 * 8


def test_pool_semantics():
    """Bump allocator of kfft_pool.c:13-96 (as pinned by test/unit/suites/sys_pool.c)."""
    lib = K.load(False)
    A = lib.kfft_pool_create(100)
    assert lib.kfft_pool_empty(A) == 100
    p1 = lib.kfft_pool_alloc(A, 60)
    p2 = lib.kfft_pool_alloc(A, 40)
    assert p1 and p2 and p2 - p1 == 60
    assert lib.kfft_pool_empty(A) == 0
    assert lib.kfft_pool_alloc(A, 1) is None
    assert lib.kfft_pool_alloc(A, 0) is None
    lib.kfft_pool_clear(A)
    assert lib.kfft_pool_empty(A) == 100
    assert lib.kfft_pool_alloc(A, 100) == p1
    lib.kfft_pool_free(A)


def test_cleanup_null():
    lib = K.load(False)
    obj = (vp * 1)(None)
    assert lib.kfft_cleanup(C.cast(obj, vp)) == K.RET_FREE_NULL


@pytest.mark.parametrize("n", [1, 2, 5, 8, 9])
def test_shift_host(n):
    """kfft_shift_cpx/_scalar (kfft_shift.c:4-61) equal numpy's fftshift / ifftshift."""
    lib = K.load(False)
    x = np.arange(n, dtype=np.float64) + 0.5
    for inverse, ref in [(False, np.fft.fftshift(x)), (True, np.fft.ifftshift(x))]:
        b = x.copy()
        lib.kfft_shift_scalar(C.c_void_p(b.ctypes.data), n, inverse, None)
        assert np.array_equal(b, ref)
        c = (x + 1j * x).copy()
        lib.kfft_shift_cpx(C.c_void_p(c.ctypes.data), n, inverse, None)
        assert np.array_equal(c, ref + 1j * ref)


def test_shift2_and_math_host():
    lib = K.load(False)
    y, x = 5, 6
    a = (np.arange(x * y) + 1j * np.arange(x * y)[::-1]).reshape(y, x).astype(np.complex128)
    b = a.copy()
    lib.kfft_shift2_cpx(C.c_void_p(b.ctypes.data), None, x, y, False, None)
    assert np.array_equal(b, np.fft.fftshift(a))
    o = np.empty(x * y, np.complex128)
    lib.kfft_math_transpose_cpx(C.c_void_p(a.ctypes.data), C.c_void_p(o.ctypes.data), x, y)
    assert np.array_equal(o.reshape(x, y), a.T)
    h = a.copy()
    lib.kfft_math_hadamard_cpx(C.c_void_p(h.ctypes.data), C.c_void_p(a.ctypes.data), x * y)
    assert np.allclose(h, a * a, rtol=1e-15)


def test_bench_radix16_stage_count():
    """bench.py's matrix-core work count: radix-16 stages of the shaped passes."""
    import bench

    assert bench._radix16_stages("direct n=65536 inner=65536 four-step 256x256") == 4
    assert bench._radix16_stages("direct n=4096 inner=4096 one-pass") == 3
    assert bench._radix16_stages("direct n=512 inner=512 one-pass") == 1  # 16, 8, 4
    assert bench._radix16_stages("direct n=32 inner=32 one-pass") == 0    # 8, 4
    assert bench._radix16_stages("rader n=1009 inner=1008 one-pass") == 0


@pytest.mark.parametrize("half", [False, True])
def test_math_helpers_host(half):
    """kfft_math_transpose_ip_*, kfft_math_magnitude(_ip), kfft_math_average
    (incs/kfft_math.h:303-350) on host buffers against the oracle's
    restatements: the in-place transpose equals the reference's cycle walk,
    magnitudes are within the reference's bisection root (5e-4 absolute) and
    within rounding of the exact root, the average is the serial mean."""
    import oracle.mathops as M

    lib = K.load(half)
    ft = np.float32 if half else np.float64
    ct = np.complex64 if half else np.complex128
    rng = np.random.default_rng(3)
    for x, y in [(3, 5), (4, 4), (7, 2), (1, 6)]:
        a = (rng.uniform(-1, 1, x * y) + 1j * rng.uniform(-1, 1, x * y)).astype(ct)
        b = a.copy()
        lib.kfft_math_transpose_ip_cpx(C.c_void_p(b.ctypes.data), x, y)
        assert np.array_equal(b, M.transpose_ip(a, x, y))
        r = a.real.copy()
        lib.kfft_math_transpose_ip_scalar(C.c_void_p(r.ctypes.data), x, y)
        assert np.array_equal(r, M.transpose_ip(a.real, x, y))
    z = (rng.uniform(-3, 3, 257) + 1j * rng.uniform(-3, 3, 257)).astype(ct)
    m = np.empty(257, ft)
    lib.kfft_math_magnitude(C.c_void_p(z.ctypes.data), C.c_void_p(m.ctypes.data), 257)
    exact = np.abs(z.astype(np.complex128))
    assert np.max(np.abs(m - exact) / exact) <= (2e-7 if half else 1e-15)
    assert np.max(np.abs(m - M.magnitude(z, ft))) <= 5e-4 + 1e-6
    zi = z.copy()
    lib.kfft_math_magnitude_ip(C.c_void_p(zi.ctypes.data), 257)
    assert np.array_equal(zi.view(ft)[:257], m)
    s = rng.uniform(-1, 1, 1001).astype(ft)
    avg = lib.kfft_math_average(C.c_void_p(s.ctypes.data), 1001)
    assert abs(avg - M.average_serial(s, np.float64)) <= (1e-7 if half else 1e-16)


def test_simd_info():
    """kfft_simd_info / kfft_simd_analize (simd/kfft_simd_check.h:58-66): the
    build uses no CPU extensions (its transforms run on the GPU); the host
    analysis reports an x86-64 CPU with at least its baseline SSE2."""
    import platform

    lib = K.load(False)
    i = lib.kfft_simd_info()
    a = lib.kfft_simd_analize()
    if platform.machine() == "x86_64":
        assert i.arch == 2 and i.ext == 0  # HW_ARCH_X64
        assert a.arch == 2
        assert a.ext & (1 << 1) and a.ext & (1 << 10)  # HW_x64, HW_SSE2


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_public_headers_compile(lang, tmp_path):
    """kfft.h with the tiny API (kfft_tiny.h) and the SIMD-info header compile
    as C and as C++ with warnings as errors."""
    import shutil
    import subprocess

    cc = "gcc" if lang == "c" else "g++"
    if not shutil.which(cc):
        pytest.skip(f"{cc} not available")
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text('#include "kfft_tiny.h"\n#include "kfft_simd_check.h"\n'
                   "int main(void) { kfft_args_norm a = {16}; (void)a;\n"
                   "  return (int)kfft_simd_align(kfft_simd_info()); }\n")
    r = subprocess.run([cc, "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
