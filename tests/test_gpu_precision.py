"""The precision form (float build): radix-16 stages as split-operand FP16 /
BF16 matrix products with FP32 accumulation, FP64-generated twiddles, and the
optional refinement step whose residual is evaluated in FP64.

Bar: relative L2 error <= 1e-6 against the FP64 oracle (the precision target
of the 2^16 x 4096 workload), on seeded uniform[-1,1) inputs and on unit noise
carrying 1e-6-amplitude tones.  With one refinement step the error drops to
the rounding of the FP32 output (<= 1.5e-7 here).
"""
import math

import numpy as np
import pytest

import kfft_amd as K
import oracle as O

pytestmark = pytest.mark.gpu

SPLIT = {"f16": K.FLAG_GPU_SPLIT_F16, "bf16": K.FLAG_GPU_SPLIT_BF16}
BAR = 1e-6
REFINED_BAR = 1.5e-7


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cplx32(rng, *shape):
    return (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)


def tones32(rng, batch, n, amp=1e-6, count=5):
    """Unit noise plus `count` tones of amplitude `amp` per instance."""
    x = cplx32(rng, batch, n).astype(np.complex128)
    t = np.arange(n)
    for b in range(batch):
        for f in rng.integers(0, n, count):
            x[b] += amp * np.exp(2j * np.pi * f * t / n)
    return x.astype(np.complex64)


# lengths with radix-16 stages: one-pass 16..8192, two-pass 2^16 and 2^20
SIZES = [16, 64, 256, 512, 1024, 2048, 4096, 8192, 1 << 16, 1 << 20]


@pytest.mark.parametrize("form", ["f16", "bf16"])
@pytest.mark.parametrize("n", SIZES)
def test_split_forward_inverse(form, n):
    rng = np.random.default_rng(n + (7 if form == "f16" else 11))
    batch = 37 if n <= 4096 else 2
    x = cplx32(rng, batch, n)
    xd = x.astype(np.complex128)
    ref = O.fft(xd)
    fwd = K.CpxPlan(n, SPLIT[form], half=True)
    X = fwd(x)
    assert rel(X, ref) <= BAR, (form, n, rel(X, ref))
    # the split arithmetic ran (not the vector-unit FP32 path): outputs differ in rounding
    plain = K.CpxPlan(n, half=True)(x)
    assert not np.array_equal(X, plain)
    inv = K.CpxPlan(n, SPLIT[form] | K.FLAG_INVERSE, half=True)
    Xr = ref.astype(np.complex64)
    assert rel(inv(Xr), O.fft(Xr.astype(np.complex128), O.FLAG_INVERSE)) <= BAR


@pytest.mark.parametrize("form", ["f16", "bf16"])
def test_split_tones_at_full_size(form):
    """2^16 points, 1e-6 tones atop unit noise: the split products keep the
    error at the FP32 level, so the tones survive as in FP64."""
    import torch

    n, batch = 1 << 16, 8
    rng = np.random.default_rng(20260103)
    x = tones32(rng, batch, n)
    ref = O.fft(x.astype(np.complex128))
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    K.CpxPlan(n, SPLIT[form], half=True)(xd, out=yd)
    assert rel(yd.cpu().numpy(), ref) <= BAR


@pytest.mark.parametrize("form", ["f16", "bf16"])
@pytest.mark.parametrize("scale", [2.0 ** -60, 2.0 ** -20, 2.0 ** 20, 2.0 ** 60])
def test_split_dynamic_range(form, scale):
    """FP16 has a narrow exponent range: the F16 form scales every butterfly
    by a power of two, so the relative error does not depend on magnitude."""
    n = 4096
    rng = np.random.default_rng(3)
    x = (cplx32(rng, 4, n) * np.float32(scale)).astype(np.complex64)
    X = K.CpxPlan(n, SPLIT[form], half=True)(x)
    assert np.all(np.isfinite(X.view(np.float32)))
    assert rel(X, O.fft(x.astype(np.complex128))) <= BAR


@pytest.mark.parametrize("form", ["f16", "bf16"])
def test_split_tail_and_columns(form):
    """Batches that leave lanes of the last workgroup idle, and the column
    pass of a two-pass length, both through the matrix cores."""
    rng = np.random.default_rng(5)
    for n, batch in [(256, 1), (256, 3), (1024, 5), (1 << 16, 1), (8192, 3)]:
        x = cplx32(rng, batch, n)
        X = K.CpxPlan(n, SPLIT[form], half=True)(x)
        assert rel(X, O.fft(x.astype(np.complex128))) <= BAR, (n, batch)


@pytest.mark.parametrize("form", ["f16", "bf16"])
@pytest.mark.parametrize("n", [1009, 65537, 1000, 6561])
def test_split_flags_on_other_shapes(form, n):
    """Rader, Bluestein and mixed-radix lengths accept the flags; their
    power-of-two inner transforms use the split products where they apply."""
    rng = np.random.default_rng(n)
    x = cplx32(rng, 2, n)
    X = K.CpxPlan(n, SPLIT[form], half=True)(x)
    assert rel(X, O.fft(x.astype(np.complex128))) <= 2e-6 * max(1, math.log2(n)) / 4


@pytest.mark.parametrize("form", [None, "f16", "bf16"])
def test_refinement_reaches_output_rounding(form):
    n, batch = 1 << 16, 4
    rng = np.random.default_rng(17)
    x = tones32(rng, batch, n)
    ref = O.fft(x.astype(np.complex128))
    base = SPLIT[form] if form else 0
    y0 = K.CpxPlan(n, base, half=True)(x)
    y1 = K.CpxPlan(n, base | K.FLAG_GPU_REFINE, half=True)(x)
    e0, e1 = rel(y0, ref), rel(y1, ref)
    # rounding the exact result to FP32 is the floor
    floor = rel(ref.astype(np.complex64), ref)
    assert e1 <= REFINED_BAR and e1 <= 1.5 * floor, (e0, e1, floor)
    assert e1 < e0 / 2


@pytest.mark.parametrize("flags", [K.FLAG_INVERSE, K.FLAG_INVERSE | K.FLAG_DISABLE_NORM])
def test_refinement_inverse_and_in_place(flags):
    import torch

    n, batch = 4096, 700
    rng = np.random.default_rng(23)
    x = cplx32(rng, batch, n)
    f = O.FLAG_INVERSE | (O.FLAG_DISABLE_NORM if flags & K.FLAG_DISABLE_NORM else 0)
    ref = O.fft(x.astype(np.complex128), f)
    p = K.CpxPlan(n, K.FLAG_GPU_SPLIT_F16 | K.FLAG_GPU_REFINE | flags, half=True)
    xd = torch.from_numpy(x).cuda()
    p(xd, out=xd)  # in place: the residual still sees the original input
    assert rel(xd.cpu().numpy(), ref) <= REFINED_BAR


def test_refinement_chunks():
    """2^16 x 600: the refinement runs in chunks of 512 transforms."""
    import torch

    n, batch = 1 << 16, 600
    g = torch.Generator(device="cuda").manual_seed(29)
    x = torch.view_as_complex(torch.rand(batch, n, 2, generator=g, device="cuda") * 2 - 1)
    x = x.contiguous()
    y = torch.empty_like(x)
    K.CpxPlan(n, K.FLAG_GPU_SPLIT_BF16 | K.FLAG_GPU_REFINE, half=True)(x, out=y)
    xs = x.cpu().numpy()
    ys = y.cpu().numpy()
    for b in (0, 511, 512, 599):
        r = O.fft(xs[b].astype(np.complex128))
        assert rel(ys[b], r) <= REFINED_BAR, b
