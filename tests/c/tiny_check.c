/*
 * Runs the tiny API (include/kfft_tiny.h: one handle for every plan kind) on
 * the GPU plans through libkfft_amd.so.  Built by __graft_entry__.build()
 * into tests/c/build/tiny_check; run by tests/test_gpu_math_shift.py.
 *
 * Each plan kind is configured through kfft_configuration, evaluated through
 * kfft_evaluation / kfft_convolution on host buffers and released through
 * kfft_finalization; results are compared with brute-force DFTs and direct
 * circular convolutions in long double.  Prints one line per case and exits
 * non-zero on any failure.
 */
#include "kfft_tiny.h"

#include <math.h>
#include <stdio.h>

#define PI_L 3.141592653589793238462643383279502884L

static int failures = 0;

static void report(const char* name, double err, double tol) {
    const int ok = err <= tol;
    printf("tiny %-28s rel err %.3e (tol %.1e) %s\n", name, err, tol, ok ? "ok" : "FAIL");
    failures += !ok;
}

/* X[k] = sum_n x[n] exp(-+2 pi i n k / N) over a strided view */
static void dft(const kfft_cpx* x, kfft_cpx* X, uint32_t n, uint32_t stride, int inverse) {
    for (uint32_t k = 0; k < n; ++k) {
        long double re = 0, im = 0;
        for (uint32_t j = 0; j < n; ++j) {
            const long double a = (inverse ? 2 : -2) * PI_L * (long double)((uint64_t)j * k % n) / n;
            const kfft_cpx v = x[(size_t)j * stride];
            re += v.r * cosl(a) - v.i * sinl(a);
            im += v.r * sinl(a) + v.i * cosl(a);
        }
        X[(size_t)k * stride].r = (kfft_scalar)re;
        X[(size_t)k * stride].i = (kfft_scalar)im;
    }
}

static double rel(const kfft_cpx* a, const kfft_cpx* b, size_t n) {
    long double num = 0, den = 0;
    for (size_t i = 0; i < n; ++i) {
        const long double dr = a[i].r - b[i].r, di = a[i].i - b[i].i;
        num += dr * dr + di * di;
        den += (long double)b[i].r * b[i].r + (long double)b[i].i * b[i].i;
    }
    return (double)sqrtl(num / (den > 0 ? den : 1));
}

static kfft_scalar val(uint32_t i, uint32_t salt) {
    return (kfft_scalar)(sin(0.37 * i + 0.11 * salt) + 0.25 * cos(1.3 * i * (salt + 1)));
}

int main(void) {
    enum { N = 60, X = 6, Y = 5 };
    static kfft_cpx in[N], out[N], ref[N], in2[N];
    static kfft_scalar r_in[N], r_out[N];

    /* 1D complex forward */
    for (uint32_t i = 0; i < N; ++i) {
        in[i].r = val(i, 1);
        in[i].i = val(i, 2);
    }
    kfft_args_norm a1 = {N};
    kfft_plan p = kfft_configuration(KFFT_PLAN_COMPLEX, KFFT_FLAG_NORMAL, &a1);
    if (!p || kfft_evaluation(p, in, out) != KFFT_RET_SUCCESS) {
        printf("tiny complex: configuration or evaluation failed\n");
        return 2;
    }
    dft(in, ref, N, 1, 0);
    report("complex 60 forward", rel(out, ref, N), 1e-13);
    kfft_finalization(p);

    /* 1D real forward (N/2+1 bins) and inverse (2x for even N, kfft_scalar.c:218-242) */
    for (uint32_t i = 0; i < N; ++i) {
        r_in[i] = val(i, 3);
        in2[i].r = r_in[i];
        in2[i].i = 0;
    }
    p = kfft_configuration(KFFT_PLAN_SCALAR, KFFT_FLAG_NORMAL, &a1);
    if (!p || kfft_evaluation(p, r_in, out) != KFFT_RET_SUCCESS) {
        printf("tiny scalar: configuration or evaluation failed\n");
        return 2;
    }
    dft(in2, ref, N, 1, 0);
    report("scalar 60 forward", rel(out, ref, N / 2 + 1), 1e-13);
    kfft_finalization(p);
    p = kfft_configuration(KFFT_PLAN_SCALAR, KFFT_FLAG_INVERSE, &a1);
    if (!p || kfft_evaluation(p, out, r_out) != KFFT_RET_SUCCESS) {
        printf("tiny scalar inverse: configuration or evaluation failed\n");
        return 2;
    }
    {
        long double num = 0, den = 0;
        for (uint32_t i = 0; i < N; ++i) {
            const long double d = r_out[i] - 2.0L * r_in[i];
            num += d * d;
            den += 4.0L * r_in[i] * r_in[i];
        }
        report("scalar 60 inverse (2x)", (double)sqrtl(num / den), 1e-13);
    }
    kfft_finalization(p);

    /* 2D complex X x Y, row-major with x fastest */
    for (uint32_t i = 0; i < X * Y; ++i) {
        in[i].r = val(i, 4);
        in[i].i = val(i, 5);
    }
    kfft_args_2d a2 = {X, Y};
    p = kfft_configuration(KFFT_PLAN_COMPLEX_2D, KFFT_FLAG_NORMAL, &a2);
    if (!p || kfft_evaluation(p, in, out) != KFFT_RET_SUCCESS) {
        printf("tiny 2d: configuration or evaluation failed\n");
        return 2;
    }
    {
        static kfft_cpx t[X * Y];
        for (uint32_t r = 0; r < Y; ++r)
            dft(in + r * X, t + r * X, X, 1, 0);
        for (uint32_t c = 0; c < X; ++c)
            dft(t + c, ref + c, Y, X, 0);
    }
    report("complex 2d 6x5 forward", rel(out, ref, X * Y), 1e-13);
    kfft_finalization(p);

    /* 1D complex circular convolution (normalised, ext/conv/kfft_cpx_conv.c:24-104) */
    for (uint32_t i = 0; i < N; ++i) {
        in[i].r = val(i, 6);
        in[i].i = val(i, 7);
        in2[i].r = val(i, 8);
        in2[i].i = val(i, 9);
    }
    p = kfft_configuration(KFFT_PLAN_COMPLEX_CONV, KFFT_FLAG_NORMAL, &a1);
    if (!p || kfft_convolution(p, in, in2, out) != KFFT_RET_SUCCESS) {
        printf("tiny conv: configuration or evaluation failed\n");
        return 2;
    }
    for (uint32_t k = 0; k < N; ++k) {
        long double re = 0, im = 0;
        for (uint32_t j = 0; j < N; ++j) {
            const kfft_cpx u = in[j], v = in2[(k + N - j) % N];
            re += (long double)u.r * v.r - (long double)u.i * v.i;
            im += (long double)u.r * v.i + (long double)u.i * v.r;
        }
        ref[k].r = (kfft_scalar)re;
        ref[k].i = (kfft_scalar)im;
    }
    report("complex conv 60", rel(out, ref, N), 1e-13);
    kfft_finalization(p);

    printf("tiny api: %s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
