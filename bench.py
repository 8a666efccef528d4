"""Benchmark of the batched 1D complex transform on one MI355X.

Workload (the headline configuration): N = 2^16 points, batch 4096, complex,
forward transform, inputs uniform[-1,1) (seeded) resident in HBM before the
timed region.  A step is one batched transform of the whole set through the
C ABI (kfft_eval_cpx_batch_async on the library stream).  Timing uses HIP
events recorded on that stream around exactly K steps.

Forms (all run in one invocation; --form picks the one reported as `value`):
  fp64       the double build (libkfft_amd.so), the plain form
  fp32       the float build (libkfft_amd_half.so): FP32 arithmetic on the
             vector units, FP64-generated twiddles rounded once
  f16, bf16  the precision form: radix-16 stages as split FP16 / BF16 matrix
             products with FP32 accumulation (KFFT_FLAG_GPU_SPLIT_F16 / _BF16)
  f16r, bf16r  the same plus one refinement step on the FP64 residual
             (KFFT_FLAG_GPU_REFINE)
The reported value is GFLOP/s = 5 N log2 N * batch / t.  Every FP32 form
carries its relative L2 error against the FP64 path on two input sets (uniform
noise, and unit noise with 1e-6-amplitude tones); the split forms carry their
matrix-core utilisation (dense FP16/BF16 products issued per second against
the 2.5 PFLOP/s peak).

The CPU baseline is the oracle (a C restatement of the reference's algorithm,
OpenMP over the batch, all host threads) timed on a bounded sample of the same
workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
MFMA_PEAK_FLOPS = 2.5e15  # MI355X dense FP16 / BF16 matrix peak
FORMS = ("fp64", "fp32", "f16", "bf16", "f16r", "bf16r")


def _form_flags(form: str) -> int:
    import kfft_amd as K

    f = 0
    if form.startswith("f16"):
        f |= K.FLAG_GPU_SPLIT_F16
    if form.startswith("bf16"):
        f |= K.FLAG_GPU_SPLIT_BF16
    if form.endswith("r"):
        f |= K.FLAG_GPU_REFINE
    return f


def _radix16_stages(describe: str) -> int:
    """Radix-16 stages of a power-of-two plan on the shaped passes: a pass of
    2^a points runs a // 4 radix-16 stages (one fewer when a % 4 == 1)."""
    import re

    m = re.search(r"inner=(\d+) (four-step (\d+)x(\d+)|one-pass)", describe)
    if not m:
        return 0
    lens = [int(m.group(3)), int(m.group(4))] if m.group(3) else [int(m.group(1))]
    total = 0
    for L in lens:
        if L < 16 or L > 4096 or L & (L - 1):
            return 0
        a = L.bit_length() - 1
        q, rr = divmod(a, 4)
        total += q - 1 if (rr == 1 and q > 0) else q
    return total


def _gpu_form(form: str, n: int, batch: int, steps: int, warmup: int, x64, x32, flags: int = 0):
    import torch

    import kfft_amd as K

    half = form != "fp64"
    plan = K.CpxPlan(n, flags | _form_flags(form), half=half)
    x = x32 if half else x64
    y = torch.empty_like(x)
    stream = torch.cuda.ExternalStream(K.load(half).kfft_gpu_stream())
    for _ in range(warmup):
        plan.run_async(x, y, batch)
    K.synchronize(half)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record(stream)
    for _ in range(steps):
        plan.run_async(x, y, batch)
    t1.record(stream)
    t1.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0.elapsed_time(t1) / steps
    return plan, y, ms, wall


def _measured_traffic(form: str, n: int, batch: int):
    """HBM bytes of one step from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary for this workload (profiles/traffic_*.json), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_c2c_{n}x{batch}.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d[form]["bytes_per_step"]
    except Exception:
        return None


def ragged_bench(count: int, steps: int, warmup: int, per_item: bool = True):
    """The batching form on the ragged workload: `count` transforms, lengths
    log-uniform over [16, 65536] (15% forced prime, 30% powers of two), half
    complex and half real input, FP32, one persistent launch per batch.  Also
    times the same transforms launched one by one through the batched API."""
    import torch

    import kfft_amd as K
    from kfft_amd.workloads import ragged_batch

    lengths, kinds = ragged_batch(count)
    t0 = time.perf_counter()
    plan = K.RaggedPlan(lengths, kinds, half=True)
    plan_s = time.perf_counter() - t0
    g = torch.Generator(device="cuda").manual_seed(20260105)
    x = torch.rand((plan.in_bytes + 3) // 4, generator=g, device="cuda") * 2 - 1
    y = torch.zeros((plan.out_bytes + 3) // 4, device="cuda")
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(K.load(True).kfft_gpu_stream())
    for _ in range(warmup):
        plan.run_async(x, y)
    K.synchronize(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        plan.run_async(x, y)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / steps
    if plan.error():
        raise RuntimeError("ragged batch: a dependency wait timed out")
    out_len = [(n // 2 + 1) if (k == 1 and n % 2 == 0) else n for n, k in zip(lengths, kinds)]
    alg = float(sum(n * (4 if k == 1 else 8) for n, k in zip(lengths, kinds)) +
                8 * sum(out_len))
    flops = float(sum((2.5 if k == 1 else 5.0) * n * math.log2(n) for n, k in zip(lengths, kinds)))
    res = {"transforms": int(count), "ms_per_batch": round(ms, 4),
           "gflops": round(flops / (ms * 1e-3) / 1e9, 2),
           "achieved_gbs": round(alg / (ms * 1e-3) / 1e9, 1),
           "hbm_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "alg_bytes": alg, "plan_s": round(plan_s, 2), "plan": plan.describe()}
    # shapes ranked by the committed cost model (trained on this GPU's timings)
    model_path = os.path.join(ROOT, "kfft_amd", "data", "cost_model_mi355x.npz")
    if os.path.exists(model_path):
        import numpy as np

        md = np.load(model_path)
        m = K.Mlp(int(md["nin"]), int(md["hidden"]))
        m.set_params(md["params"])
        m.use_as_cost_model()
        try:
            t0 = time.perf_counter()
            tuned = K.RaggedPlan(lengths, kinds, flags=K.FLAG_GPU_TUNED, half=True)
            tplan_s = time.perf_counter() - t0
        finally:
            K.clear_cost_model()
        for _ in range(warmup):
            tuned.run_async(x, y)
        K.synchronize(True)
        e0.record(stream)
        for _ in range(steps):
            tuned.run_async(x, y)
        e1.record(stream)
        e1.synchronize()
        tms = e0.elapsed_time(e1) / steps
        if tuned.error():
            raise RuntimeError("tuned ragged batch: a dependency wait timed out")
        changed = sum(1 for c in tuned.classes() if tuple(c[3:]) != (16, 0, 0))
        res["tuned"] = {"ms_per_batch": round(tms, 4), "speedup_vs_default": round(ms / tms, 3),
                        "plan_s": round(tplan_s, 2), "classes_not_default": changed,
                        "model": "kfft_amd/data/cost_model_mi355x.npz"}
    if per_item:
        # the method without the executor: one asynchronous batched call per transform
        plans = {}
        for n, k in set(zip(lengths.tolist(), kinds.tolist())):
            plans[(n, k)] = K.ScalarPlan(n, half=True) if k == 1 else K.CpxPlan(n, half=True)
        xs = [x[int(o) // 4:] for o in plan.in_off]
        ys = [y[int(o) // 4:] for o in plan.out_off]
        L = K.load(True)

        def one_by_one():
            for i, (n, k) in enumerate(zip(lengths.tolist(), kinds.tolist())):
                p = plans[(n, k)]
                if k == 1:
                    L.kfft_eval_scalar_batch_async(p.handle, xs[i].data_ptr(), ys[i].data_ptr(),
                                                   1, None)
                else:
                    L.kfft_eval_cpx_batch_async(p.handle, xs[i].data_ptr(), ys[i].data_ptr(), 1,
                                                None)
        one_by_one()
        K.synchronize(True)
        e0.record(stream)
        one_by_one()
        e1.record(stream)
        e1.synchronize()
        ms1 = e0.elapsed_time(e1)
        res["one_launch_per_transform_ms"] = round(ms1, 3)
        res["speedup_vs_one_launch_per_transform"] = round(ms1 / ms, 2)
    return res


def cpu_baseline(n: int, sample_batch: int):
    import numpy as np

    import oracle as O

    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, (sample_batch, n)) + 1j * rng.uniform(-1, 1, (sample_batch, n))
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except Exception:
        pass
    threads = min(threads, 16)
    p = O.Plan(n)
    p(x[:1], threads=threads)  # warm
    best = None
    for _ in range(3):
        t = time.perf_counter()
        p(x, threads=threads)
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    gflops = 5 * n * math.log2(n) * sample_batch / best / 1e9
    return {"value": round(gflops, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 C2C N={n} x {sample_batch} transforms, best of 3 "
                      f"({best:.2f} s), OpenMP over the batch"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 16)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--form", choices=list(FORMS), default="f16")
    ap.add_argument("--forms", default=",".join(FORMS),
                    help="comma-separated forms to run (the reported one is always included)")
    ap.add_argument("--only", action="store_true",
                    help="time the reported form alone (no accuracy sets, no ragged batch, no "
                         "CPU baseline): for profiler runs")
    ap.add_argument("--cpu-sample", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ragged", type=int, default=20000,
                    help="transforms of the ragged batching-form workload (0: skip)")
    ap.add_argument("--onchip", action="store_true",
                    help="plan with KFFT_FLAG_GPU_ONCHIP (fused persistent two-pass kernel)")
    ap.add_argument("--copy-ref", action="store_true",
                    help="after timing, run one plain device copy of the FP64 input (4 GiB at the "
                         "default size) so that profiler byte counters can be calibrated")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; the batch shards across GPUs "
                          "with no exchange, but multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    n, batch = args.n, args.batch
    forms = [f for f in FORMS if f in args.forms.split(",") or f in (args.form, "fp64")]
    if args.only:
        forms = [args.form]
        args.ragged, args.no_cpu = 0, True
    g = torch.Generator(device="cuda").manual_seed(20260103)
    # set A: uniform[-1,1) re/im, rounded to FP32 once so that every form sees the same values
    x32 = torch.view_as_complex(
        torch.rand(batch, n, 2, generator=g, device="cuda", dtype=torch.float32) * 2 - 1
    ).contiguous()
    x64 = x32.to(torch.complex128)

    import kfft_amd as K

    flags = K.FLAG_GPU_ONCHIP if args.onchip else 0
    res = {}
    for form in forms:
        plan, y, ms, wall = _gpu_form(form, n, batch, args.steps, args.warmup, x64, x32, flags)
        if plan.error():
            raise RuntimeError(f"{form}: the fused transform reported a failed dependency wait")
        res[form] = {"plan": plan, "ms": ms, "y": y}
        print(f"# {form}: {ms:.4f} ms/step ({plan.describe()})", file=sys.stderr, flush=True)

    def rel_err(y, ref):
        return (torch.linalg.vector_norm(y.to(torch.complex128) - ref) /
                torch.linalg.vector_norm(ref)).item()

    # accuracy on set A against the FP64 path
    y64 = res["fp64"]["y"] if "fp64" in res else None
    for form in forms:
        if form != "fp64" and y64 is not None:
            res[form]["err_uniform"] = rel_err(res[form]["y"], y64)
        del res[form]["y"]
    del y64
    # set B: unit noise plus five 1e-6-amplitude tones per transform
    if not args.only:
        t = torch.arange(n, device="cuda", dtype=torch.float64)
        freqs = torch.randint(0, n, (batch, 5), generator=g, device="cuda")
        xb = x64.clone()
        for j in range(5):
            ph = 2 * math.pi * (freqs[:, j:j + 1].to(torch.float64) * t[None, :] % n) / n
            xb += 1e-6 * torch.polar(torch.ones_like(ph), ph)
        xb32 = xb.to(torch.complex64)
        xb = xb32.to(torch.complex128)
        del t, freqs
        torch.cuda.synchronize()
        yb64 = torch.empty_like(xb)
        K.CpxPlan(n, 0)(xb, out=yb64)
        yb = torch.empty_like(xb32)
        for form in forms:
            if form != "fp64":
                K.CpxPlan(n, _form_flags(form), half=True)(xb32, out=yb)
                res[form]["err_tones"] = rel_err(yb, yb64)
        del yb, yb64, xb, xb32
    del x64

    form = args.form
    plan, ms = res[form]["plan"], res[form]["ms"]
    esz = 8 if form != "fp64" else 16
    flops = 5.0 * n * math.log2(n) * batch
    value = flops / (ms * 1e-3) / 1e9
    alg_bytes = 2.0 * n * esz * batch
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    r16 = _radix16_stages(res["fp32"]["plan"].describe()) if "fp32" in res else 0

    def form_entry(f, r):
        e = {"ms_per_step": round(r["ms"], 4),
             "gflops": round(flops / (r["ms"] * 1e-3) / 1e9, 2),
             "hbm_frac": round(2.0 * n * (8 if f != "fp64" else 16) * batch /
                               (r["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
             "speedup_vs_fp64": round(res["fp64"]["ms"] / r["ms"], 3) if "fp64" in res else None}
        if "err_uniform" in r:
            e["rel_l2_vs_fp64"] = {"uniform": r["err_uniform"], "tones_1e-6": r["err_tones"]}
        if f[:3] in ("f16", "bf1"):
            # per 64 butterflies: 2 groups x (3 products (F16) or 6 (BF16)) x 2 halves (Re, Im)
            per_wave = 12 if f.startswith("f16") else 24
            mf = r16 * (n // 16 // 64) * per_wave * (32 * 32 * 16 * 2) * batch
            passes = 2 if f.endswith("r") else 1
            e["mfma"] = {"flop_per_step": mf * passes,
                         "util": round(mf * passes / (r["ms"] * 1e-3) / MFMA_PEAK_FLOPS, 4),
                         "radix16_stages": r16}
        return e

    out = {
        "metric": "batched C2C GFLOP/s (5N*log2N/t) & %HBM peak, FP32-split vs FP64 at <=1e-6",
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if form != "fp64" else "f64",
        "data": "synthetic uniform[-1,1) seeded, resident in HBM",
        "config": {"workload": f"batched 1D C2C forward N={n} batch={batch}", "n": n,
                   "batch": batch, "form": form, "plan": plan.describe(),
                   "onchip": bool(args.onchip)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole transform (all passes of one step)",
                     "alg_bytes_per_step": alg_bytes},
        "forms": {f: form_entry(f, r) for f, r in res.items()},
    }
    if args.copy_ref:
        ref = torch.empty_like(x64)
        ref.copy_(x64)
        torch.cuda.synchronize()
        out["copy_ref_bytes"] = 2 * x64.numel() * x64.element_size()
        del ref
    if args.ragged:
        out["ragged"] = ragged_bench(args.ragged, max(3, args.steps // 2), 1)
        print(f"# ragged: {out['ragged']['ms_per_batch']} ms per batch", file=sys.stderr, flush=True)
    traffic = None if args.onchip else _measured_traffic(form, n, batch)
    if traffic is not None:
        out["roofline"]["traffic"] = traffic
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
