"""Benchmark: Zernike GLQ forward + backward transforms per second at order 128,
mixed precision (FP32 FFT + FP32-MFMA contractions) against FP64.

Workload: 64 independent Zernike expansions at order N = 128 (grids
255 x 129 x 136 doubles each), conventions (unnormed, geo, none) as zest's own
benchmarks use (GLQTransformerGeo). Coefficients are N(0,1)/(1+n) from the
portable SplitMix64 stream, seeds 0..63. One step = one backward transform of
all 64 expansions followed by one forward transform of the resulting grids,
with every input already resident in HBM. value = instances * steps / time of
the mixed form; the FP64 form is timed the same way beside it.

Two mixed forms are timed: with the grids stored in FP64 (zgl_forward /
zgl_backward of an FP32 plan) and with the grids stored in FP32
(zgl_forward_grid32 / zgl_backward_grid32); tables and coefficients are FP64
in both. Accuracy gate (err <= 1e-6): max over instances of
max|c - c64| / max|c64|, where c64 is the FP64 forward of the FP64 backward
grids and c the mixed forward of the same grids as that form stores them (so
the FP32-grid form's error includes rounding the grids to FP32). The plain
FP32 forward is used when it meets the gate, else the forward with one
FP64-residual refinement step. The headline is the faster form that meets
the gate; the other is reported beside it.

The dominant kernel's roofline comes from HIP events recorded around each
kernel launch on its own stream during the timed region. The CPU baseline
times the upstream zest library (compiled from its sources into oracle/_ref)
on the host cores over a bounded sample, or the C oracle when that library is
absent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}


def stage_work(order: int, real_bytes: int, grid_bytes: int = 8):
    """Algorithmic bytes / flops of one instance per stage (one direction).
    The grid holds `grid_bytes` reals per point; the spectrum holds complex
    values of the compute type."""
    nphi, nth, nr = 2 * order - 1, order + 1, (order + 1 + 7) // 8 * 8
    F = nphi // 2 + 1
    ncoef = (order + 1) * (order + 3) * (2 * order + 1) // 24
    fft_bytes = nphi * nth * nr * grid_bytes + F * nth * nr * 2 * real_bytes
    legendre_flops = 2 * (order * (order + 1) // 2) * nth * nr
    radial_flops = 4 * ncoef * nr
    return {
        "fft_forward": ("hbm", fft_bytes), "fft_backward": ("hbm", fft_bytes),
        "legendre_forward": ("mfma", legendre_flops), "legendre_backward": ("mfma", legendre_flops),
        "radial_forward": ("mfma", radial_flops), "radial_backward": ("mfma", radial_flops),
    }


def stage_table(stats, order, dtype, grid_bytes=8):
    work = stage_work(order, 8 if dtype == "f64" else 4, grid_bytes)
    stages = {}
    for name, (launches, inst, ms) in stats.items():
        if name not in work or launches == 0:
            continue
        bound, per_inst = work[name]
        per_launch = per_inst * inst / launches
        avg_s = ms / launches / 1e3
        if bound == "hbm":
            achieved, peak, unit = per_launch / avg_s / 1e9, HBM_PEAK_GBS, "GB/s"
        else:
            achieved, peak, unit = per_launch / avg_s / 1e12, PEAK_TFLOPS[dtype], "TFLOP/s"
        stages[name] = {"ms_per_launch": round(ms / launches, 5), "launches": launches,
                        "instances_per_launch": round(inst / launches, 3), "bound": bound,
                        "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                        "frac": round(achieved / peak, 4), "total_ms": ms}
    total = sum(v["total_ms"] for v in stages.values()) or 1.0
    for v in stages.values():
        v["share"] = round(v.pop("total_ms") / total, 4)
    return stages


def cpu_baseline(order: int, seconds: float):
    """zest (or the oracle) fwd+bwd pairs/s on the host, one transformer per thread."""
    import numpy as np
    import oracle
    from zest_amd import synth
    threads = max(1, min(16, os.cpu_count() or 1))
    kind = "reference" if oracle.REF is not None else "port"
    c = synth.zernike_coeffs(order, 0)
    counts = [0] * threads
    deadline = [0.0]

    def work(t):
        if kind == "reference":
            z = oracle.ZestZernike(order, 1, 0, -1)
            g = np.zeros(oracle.zt_shape(order))
            co = np.zeros_like(c)
            while time.perf_counter() < deadline[0]:
                z.backward(c, out=g)
                z.forward(g, out=co)
                counts[t] += 1
        else:
            while time.perf_counter() < deadline[0]:
                g = oracle.zt_backward(c, order)
                oracle.zt_forward(g, order)
                counts[t] += 1

    # ctypes releases the GIL during each call, so the threads run in parallel.
    t0 = time.perf_counter()
    deadline[0] = t0 + seconds
    ths = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    elapsed = time.perf_counter() - t0
    total = sum(counts)
    return {
        "value": round(total / elapsed, 3), "unit": "transforms/s", "cores": threads, "kind": kind,
        "sample": (f"{total} Zernike backward+forward pairs at order {order} in {elapsed:.1f} s, "
                   f"{threads} threads, one transformer each, FP64; "
                   + ("zest compiled from its sources (-O3 -march=x86-64-v3)" if kind == "reference"
                      else "C restatement of zest")),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--order", type=int, default=128)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--only", choices=["fp64", "mixed", "mixed_grid32"], default=None,
                    help="time one form only (profiling runs)")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the accuracy phase (profiling runs): the mixed forms then use the "
                         "plain FP32 forward and no accuracy is reported")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the transforms are "
              "independent per instance, so each GPU would run this same workload on its own "
              "share of instances; no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import zest_amd
    from zest_amd import synth

    dev = torch.device("cuda:0")
    N, B = args.order, args.batch
    coeffs = np.stack([synth.zernike_coeffs(N, s) for s in range(B)])
    dc = torch.from_numpy(coeffs).to(dev)
    dg = torch.empty((B,) + zest_amd.zt_shape(N), dtype=torch.float64, device=dev)
    dco = torch.zeros_like(dc)
    stream = torch.cuda.current_stream()
    conv = dict(zernike_norm=1, sh_norm=0, sh_phase=-1)
    p64 = zest_amd.Plan(zest_amd.ZERNIKE, N, precision=zest_amd.FP64, **conv)
    p32 = zest_amd.Plan(zest_amd.ZERNIKE, N, precision=zest_amd.FP32, **conv)
    p32r = zest_amd.Plan(zest_amd.ZERNIKE, N, precision=zest_amd.FP32_REFINE, **conv)

    def relerr(a, ref):
        dims = tuple(range(1, ref.dim()))
        return float(((a.double() - ref).abs().amax(dim=dims) / ref.abs().amax(dim=dims)).max())

    # accuracy of the mixed forms against FP64 on the FP64 backward grids
    if not args.no_accuracy:
        p64.backward(dc, dg, N, batch=B, stream=stream)
        c64 = torch.zeros_like(dc)
        p64.forward(dg, c64, N, batch=B, stream=stream)
    acc = {"fp64_roundtrip_max_rel_err": None}
    forms = {}
    if args.no_accuracy:
        forms = {"mixed": (torch.float64, p32, "fp32", True),
                 "mixed_grid32": (torch.float32, p32, "fp32", True)}
    for form, gdtype in (() if args.no_accuracy else
                         (("mixed", torch.float64), ("mixed_grid32", torch.float32))):
        gin = dg.to(gdtype)
        c32 = torch.zeros_like(dc)
        p32.forward(gin, c32, N, batch=B, stream=stream)
        c32r = torch.zeros_like(dc)
        p32r.forward(gin, c32r, N, batch=B, stream=stream)
        gout = torch.empty_like(gin)
        p32.backward(dc, gout, N, batch=B, stream=stream)
        torch.cuda.synchronize()
        e32, e32r, ebwd = relerr(c32, c64), relerr(c32r, c64), relerr(gout, dg)
        acc[form] = {"fwd_max_rel_err_vs_fp64": e32, "refined_fwd_max_rel_err_vs_fp64": e32r,
                     "bwd_max_rel_err_vs_fp64": ebwd}
        fwd_plan, fwd_name = (p32, "fp32") if e32 <= 1e-6 else (p32r, "fp32+refine")
        gate = min(e32, e32r) <= 1e-6
        forms[form] = (gdtype, fwd_plan, fwd_name, gate)
        del gin, c32, c32r, gout
    if not args.no_accuracy:
        acc["fp64_roundtrip_max_rel_err"] = relerr(c64, dc)
    else:
        acc = None
    dg32 = torch.empty(dg.shape, dtype=torch.float32, device=dev)

    def timed(plan_bwd, plan_fwd, grid):
        def step():
            plan_bwd.backward(dc, grid, N, batch=B, stream=stream)
            plan_fwd.forward(grid, dco, N, batch=B, stream=stream)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        for p in (plan_bwd, plan_fwd):
            p.profile(True)
            p.stage_stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        stats = {}
        for p in (plan_bwd, plan_fwd) if plan_bwd is not plan_fwd else (plan_bwd,):
            for k, v in p.stage_stats().items():
                a = stats.get(k, (0, 0, 0.0))
                stats[k] = (a[0] + v[0], a[1] + v[1], a[2] + v[2])
            p.profile(False)
        return elapsed, stats

    results = {}
    if args.only in (None, "fp64"):
        results["fp64"] = timed(p64, p64, dg)
    for form, (gdtype, fwd_plan, _, gate) in forms.items():
        if gate and args.only in (None, form):
            results[form] = timed(p32, fwd_plan, dg if gdtype == torch.float64 else dg32)

    mixed = [f for f in ("mixed", "mixed_grid32") if f in results]
    head = max(mixed, key=lambda f: -results[f][0]) if mixed else "fp64"
    elapsed, stats = results[head]
    dtype = "f32" if head != "fp64" else "f64"
    gbytes = 4 if head == "mixed_grid32" else 8
    value = B * args.steps / elapsed
    stages = stage_table(stats, N, dtype, gbytes)
    dom_name = max(stages, key=lambda k: stages[k]["share"])
    dom = stages[dom_name]
    out = {
        "metric": "Zernike fwd+bwd transforms/s @order128: FP32-MFMA vs FP64, err<=1e-6",
        "value": round(value, 2),
        "unit": "transforms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": "zernike_glq_backward_forward", "order": N, "instances": B,
                   "grid": list(zest_amd.zt_shape(N)), "conventions": "unnormed,geo,none",
                   "form": head,
                   "grid_storage": "f32" if head == "mixed_grid32" else "f64",
                   "forward": forms[head][2] if head in forms else "fp64"},
        "roofline": {"bound": dom["bound"], "achieved": dom["achieved"], "peak": dom["peak"],
                     "unit": dom["unit"], "frac": dom["frac"], "traffic": None,
                     "kernel": dom_name},
        "accuracy": acc,
        "stages": stages,
    }
    for form in mixed:
        if form == head:
            continue
        e, st = results[form]
        out[form] = {"value": round(B * args.steps / e, 2),
                     "ms_per_step": round(e / args.steps * 1e3, 4), "forward": forms[form][2],
                     "stages": stage_table(st, N, "f32", 4 if form == "mixed_grid32" else 8)}
    if "fp64" in results:
        e64, s64 = results["fp64"]
        out["fp64_value"] = round(B * args.steps / e64, 2)
        out["fp64_ms_per_step"] = round(e64 / args.steps * 1e3, 4)
        out["fp64_stages"] = stage_table(s64, N, "f64")
        if head != "fp64":
            out["mixed_vs_fp64"] = round(value / out["fp64_value"], 3)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
