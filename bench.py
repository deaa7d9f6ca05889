"""Benchmark: Zernike GLQ forward + backward transforms per second at order 128.

Workload: 64 independent Zernike expansions at order N = 128 (grids
255 x 129 x 136 doubles each), conventions (unnormed, geo, none) as zest's own
benchmarks use (GLQTransformerGeo). Coefficients are N(0,1)/(1+n) from the
portable SplitMix64 stream, seeds 0..63. One step = one backward transform of
all 64 expansions followed by one forward transform of the resulting grids,
with every input already resident in HBM. value = instances * steps / time.

The dominant kernel's roofline is taken from HIP events recorded around each
kernel launch on its own stream during the timed region (zgl_profile_enable).
The CPU baseline times the upstream zest library (compiled from its sources
into oracle/_ref) on the host cores over a bounded sample, or the C oracle
when that library is absent.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def stage_work(order: int):
    """Algorithmic bytes / flops of one instance per stage (one direction)."""
    nphi, nth, nr = 2 * order - 1, order + 1, (order + 1 + 7) // 8 * 8
    F = nphi // 2 + 1
    ncoef = (order + 1) * (order + 3) * (2 * order + 1) // 24
    fft_bytes = nphi * nth * nr * 8 + F * nth * nr * 16
    legendre_flops = 2 * (order * (order + 1) // 2) * nth * nr
    radial_flops = 4 * ncoef * nr
    return {
        "fft_forward": ("hbm", fft_bytes),
        "fft_backward": ("hbm", fft_bytes),
        "legendre_forward": ("mfma", legendre_flops),
        "legendre_backward": ("mfma", legendre_flops),
        "radial_forward": ("mfma", radial_flops),
        "radial_backward": ("mfma", radial_flops),
    }


def read_stage_stats(zest_amd, plan):
    class Stats(ctypes.Structure):
        _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_longlong),
                    ("instances", ctypes.c_longlong), ("total_ms", ctypes.c_double)]
    lib = zest_amd.LIB
    lib.zgl_profile_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats), ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int)]
    arr = (Stats * 16)()
    n = ctypes.c_int()
    zest_amd._check(lib.zgl_profile_stats(plan._h, arr, 16, ctypes.byref(n)))
    return {arr[i].name.decode(): (arr[i].launches, arr[i].instances, arr[i].total_ms)
            for i in range(n.value)}


def cpu_baseline(order: int, seconds: float = 12.0):
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
        "value": total / elapsed,
        "unit": "transforms/s",
        "cores": threads,
        "kind": kind,
        "sample": (f"{total} Zernike backward+forward pairs at order {order} in {elapsed:.1f} s, "
                   f"{threads} threads, one transformer each; "
                   + ("zest compiled from source (-O3 -march=x86-64-v3)" if kind == "reference"
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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the transforms are "
              "independent per instance, so several GPUs would each run this same workload on "
              "their own share; no multi-GPU figure is produced.", file=sys.stderr)
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
    plan = zest_amd.Plan(zest_amd.ZERNIKE, N, zernike_norm=1, sh_norm=0, sh_phase=-1)
    stream = torch.cuda.current_stream()

    def step():
        plan.backward(dc, dg, N, batch=B, stream=stream)
        plan.forward(dg, dco, N, batch=B, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # accuracy of the round trip (forward o backward = identity on band-limited input)
    err = float(((dco - dc).abs().amax(dim=(1, 2)) / dc.abs().amax(dim=(1, 2))).max())

    zest_amd.LIB.zgl_profile_enable(plan._h, 1)
    read_stage_stats(zest_amd, plan)  # discard anything recorded before
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stats = read_stage_stats(zest_amd, plan)
    zest_amd.LIB.zgl_profile_enable(plan._h, 0)

    value = B * args.steps / elapsed
    work = stage_work(N)
    stages = {}
    for name, (launches, inst, ms) in stats.items():
        bound, per_inst = work[name]
        per_launch = per_inst * inst / launches
        avg_s = ms / launches / 1e3
        if bound == "hbm":
            achieved, peak, unit = per_launch / avg_s / 1e9, HBM_PEAK_GBS, "GB/s"
        else:
            achieved, peak, unit = per_launch / avg_s / 1e12, FP64_PEAK_TFLOPS, "TFLOP/s"
        stages[name] = {"ms_per_launch": ms / launches, "launches": launches,
                        "instances_per_launch": inst / launches, "bound": bound,
                        "achieved": achieved, "peak": peak, "unit": unit,
                        "frac": achieved / peak, "share": 0.0}
    total_ms = sum(v["ms_per_launch"] * v["launches"] for v in stages.values())
    for v in stages.values():
        v["share"] = v["ms_per_launch"] * v["launches"] / total_ms
    dom_name = max(stages, key=lambda k: stages[k]["share"])
    dom = stages[dom_name]
    roofline = {"bound": dom["bound"], "achieved": round(dom["achieved"], 3), "peak": dom["peak"],
                "unit": dom["unit"], "frac": round(dom["frac"], 4), "traffic": None,
                "kernel": dom_name}

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
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "zernike_glq_backward_forward", "order": N, "instances": B,
                   "grid": list(zest_amd.zt_shape(N)), "conventions": "unnormed,geo,none",
                   "precision": "fp64"},
        "max_rel_err_roundtrip": err,
        "roofline": roofline,
        "stages": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                   for k, v in stages.items()},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
