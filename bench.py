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
in both. Accuracy gate (err <= 1e-6), normwise: max over instances of
max|x - x64| / max|x64|. A form qualifies when (1) its forward of the FP64
backward grids (as the form stores them) and (2) its backward are within the
gate against FP64, and (3) the output of its timed step itself -- the mixed
backward into the form's grid, then the mixed forward of that grid -- is
within the gate against the FP64 step's output. The plain FP32 forward is
used when it passes, else the forward with one FP64-residual refinement step.
The headline is the faster qualifying form; the other is reported beside it.
Elementwise errors with a floor, max |x - x64| / max(|x64|, 1e-3 max|x64|),
are reported beside the normwise ones.

value comes from a run of the steps without instrumentation; the same steps
then run again with HIP events recorded around each kernel launch on its own
stream, and the dominant kernel's roofline comes from those; every stage also
reports its binding roof, max(bytes / HBM peak, flops / MFMA peak), and the
step the sum of those roofs against its measured time. The CPU baseline
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
# HBM bytes per launch of a stage kernel, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (in separate runs, scaled by what the counters read for a 512 MiB device
# copy: x2.0 and x1.0) on the workload named by the key: (stage, order, instances, form)
# -> (fetch bytes, write bytes). Reported as roofline.traffic only for that workload;
# the commands and the full table are in profiles/zernike_n128_b64_kernels.md.
MEASURED_TRAFFIC = {
    ("fft_forward", 128, 64, "mixed_grid32"): (1192e6, 1160e6),
    ("fft_backward", 128, 64, "mixed_grid32"): (1168e6, 1155e6),
    ("legendre_forward", 128, 64, "mixed_grid32"): (1303e6, 589e6),
    ("legendre_backward", 128, 64, "mixed_grid32"): (688e6, 1183e6),
    ("radial_forward", 128, 64, "mixed_grid32"): (596e6, 196e6),
    ("radial_backward", 128, 64, "mixed_grid32"): (225e6, 589e6),
    ("fft_forward", 128, 64, "fp64"): (2326e6, 2317e6),
    ("fft_backward", 128, 64, "fp64"): (2318e6, 2308e6),
    ("legendre_forward", 128, 64, "fp64"): (2511e6, 1150e6),
    ("legendre_backward", 128, 64, "fp64"): (1298e6, 2300e6),
    ("radial_forward", 128, 64, "fp64"): (1187e6, 199e6),
    ("radial_backward", 128, 64, "fp64"): (284e6, 1152e6),
}
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}


def stage_work(order: int, real_bytes: int, grid_bytes: int = 8):
    """Algorithmic HBM bytes and flops of one instance per stage (one
    direction). The grid holds `grid_bytes` reals per point; the spectrum and
    flm hold complex values of the compute type; coefficients are FP64 pairs.
    Flops count the parity-split contractions (a real factor times a complex
    value: 4 flops per term) and nothing for the FFT (its arithmetic is far
    below the machine balance)."""
    nphi, nth, nr = 2 * order - 1, order + 1, (order + 1 + 7) // 8 * 8
    F = nphi // 2 + 1
    ncoef = (order + 1) * (order + 3) * (2 * order + 1) // 24
    ntri = order * (order + 1) // 2
    spec = F * nth * nr * 2 * real_bytes
    flm = ntri * nr * 2 * real_bytes
    fft = (nphi * nth * nr * grid_bytes + spec, 0)
    leg = (spec + flm, 2 * ntri * nth * nr)
    rad = (flm + 16 * ncoef, 4 * ncoef * nr)
    fused = (spec + 16 * ncoef, leg[1] + rad[1])
    return {"fft_forward": fft, "fft_backward": fft, "legendre_forward": leg, "legendre_backward": leg,
            "radial_forward": rad, "radial_backward": rad,
            "legendre_radial_forward": fused, "radial_legendre_backward": fused}


def _traffic(stage, order, instances, form):
    """Measured HBM bytes per launch (fetch + write) for the workload, or None."""
    t = MEASURED_TRAFFIC.get((stage, order, instances, form))
    return None if t is None else round(t[0] + t[1])


def stage_table(stats, order, dtype, grid_bytes=8):
    """Per stage: the achieved HBM bandwidth and MFMA rate against their peaks,
    the binding roof max(bytes / HBM peak, flops / MFMA peak) and the fraction
    of it reached (roof time / measured time)."""
    work = stage_work(order, 8 if dtype == "f64" else 4, grid_bytes)
    stages = {}
    for name, (launches, inst, ms) in stats.items():
        if name not in work or launches == 0:
            continue
        nbytes, flops = work[name]
        avg_s = ms / launches / 1e3
        b_launch, f_launch = nbytes * inst / launches, flops * inst / launches
        t_hbm = b_launch / (HBM_PEAK_GBS * 1e9)
        t_mfma = f_launch / (PEAK_TFLOPS[dtype] * 1e12)
        bound = "hbm" if t_hbm >= t_mfma else "mfma"
        if bound == "hbm":
            achieved, peak, unit = b_launch / avg_s / 1e9, HBM_PEAK_GBS, "GB/s"
        else:
            achieved, peak, unit = f_launch / avg_s / 1e12, PEAK_TFLOPS[dtype], "TFLOP/s"
        stages[name] = {"ms_per_launch": round(ms / launches, 5), "launches": launches,
                        "instances_per_launch": round(inst / launches, 3), "bound": bound,
                        "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                        "frac": round(achieved / peak, 4),
                        "hbm_gbs": round(b_launch / avg_s / 1e9, 1),
                        "mfma_tflops": round(f_launch / avg_s / 1e12, 2),
                        "roof_ms": round(max(t_hbm, t_mfma) * 1e3, 5), "total_ms": ms}
    total = sum(v["total_ms"] for v in stages.values()) or 1.0
    for v in stages.values():
        v["share"] = round(v.pop("total_ms") / total, 4)
    return stages


def pipeline_roofline(stages, steps, ms_per_step):
    """The summed roof time of every stage launch in the timed region (the
    sum over stages of max(bytes / HBM peak, flops / MFMA peak)) per step,
    against the measured step time."""
    if not stages or not ms_per_step:
        return None
    roof = sum(v["roof_ms"] * v["launches"] for v in stages.values()) / steps
    return {"roof_ms_per_step": round(roof, 4), "ms_per_step": ms_per_step,
            "frac": round(roof / ms_per_step, 4)}


def cpu_info():
    """The host CPU as this process sees it: model name, logical CPUs, the
    CPUs this process may run on (the box's share), and the ISA of the zest
    build that the baseline loads."""
    import oracle
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "zest_build": oracle.REF_ISA}


def cpu_baseline(order: int, seconds: float):
    """zest (or the oracle) fwd+bwd pairs/s on the host, one transformer per
    thread, on every CPU the process may use and on one: after a warm-up
    second, the pairs completed in each one-second epoch (nanobench's
    minEpochTime(1s), benchmarks/benchmark_zernike_glq_forward_transform.cpp:52-54),
    the median epoch reported."""
    import numpy as np
    import oracle
    from zest_amd import synth
    info = cpu_info()
    kind = "reference" if oracle.REF is not None else "port"
    c = synth.zernike_coeffs(order, 0)

    def run(threads, epochs):
        stamps = [[] for _ in range(threads)]
        t0 = time.perf_counter()
        end = t0 + 1.0 + epochs

        def work(t):
            if kind == "reference":
                z = oracle.ZestZernike(order, 1, 0, -1)
                g = np.zeros(oracle.zt_shape(order))
                co = np.zeros_like(c)
                while time.perf_counter() < end:
                    z.backward(c, out=g)
                    z.forward(g, out=co)
                    stamps[t].append(time.perf_counter())
            else:
                while time.perf_counter() < end:
                    g = oracle.zt_backward(c, order)
                    oracle.zt_forward(g, order)
                    stamps[t].append(time.perf_counter())

        # ctypes releases the GIL during each call, so the threads run in parallel
        ths = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        allst = np.sort(np.concatenate([np.asarray(x) for x in stamps if x] or [np.zeros(0)]))
        rates = [int(((allst >= t0 + 1.0 + k) & (allst < t0 + 2.0 + k)).sum()) for k in range(epochs)]
        return float(np.median(rates)), len(allst), rates

    threads = max(1, info["affinity"])
    epochs_all = max(3, int(seconds * 0.6))
    epochs_one = max(3, int(seconds * 0.3))
    v_all, n_all, r_all = run(threads, epochs_all)
    v_one, n_one, r_one = run(1, epochs_one)
    build = (f"zest compiled from its sources (-O3 -march={oracle.REF_ISA})" if kind == "reference"
             else "C restatement of zest")
    return {
        "value": round(v_all, 3), "unit": "transforms/s", "cores": threads, "kind": kind,
        "single_core_value": round(v_one, 3),
        "cpu": info,
        "sample": (f"Zernike backward+forward pairs at order {order}, FP64, one transformer per "
                   f"thread: median of {epochs_all} one-second epochs on {threads} threads "
                   f"(pairs per epoch {r_all}; {n_all} in all, warm-up second included) and of "
                   f"{epochs_one} on 1 thread ({r_one}); {build}"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--order", type=int, default=128)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--instances", type=int, default=0,
                    help="sh512: instances (default 4096; profiling runs take fewer)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--only", choices=["fp64", "mixed", "mixed_grid32"], default=None,
                    help="time one form only (profiling runs)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tile shapes (no timing of candidates on first use)")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the accuracy phase (profiling runs): the mixed forms then use the "
                         "plain FP32 forward and no accuracy is reported")
    ap.add_argument("--config", choices=["zernike128", "ragged", "sh512", "grad", "eval", "rotate", "spectra"],
                    default="zernike128",
                    help="zernike128: the headline (64 instances at order 128); ragged: 10,000 "
                         "instances of orders U{8..96}; sh512: 4096 SH instances at L=512, "
                         "precision report; grad: 256 least-squares gradients at order 64; "
                         "eval: grid evaluation of 64 order-128 Zernike expansions on "
                         "256 x 128 x 128 points (and 64 SH expansions, L=512, on 1024 x 512); "
                         "rotate: ZYZ rotations of 64 order-128 Zernike expansions (and 64 SH "
                         "expansions at L=512); spectra: power spectra of 512 order-128 Zernike "
                         "expansions (and 512 SH expansions at L=512)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the transforms are "
              "independent per instance, so each GPU would run this same workload on its own "
              "share of instances; no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)
    {"zernike128": run_headline, "ragged": run_ragged, "sh512": run_sh512,
     "grad": run_grad, "eval": run_eval, "rotate": run_rotate, "spectra": run_spectra}[args.config](args)


def run_headline(args):
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
    if not args.no_autotune:
        for p in (p64, p32, p32r):
            p.autotune(True)

    def relerr(a, ref):
        """max over instances of max|a - ref| / max|ref| (normwise)"""
        dims = tuple(range(1, ref.dim()))
        return float(((a.double() - ref).abs().amax(dim=dims) / ref.abs().amax(dim=dims)).max())

    def elemerr(a, ref, floor=1e-3):
        """max over entries of |a - ref| / max(|ref|, floor * max|ref| of the instance)"""
        dims = tuple(range(1, ref.dim()))
        scale = torch.maximum(ref.abs(), floor * ref.abs().amax(dim=dims, keepdim=True))
        return float(((a.double() - ref).abs() / scale).max())

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
                     "bwd_max_rel_err_vs_fp64": ebwd,
                     "fwd_elementwise_err_vs_fp64": elemerr(c32, c64),
                     "bwd_elementwise_err_vs_fp64": elemerr(gout, dg)}
        fwd_plan, fwd_name = (p32, "fp32") if e32 <= 1e-6 else (p32r, "fp32+refine")
        # the timed step's own output is gated after the timed loop as well
        gate = min(e32, e32r) <= 1e-6 and ebwd <= 1e-6
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
        # the throughput run carries no instrumentation: HIP events recorded
        # between the stage kernels add about 30 us per step
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        # then the same steps again with events around every stage launch
        # (on the launch stream): the per-stage durations and rooflines
        for p in (plan_bwd, plan_fwd):
            p.profile(True)
            p.stage_stats()
        torch.cuda.synchronize()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        stats = {}
        for p in (plan_bwd, plan_fwd) if plan_bwd is not plan_fwd else (plan_bwd,):
            for k, v in p.stage_stats().items():
                a = stats.get(k, (0, 0, 0.0))
                stats[k] = (a[0] + v[0], a[1] + v[1], a[2] + v[2])
            p.profile(False)
        # the error of what the timed step produced (backward into this form's
        # grid, forward of that grid) against the FP64 step's output
        step_err = None if args.no_accuracy else (relerr(dco, c64), elemerr(dco, c64))
        return elapsed, stats, step_err

    results = {}
    if args.only in (None, "fp64"):
        results["fp64"] = timed(p64, p64, dg)
    for form, (gdtype, fwd_plan, fwd_name, gate) in forms.items():
        if gate and args.only in (None, form):
            r = timed(p32, fwd_plan, dg if gdtype == torch.float64 else dg32)
            if r[2] is not None and r[2][0] > 1e-6 and fwd_plan is p32:
                # the plain FP32 step misses the gate: time the refined forward instead
                forms[form] = (gdtype, p32r, "fp32+refine", gate)
                r = timed(p32, p32r, dg if gdtype == torch.float64 else dg32)
            if r[2] is not None:
                acc[form]["timed_step_max_rel_err_vs_fp64"] = r[2][0]
                acc[form]["timed_step_elementwise_err_vs_fp64"] = r[2][1]
            if r[2] is None or r[2][0] <= 1e-6:
                results[form] = r

    mixed = [f for f in ("mixed", "mixed_grid32") if f in results]
    head = max(mixed, key=lambda f: -results[f][0]) if mixed else "fp64"
    elapsed, stats, _ = results[head]
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
                     "unit": dom["unit"], "frac": dom["frac"],
                     "traffic": _traffic(dom_name, N, B, head),
                     "kernel": dom_name},
        "accuracy": acc,
        "pipeline_roofline": pipeline_roofline(stages, args.steps, round(elapsed / args.steps * 1e3, 4)),
        "stages": stages,
        "tuning": {"fp64": [list(t) for t in p64.tuning()], "fp32": [list(t) for t in p32.tuning()]},
    }
    for form in mixed:
        if form == head:
            continue
        e, st, _ = results[form]
        ms = round(e / args.steps * 1e3, 4)
        stt = stage_table(st, N, "f32", 4 if form == "mixed_grid32" else 8)
        out[form] = {"value": round(B * args.steps / e, 2), "ms_per_step": ms, "forward": forms[form][2],
                     "pipeline_roofline": pipeline_roofline(stt, args.steps, ms), "stages": stt}
    if "fp64" in results:
        e64, s64, _ = results["fp64"]
        out["fp64_value"] = round(B * args.steps / e64, 2)
        out["fp64_ms_per_step"] = round(e64 / args.steps * 1e3, 4)
        out["fp64_stages"] = stage_table(s64, N, "f64")
        out["fp64_pipeline_roofline"] = pipeline_roofline(out["fp64_stages"], args.steps,
                                                          out["fp64_ms_per_step"])
        if head != "fp64":
            out["mixed_vs_fp64"] = round(value / out["fp64_value"], 3)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def _timed_steps(step, steps, warmup):
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def run_ragged(args):
    """10,000 Zernike instances of orders N ~ U{8..96} (seed 2), coefficients
    N(0,1)/(1+n): one ragged backward + forward per step through
    zgl_backward_ragged / zgl_forward_ragged (the default lanes pipeline: an
    order-sorted queue of per-order pipelines over four internal streams; HIP
    graph captured on the first call, replayed after), against the same work
    as per-order batched zgl_backward / zgl_forward calls on one stream and as
    the persistent pipeline (ZGL_RAGGED=persistent: the queue cut into
    scratch-sized waves, one persistent launch per stage and wave over all
    orders, items dealt to the resident workgroups longest first)."""
    import numpy as np
    import torch
    import zest_amd
    from zest_amd import synth
    dev = torch.device("cuda:0")
    ninst = 10000
    orders = (8 + np.floor(synth.uniform(2, ninst) * 89)).astype(int)
    gen = torch.Generator(device=dev)
    gen.manual_seed(2)
    groups = {}
    for o in sorted(set(orders.tolist())):
        ids = np.nonzero(orders == o)[0]
        n, l, m = synth.zernike_nlm(o)
        scale = torch.from_numpy(1.0 / (1.0 + n)).to(dev)
        c = torch.randn((len(ids), len(n), 2), generator=gen, dtype=torch.float64, device=dev)
        c *= scale[None, :, None]
        c[:, torch.from_numpy(m == 0).to(dev), 1] = 0.0
        g = torch.empty((len(ids),) + zest_amd.zt_shape(o), dtype=torch.float64, device=dev)
        groups[o] = (ids, c, g, torch.zeros_like(c))
    # jobs in instance order (not sorted): views into the per-order buffers
    bwd, fwd = [None] * ninst, [None] * ninst
    for o, (ids, c, g, co) in groups.items():
        for k, i in enumerate(ids):
            bwd[i] = (o, g[k], c[k])
            fwd[i] = (o, g[k], co[k])
    jb, jf = zest_amd.make_jobs(bwd), zest_amd.make_jobs(fwd)
    points = sum(int(np.prod(zest_amd.zt_shape(o))) * len(v[0]) for o, v in groups.items())
    out = {"metric": "Zernike ragged fwd+bwd transforms/s (10,000 instances, N~U{8..96})",
           "unit": "transforms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "data": "synthetic",
           "config": {"workload": "zernike_ragged_backward_forward", "instances": ninst,
                      "orders": "U{8..96}, seed 2", "distinct_orders": len(groups),
                      "grid_points": points, "conventions": "unnormed,geo,none"}}
    stream = torch.cuda.Stream()
    for name, prec in (("fp64", zest_amd.FP64), ("mixed", zest_amd.FP32)):
        plan = zest_amd.Plan(zest_amd.ZERNIKE, 8, precision=prec)

        def ragged_step():
            plan.backward_ragged(jb, stream=stream)
            plan.forward_ragged(jf, stream=stream)

        def plain_step():
            with torch.cuda.stream(stream):
                for o, (ids, c, g, co) in groups.items():
                    plan.backward(c, g, o, batch=len(ids), stream=stream)
                    plan.forward(g, co, o, batch=len(ids), stream=stream)

        t_rag = _timed_steps(ragged_step, args.steps, args.warmup)
        costs = plan.ragged_lane_costs()
        t_plain = _timed_steps(plain_step, args.steps, args.warmup)
        # the persistent pipeline (one persistent launch per stage and wave over
        # all orders) on the same jobs
        os.environ["ZGL_RAGGED"] = "persistent"
        pers_plan = zest_amd.Plan(zest_amd.ZERNIKE, 8, precision=prec)

        def pers_step():
            pers_plan.backward_ragged(jb, stream=stream)
            pers_plan.forward_ragged(jf, stream=stream)

        t_pers = _timed_steps(pers_step, args.steps, args.warmup)
        pcosts = pers_plan.ragged_lane_costs()
        del os.environ["ZGL_RAGGED"]
        del pers_plan
        out[name] = {"value": round(ninst * args.steps / t_rag, 2),
                     "ms_per_step": round(t_rag / args.steps * 1e3, 3),
                     "per_order_value": round(ninst * args.steps / t_plain, 2),
                     "per_order_ms_per_step": round(t_plain / args.steps * 1e3, 3),
                     "speedup_vs_per_order": round(t_plain / t_rag, 3),
                     "lane_cost_max_over_mean": round(max(costs) / (sum(costs) / len(costs)), 4),
                     "persistent_value": round(ninst * args.steps / t_pers, 2),
                     "persistent_ms_per_step": round(t_pers / args.steps * 1e3, 3),
                     "persistent_speedup_vs_per_order": round(t_plain / t_pers, 3),
                     "persistent_workgroups": len(pcosts),
                     "persistent_workgroup_cost_max_over_mean":
                         round(max(pcosts) / (sum(pcosts) / len(pcosts)), 4)}
        del plan
        torch.cuda.synchronize()
    # accuracy of the mixed form on a sample: the FP64 grids forwarded in FP32
    p64 = zest_amd.Plan(zest_amd.ZERNIKE, 8)
    p32 = zest_amd.Plan(zest_amd.ZERNIKE, 8, precision=zest_amd.FP32)
    worst = 0.0
    for o in (8, 50, 96):
        ids, c, g, co = groups[o]
        p64.backward(c, g, o, batch=len(ids))
        a, b = torch.zeros_like(c), torch.zeros_like(c)
        p64.forward(g, a, o, batch=len(ids))
        p32.forward(g, b, o, batch=len(ids))
        torch.cuda.synchronize()
        worst = max(worst, float(((b - a).abs().amax(dim=(1, 2)) / a.abs().amax(dim=(1, 2))).max()))
    out["mixed"]["fwd_max_rel_err_vs_fp64"] = worst
    out["value"] = out["fp64"]["value"]
    out["ms_per_step"] = out["fp64"]["ms_per_step"]
    out["dtype"] = "f64"
    print(json.dumps(out), flush=True)


def run_sh512(args):
    """4096 spherical-harmonic expansions at L = 512 (LonLat grids 1023 x 512),
    coefficients N(0,1) (1+l)^-1.5 (torch generator, seed 1); grids are the FP64
    backward transforms. Reports the forward error of FP32 and FP32 + one
    FP64-residual refinement against FP64 (max and RMS relative, per instance,
    worst instance), and backward+forward pairs/s of each form."""
    import numpy as np
    import torch
    import zest_amd
    from zest_amd import synth
    dev = torch.device("cuda:0")
    L, B = 512, args.instances or 4096
    l, m = synth.sh_lm(L)
    gen = torch.Generator(device=dev)
    gen.manual_seed(1)
    c = torch.randn((B, len(l), 2), generator=gen, dtype=torch.float64, device=dev)
    c *= torch.from_numpy((1.0 + l) ** -1.5).to(dev)[None, :, None]
    c[:, torch.from_numpy(m == 0).to(dev), 1] = 0.0
    g = torch.empty((B,) + zest_amd.st_shape(L), dtype=torch.float64, device=dev)
    # fp32_grid32: FP32 arithmetic on grids stored in FP32 (the forward reads
    # the FP32 rounding of the FP64 backward's grid)
    forms = (("fp64", zest_amd.FP64), ("fp32", zest_amd.FP32), ("fp32_grid32", zest_amd.FP32),
             ("fp32+refine", zest_amd.FP32_REFINE))
    if args.only == "fp64":  # profiling runs: the FP64 form alone
        forms = forms[:1]
    plans = {name: zest_amd.Plan(zest_amd.SH, L, precision=p) for name, p in forms}
    plans["fp64"].backward(c, g, L, batch=B)
    c64 = torch.zeros_like(c)
    plans["fp64"].forward(g, c64, L, batch=B)
    out = {"metric": "SH L=512 fwd+bwd transforms/s and FP32 / FP32+refine error vs FP64",
           "unit": "transforms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "data": "synthetic",
           "config": {"workload": "sh_glq_backward_forward", "order": L, "instances": B,
                      "grid": list(zest_amd.st_shape(L)), "conventions": "geo,none"}}
    cw = torch.zeros_like(c)
    for name, plan in plans.items():
        gin = g.float() if name == "fp32_grid32" else g
        gw = torch.empty_like(gin)
        cx = torch.zeros_like(c)
        plan.forward(gin, cx, L, batch=B)
        torch.cuda.synchronize()
        del gin
        d = cx - c64
        maxrel = float((d.abs().amax(dim=(1, 2)) / c64.abs().amax(dim=(1, 2))).max())
        rms = float((d.square().sum(dim=(1, 2)).sqrt() / c64.square().sum(dim=(1, 2)).sqrt()).max())

        def step():
            plan.backward(c, gw, L, batch=B)
            plan.forward(gw, cw, L, batch=B)

        t = _timed_steps(step, args.steps, args.warmup)
        # one more step with HIP events around every stage launch (outside the timing)
        plan.profile(True)
        plan.stage_stats()
        step()
        stages = {k: round(v[2] * 1e3, 1) for k, v in plan.stage_stats().items() if v[0]}
        plan.profile(False)
        out[name] = {"value": round(B * args.steps / t, 2), "ms_per_step": round(t / args.steps * 1e3, 3),
                     "fwd_max_rel_err_vs_fp64": maxrel, "fwd_rms_rel_err_vs_fp64": rms,
                     "stage_us_per_step": stages}
    top = "fp32+refine" if "fp32+refine" in out else "fp64"
    out["value"] = out[top]["value"]
    out["ms_per_step"] = out[top]["ms_per_step"]
    out["dtype"] = "f32" if top != "fp64" else "f64"
    print(json.dumps(out), flush=True)


def run_grad(args):
    """256 least-squares fits at order 64 (seed 3): targets g = r exp(-(a.x)^2)
    with a ~ U(0.5, 2)^3 on the ball grid; one step evaluates the loss
    sum_x W_x (B c - g)_x^2 and its gradient 2 B^T (W (B c - g)) through the
    hand-written transpose (zgl_backward_adjoint). The gradient is compared
    with the closed form 2 diag(1/nu) F(B c - g)."""
    import numpy as np
    import torch
    import zest_amd
    from zest_amd import synth
    dev = torch.device("cuda:0")
    N, B = 64, 256
    nphi, nth, nr = zest_amd.zt_shape(N)
    # the quadrature nodes of the ball grid (ascending, as zest's)
    z, wz = np.polynomial.legendre.leggauss(nth)
    x, wx = np.polynomial.legendre.leggauss(nr)
    r = 0.5 * (1.0 + x)
    phi = 2.0 * np.pi * np.arange(nphi) / nphi
    st = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    X = torch.from_numpy(r[None, None, :] * st[None, :, None] * np.cos(phi)[:, None, None]).to(dev)
    Y = torch.from_numpy(r[None, None, :] * st[None, :, None] * np.sin(phi)[:, None, None]).to(dev)
    Z = torch.from_numpy(r[None, None, :] * z[None, :, None] * np.ones(nphi)[:, None, None]).to(dev)
    R = torch.from_numpy(np.broadcast_to(r[None, None, :], (nphi, nth, nr)).copy()).to(dev)
    a = torch.from_numpy(synth.uniform(3, 3 * B).reshape(B, 3) * 1.5 + 0.5).to(dev)
    g = torch.stack([R * torch.exp(-(a[i, 0] * X + a[i, 1] * Y + a[i, 2] * Z) ** 2) for i in range(B)])
    W = torch.from_numpy(wz[None, :, None] * (r * r * wx)[None, None, :] * np.ones(nphi)[:, None, None]).to(dev)
    n, l, m = synth.zernike_nlm(N)
    nu = torch.from_numpy((1.0 / (4.0 * np.pi)) * 0.5 * (2.0 * np.pi / nphi) * (2.0 * n + 3.0)).to(dev)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, N)
    c = torch.zeros((B, len(n), 2), dtype=torch.float64, device=dev)
    plan.forward(g, c, N, batch=B)  # start from the projection of the target
    c *= 0.9
    bc = torch.empty_like(g)
    grad = torch.zeros_like(c)
    loss = torch.empty(B, dtype=torch.float64, device=dev)

    def step():
        plan.backward(c, bc, N, batch=B)
        res = bc - g
        torch.sum((W * res * res).view(B, -1), dim=1, out=loss)
        plan.backward_adjoint(W * res, grad, N, batch=B)
        grad.mul_(2.0)

    t = _timed_steps(step, args.steps, args.warmup)
    closed = torch.zeros_like(c)
    plan.forward(bc - g, closed, N, batch=B)
    closed = 2.0 * closed / nu[None, :, None]
    closed[:, torch.from_numpy(m == 0).to(dev), 1] = 0.0
    torch.cuda.synchronize()
    err = float(((grad - closed).abs().amax(dim=(1, 2)) / closed.abs().amax(dim=(1, 2))).max())
    out = {"metric": "Zernike least-squares loss+gradient evaluations/s (256 fits, order 64)",
           "value": round(B * args.steps / t, 2), "unit": "gradients/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t / args.steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "zernike_lsq_gradient", "order": N, "fits": B,
                      "target": "r exp(-(a.x)^2), a ~ U(0.5,2)^3, seed 3"},
           "gradient_max_rel_diff_vs_closed_form": err}
    print(json.dumps(out), flush=True)


def eval_work(order: int, nlon: int, nlat: int, nrad: int | None):
    """Algorithmic flops of one evaluation per stage (FP64): a real factor
    times a complex value is 2 multiply-adds, 4 flops."""
    ntri = order * (order + 1) // 2
    cols = nrad if nrad else 1
    w = {"eval_legendre": 4 * ntri * nlat * cols, "eval_longitude": 4 * order * nlon * nlat * cols}
    if nrad:
        w["eval_radial"] = 4 * ((order + 1) * (order + 3) * (2 * order + 1) // 24) * nrad
    return w


def _eval_grid(nlon, nlat, nrad):
    import numpy as np
    lon = 2.0 * np.pi * np.arange(nlon) / nlon
    col = np.pi * (np.arange(nlat) + 0.5) / nlat
    r = (np.arange(nrad) + 0.5) / nrad if nrad else None
    return lon, col, r


def run_eval(args):
    """zt::GridEvaluator::evaluate on a plotting-style grid: 64 Zernike
    expansions at order 128 (coefficients N(0,1)/(1+n), SplitMix64 seeds
    0..63, conventions unnormed, geo, none) on longitudes 2 pi i / 256,
    colatitudes pi (j + 1/2) / 128 and radii (k + 1/2) / 128. One step = one
    zgl_evaluate call for all 64 instances, inputs resident in HBM. Beside it,
    st::GridEvaluator for 64 SH expansions at L = 512 on 1024 x 512 points.
    The roofline is the longitude sum's (FP64 MFMA), timed per launch with HIP
    events; the CPU baseline is zest's GridEvaluator (oracle/_ref) on the host
    cores over a bounded sample."""
    import numpy as np
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    dev = torch.device("cuda:0")
    N, B, nlon, nlat, nrad = 128, 64, 256, 128, 128
    lon, col, r = _eval_grid(nlon, nlat, nrad)
    cs = np.stack([synth.zernike_coeffs(N, i) for i in range(B)])
    dc = torch.from_numpy(cs).to(dev)
    dl, dt, dr = (torch.from_numpy(a).to(dev) for a in (lon, col, r))
    vals = torch.empty((B, nlon, nlat, nrad), dtype=torch.float64, device=dev)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, N, zernike_norm=1, sh_norm=0, sh_phase=-1)

    def step():
        plan.evaluate(dc, N, dl, dt, dr, values=vals, batch=B)

    t = _timed_steps(step, args.steps, args.warmup)
    # accuracy: two instances against the oracle on a sub-grid
    sub = (slice(0, nlon, 16), slice(0, nlat, 8), slice(0, nrad, 8))
    worst = 0.0
    for b in (0, B - 1):
        ref = oracle.zt_evaluate(cs[b], N, lon[sub[0]], col[sub[1]], r[sub[2]])
        got = vals[b][sub].cpu().numpy()
        worst = max(worst, float(np.abs(got - ref).max() / np.abs(ref).max()))
    # per-stage kernel times over a few profiled steps
    plan.profile(True)
    for _ in range(3):
        step()
    stats = plan.stage_stats()
    plan.profile(False)
    work = eval_work(N, nlon, nlat, nrad)
    stages = {}
    for name, (launches, inst, ms) in stats.items():
        if name not in work or launches == 0:
            continue
        per_launch = work[name] * inst / launches
        ach = per_launch / (ms / launches / 1e3) / 1e12
        stages[name] = {"ms_per_launch": round(ms / launches, 4), "launches": launches,
                        "instances_per_launch": round(inst / launches, 3), "bound": "mfma",
                        "achieved": round(ach, 3), "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                        "frac": round(ach / PEAK_TFLOPS["f64"], 4)}
    lon_st = stages["eval_longitude"]
    out = {"metric": "Zernike grid evaluations/s (order 128, 256x128x128 points)",
           "value": round(B * args.steps / t, 2), "unit": "evaluations/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t / args.steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "zernike_grid_evaluate", "order": N, "instances": B,
                      "grid": [nlon, nlat, nrad], "conventions": "unnormed,geo,none"},
           "max_rel_err_vs_oracle_subgrid": worst,
           "stages": stages,
           "roofline": {"bound": "mfma", "kernel": "eval_lon_kernel",
                        "achieved": lon_st["achieved"], "peak": PEAK_TFLOPS["f64"],
                        "unit": "TFLOP/s", "frac": lon_st["frac"], "traffic": None}}
    del vals
    torch.cuda.synchronize()
    # SH: 64 expansions at L = 512 on 1024 x 512
    L, SB, slon, slat = 512, 64, 1024, 512
    slon_a, scol_a, _ = _eval_grid(slon, slat, 0)
    sc = np.stack([synth.sh_coeffs(L, i) for i in range(SB)])
    dsc = torch.from_numpy(sc).to(dev)
    dsl, dst = torch.from_numpy(slon_a).to(dev), torch.from_numpy(scol_a).to(dev)
    svals = torch.empty((SB, slon, slat), dtype=torch.float64, device=dev)
    splan = zest_amd.Plan(zest_amd.SH, 8)

    def sstep():
        splan.evaluate(dsc, L, dsl, dst, values=svals, batch=SB)

    ts = _timed_steps(sstep, max(2, args.steps // 4), 1)
    ref = oracle.st_evaluate(sc[0], L, slon_a[::32], scol_a[::16])
    serr = float(np.abs(svals[0][::32, ::16].cpu().numpy() - ref).max() / np.abs(ref).max())
    out["sh"] = {"value": round(SB * max(2, args.steps // 4) / ts, 2), "unit": "evaluations/s",
                 "order": L, "instances": SB, "grid": [slon, slat],
                 "max_rel_err_vs_oracle_subgrid": serr}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_eval_baseline(N, cs[0], lon, col, r, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def cpu_eval_baseline(order, c, lon, col, r, seconds):
    """zest's zt::GridEvaluator (or the oracle) on the host, one evaluator per thread."""
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    kind = "reference" if oracle.REF is not None else "port"
    fn = oracle.ref_zt_evaluate if kind == "reference" else oracle.zt_evaluate
    counts = [0] * threads
    t0 = time.perf_counter()
    deadline = t0 + seconds

    def work(i):
        while time.perf_counter() < deadline:
            fn(c, order, lon, col, r, 1, 0, -1)
            counts[i] += 1

    ths = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    elapsed = time.perf_counter() - t0
    total = sum(counts)
    return {"value": round(total / elapsed, 4), "unit": "evaluations/s", "cores": threads, "kind": kind,
            "sample": (f"{total} evaluations of one order-{order} expansion on the "
                       f"{len(lon)}x{len(col)}x{len(r)} grid in {elapsed:.1f} s, {threads} threads, "
                       + ("zest GridEvaluator compiled from its sources" if kind == "reference"
                          else "C restatement of zest"))}


def rotation_flops(zt: bool, order: int, lmax: int = 255) -> int:
    """Algorithmic flops of one rotation (FP64) over the degrees 1 <= l <=
    lmax: per block of degree l, two X steps, each (l+1)^2 multiply-adds (the
    two parity-masked products)."""
    total = 0
    for l in range(1, min(order, lmax + 1)):
        blocks = (order - l + 1) // 2 if zt else 1
        total += blocks * 2 * 2 * (l + 1) ** 2
    return total


def run_spectra(args):
    """zt::power_spectrum of 512 Zernike expansions at order 128 (coefficients
    N(0,1)/(1+n), torch generator seed 7; 2.9 MB each), one zgl_power_spectrum
    call for all of them per step, inputs resident in HBM; beside it the SH
    power spectra of 512 expansions at L = 512. The kernel reads every
    coefficient once and writes one value per (n, l): HBM-bound; the roofline
    is its algorithmic bytes per launch over the launch time (HIP events on
    the launch stream). The CPU baseline runs zest's zt::power_spectrum
    (oracle/_ref) on the host cores over a bounded sample."""
    import numpy as np
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    dev = torch.device("cuda:0")
    N, B = 128, 512
    n, _, m = synth.zernike_nlm(N)
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    c = torch.randn((B, n.size, 2), generator=gen, dtype=torch.float64, device=dev)
    c *= torch.from_numpy(1.0 / (1.0 + n)).to(dev)[None, :, None]
    c[:, torch.from_numpy(m == 0).to(dev), 1] = 0.0
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 4)
    nent = (N + 1) ** 2 // 4
    out = torch.empty((B, nent), dtype=torch.float64, device=dev)
    plan.power_spectrum(c, N, out, batch=B)
    torch.cuda.synchronize()
    host = c[:2].cpu().numpy()
    got = out[:2].cpu().numpy()
    worst = max(float(np.abs(got[b] - oracle.zt_power_spectrum(host[b], N)).max()
                      / np.abs(oracle.zt_power_spectrum(host[b], N)).max()) for b in range(2))

    def step():
        plan.power_spectrum(c, N, out, batch=B)

    t = _timed_steps(step, args.steps, args.warmup)
    plan.profile(True)
    plan.stage_stats()
    for _ in range(3):
        step()
    launches, inst, ms = plan.stage_stats()["power_spectrum"]
    plan.profile(False)
    per_inst = 16 * n.size + 8 * nent  # coefficient pairs read, one value written per (n, l)
    per_launch_ms = ms / launches
    ach = per_inst * inst / launches / (per_launch_ms / 1e3) / 1e9
    res = {"metric": "Zernike power spectra/s (order 128)",
           "value": round(B * args.steps / t, 2), "unit": "spectra/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t / args.steps * 1e3, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "zernike_power_spectrum", "order": N, "instances": B,
                      "coefficients": "N(0,1)/(1+n), torch seed 7"},
           "max_rel_err_vs_oracle": worst,
           "roofline": {"bound": "hbm", "kernel": "spectrum_kernel", "achieved": round(ach, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": None, "ms_per_launch": round(per_launch_ms, 4),
                        "bytes_per_launch": per_inst * B}}
    # SH, L = 512
    L = 512
    ls, ms_ = synth.sh_lm(L)
    sc = torch.randn((B, ls.size, 2), generator=gen, dtype=torch.float64, device=dev)
    splan = zest_amd.Plan(zest_amd.SH, 8)
    sout = torch.empty((B, L), dtype=torch.float64, device=dev)
    ts = _timed_steps(lambda: splan.power_spectrum(sc, L, sout, batch=B), args.steps, 1)
    shost = sc[0].cpu().numpy()
    serr = float(np.abs(sout[0].cpu().numpy() - oracle.st_power_spectrum(shost, L)).max()
                 / np.abs(oracle.st_power_spectrum(shost, L)).max())
    res["sh"] = {"value": round(B * args.steps / ts, 2), "unit": "spectra/s", "order": L, "instances": B,
                 "max_rel_err_vs_oracle": serr}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_spectrum_baseline(N, c[:8].cpu().numpy(), args.cpu_seconds)
    print(json.dumps(res), flush=True)


def cpu_spectrum_baseline(order, cs, seconds):
    """zest's zt::power_spectrum (or the oracle) on the host, one thread per core."""
    import numpy as np
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    kind = "reference" if oracle.REF is not None else "port"
    counts = [0] * threads
    nent = (order + 1) ** 2 // 4
    t0 = time.perf_counter()
    deadline = t0 + seconds

    def work(i):
        c = np.ascontiguousarray(cs[i % len(cs)])
        o = np.empty(nent)
        while time.perf_counter() < deadline:
            if kind == "reference":
                oracle.REF.zref_zt_power_spectrum(1, 0, order, oracle._p(c), oracle._p(o))
            else:
                o = oracle.zt_power_spectrum(c, order)
            counts[i] += 1

    ths = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    elapsed = time.perf_counter() - t0
    total = sum(counts)
    return {"value": round(total / elapsed, 2), "unit": "spectra/s", "cores": threads, "kind": kind,
            "sample": f"{total} zt::power_spectrum calls at order {order} in {elapsed:.1f} s, "
                      f"{threads} threads"}


def run_rotate(args):
    """Rotor::rotate of 64 Zernike expansions at order 128 (coefficients
    N(0,1)/(1+n), SplitMix64 seeds 0..63), each by its own Euler angles
    (uniform in [-pi, pi), seed 5), object rotation. One step = one
    zgl_rotate call for all 64 instances in place, inputs resident in HBM.
    Beside it 64 SH expansions at L = 512. The roofline is the rotation
    kernel's (FP64 MFMA), per launch with HIP events; the CPU baseline runs
    zest's Rotor (oracle/_ref) on the host cores over a bounded sample."""
    import numpy as np
    import torch
    import oracle
    import zest_amd
    from zest_amd import synth
    dev = torch.device("cuda:0")
    N, B = 128, 64
    cs = np.stack([synth.zernike_coeffs(N, i) for i in range(B)])
    ang = synth.uniform(5, 3 * B).reshape(B, 3) * 2 * np.pi - np.pi
    dc, da = torch.from_numpy(cs).to(dev), torch.from_numpy(ang).to(dev)
    plan = zest_amd.Plan(zest_amd.ZERNIKE, 4)
    # accuracy of one rotation against the oracle (two instances)
    plan.rotate(dc, N, da, 0, batch=B)
    got = dc.cpu().numpy()
    worst = max(float(np.abs(got[b] - oracle.rotate(cs[b], N, oracle.ZERNIKE_KIND, ang[b], 0)).max()
                      / np.abs(cs[b]).max()) for b in (0, B - 1))

    def step():
        plan.rotate(dc, N, da, 0, batch=B)

    t = _timed_steps(step, args.steps, args.warmup)
    plan.profile(True)
    for _ in range(3):
        step()
    stats = plan.stage_stats()
    plan.profile(False)
    launches, inst, ms = stats["rotate"]
    per_launch = rotation_flops(True, N) * inst / launches  # the fused kernel: the whole rotation
    ach = per_launch / (ms / launches / 1e3) / 1e12
    out = {"metric": "Zernike rotations/s (order 128, ZYZ Euler angles per instance)",
           "value": round(B * args.steps / t, 2), "unit": "rotations/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t / args.steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "zernike_rotate", "order": N, "instances": B,
                      "angles": "U[-pi, pi)^3 per instance, seed 5", "rotation_type": "object"},
           "max_rel_err_vs_oracle": worst,
           "roofline": {"bound": "mfma", "kernel": "rotate_kernel", "achieved": round(ach, 3),
                        "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                        "frac": round(ach / PEAK_TFLOPS["f64"], 4), "traffic": None,
                        "ms_per_launch": round(ms / launches, 4)}}
    # SH, L = 512
    L = 512
    sc = np.stack([synth.sh_coeffs(L, i) for i in range(B)])
    sang = synth.uniform(6, 3 * B).reshape(B, 3) * 2 * np.pi - np.pi
    dsc, dsa = torch.from_numpy(sc).to(dev), torch.from_numpy(sang).to(dev)
    splan = zest_amd.Plan(zest_amd.SH, 8)
    splan.rotate(dsc, L, dsa, 0, batch=B)
    serr = float(np.abs(dsc[0].cpu().numpy() - oracle.rotate(sc[0], L, oracle.SH_KIND, sang[0], 0)).max()
                 / np.abs(sc[0]).max())
    ts = _timed_steps(lambda: splan.rotate(dsc, L, dsa, 0, batch=B), args.steps, 1)
    out["sh"] = {"value": round(B * args.steps / ts, 2), "unit": "rotations/s", "order": L, "instances": B,
                 "max_rel_err_vs_oracle": serr}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_rotate_baseline(N, cs, ang, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def cpu_rotate_baseline(order, cs, ang, seconds):
    """zest's Rotor (or the oracle) on the host, one Rotor per thread."""
    import numpy as np
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    kind = "reference" if oracle.REF is not None else "port"
    counts = [0] * threads
    t0 = time.perf_counter()
    deadline = t0 + seconds

    per_call = 8  # rotations per call: one d table and one Rotor serve them

    def work(i):
        c = np.array(cs[:per_call], copy=True)
        a = np.ascontiguousarray(ang[:per_call])
        while time.perf_counter() < deadline:
            if kind == "reference":
                oracle.REF.zref_rotate_many(0, order, oracle._p(c), c[0].size, per_call, oracle._p(a), 0)
            else:
                for k in range(per_call):
                    c[k] = oracle.rotate(c[k], order, oracle.ZERNIKE_KIND, a[k], 0)
            counts[i] += per_call

    ths = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    elapsed = time.perf_counter() - t0
    total = sum(counts)
    return {"value": round(total / elapsed, 3), "unit": "rotations/s", "cores": threads, "kind": kind,
            "sample": (f"{total} rotations of order-{order} Zernike expansions in {elapsed:.1f} s, "
                       f"{threads} threads, "
                       + ("zest Rotor compiled from its sources, one WignerdPiHalfCollection per 8 "
                          "rotations" if kind == "reference" else "C restatement of zest"))}


if __name__ == "__main__":
    main()
