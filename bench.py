#!/usr/bin/env python3
"""PatchMatch stereo benchmark: Mpix*iter/s of PMHuberStereo::depthMap on one MI355X.

Workload (default): one 1242x375 fp32 gray pair, 16 slanted planes, disparity
0-192, noise sigma 0.01, window 35x35, 4 iterations of red-black propagation
+ random refinement, L/R tolerance 1 px.  A step is one depthMap call over the
pair (inputs resident in HBM, disparity map resident in HBM at the end).

Prints one JSON line with the metric, the roofline of the dominant kernel
(propagate: window cost, VALU-bound, timed live with HIP events on the
context's stream) and the CPU baseline (the oracle on a bounded sample of the
same workload, timed on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak (published)
FLOPS_PER_TAP = 42         # algorithmic flops of one window tap (FMA = 2)
CANDIDATES = 18            # candidates per pixel per view per iteration, lower bound


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-rows", type=int, default=128,
                    help="rows of the pair in the CPU-baseline sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): pairs shard across GPUs "
              "with no communication, so the per-GPU figure is the measurement", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from cvt_amd import _lib, ops, synth

    W, H, n_planes, dmax, noise, P, iters = 1242, 375, 16, 192.0, 0.01, 17, 4
    pair = synth.make_pair(W, H, n_planes, dmax, noise, seed=2)
    dev = torch.device("cuda")
    L = torch.from_numpy(pair.left).to(dev)
    R = torch.from_numpy(pair.right).to(dev)
    opts = _lib.default_options(sweep=_lib.SWEEP_REDBLACK, maxdispdiff=1.0)
    ctx = ops.context()

    def step():
        return ops.depthmap(L, R, P, dmax, iters, 4, 1.0, opts)["dmap"]

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dmap = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.profile(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = W * H * iters / (ms_per_step * 1e-3) / 1e6
    bad = synth.bad_pixels(dmap.cpu().numpy(), pair)

    kernels = {}
    for k in ("propagate", "pd_step", "consistency", "fill_state", "viewbuf", "init", "prep",
              "fill_depthmap"):
        ms, n = ctx.profile_read(k)
        kernels[k] = {"ms_per_step": round(ms / args.steps, 4), "launches_per_step": n // args.steps,
                      "avg_us": round(1e3 * ms / max(n, 1), 2)}
    prop_ms, prop_n = ctx.profile_read("propagate")
    avg_s = prop_ms * 1e-3 / prop_n
    px_per_launch = W * H / 2  # red-black: one colour per launch
    flops = px_per_launch * CANDIDATES * (2 * P + 1) ** 2 * FLOPS_PER_TAP
    achieved = flops / avg_s / 1e12
    roofline = {"bound": "valu", "kernel": "propagate_kernel", "achieved": round(achieved, 3),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                "avg_launch_ms": round(avg_s * 1e3, 4),
                "flops_per_launch": flops}

    cpu = None
    if not args.no_cpu:
        import oracle
        rows = min(args.cpu_rows, H)
        sample_l, sample_r = pair.left[:rows].copy(), pair.right[:rows].copy()
        t0 = time.perf_counter()
        oracle.depthmap(sample_l, sample_r, P, dmax, 1, 1.0, maxdispdiff=1.0, sweep=1)
        t1 = time.perf_counter()
        cpu = {"value": round(W * rows * 1 / (t1 - t0) / 1e6, 5), "unit": "Mpix*iter/s",
               "cores": oracle.num_threads(), "kind": "port",
               "sample": f"{W}x{rows} band of the {W}x{H} pair, 1 iteration, red-black, "
                         f"window 35x35, full pipeline (init, propagate, consistency, fill, "
                         f"250-step Huber-ROF x2); {t1 - t0:.1f} s on {cpu_model()}"}

    out = {
        "metric": "PatchMatch Mpix*iter/s on 1 GPU (plain method, red-black)",
        "value": round(value, 4), "unit": "Mpix*iter/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded slanted-plane Voronoi scene, band-limited texture)",
        "config": {"workload": "plain GPU PatchMatch, 1242x375 pair, 16 planes, disparity 0-192, "
                               "noise 0.01, window 35x35, 4 iterations red-black, L/R tol 1 px",
                   "width": W, "height": H, "patch_half": P, "iterations": iters,
                   "bad_1px_percent": round(bad, 3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
