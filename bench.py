#!/usr/bin/env python3
"""PatchMatch stereo benchmark: Mpix*iter/s of PMHuberStereo::depthMap on one MI355X.

Workload (default): one 1242x375 fp32 gray pair, 16 slanted planes, disparity
0-192, noise sigma 0.01, window 35x35, 4 iterations of red-black propagation
+ random refinement, L/R tolerance 1 px.  A step is one depthMap call over the
pair (inputs resident in HBM, disparity map resident in HBM at the end).

Prints one JSON line with the metric, the roofline of the dominant kernel
(propagate: window cost, VALU-bound, timed live with HIP events on the
context's stream), the CPU baseline (the oracle on a bounded sample of the
same workload, timed on this host's cores) and the batched method (batch
launches, coarse-to-fine start, gradient plane refinement, graph replay) on
the same pair beside the plain one at their bad-1.0px errors.

--workload c4 / c5 compare the plain method (one depthMap call per pair) with
the batched method (one pmh_batch_depthmap call over all pairs) on 64 uneven
pairs (W 320-1920, H 240-1080, disparity 32-256, 4 iterations) or on 8
2880x1988 pairs (32 planes, disparity 0-400, noise 0.02, 8 iterations), and
print one JSON line in Mpix/s (pixels of all pairs per second of wall time).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

METRIC = "PatchMatch Mpix\u00b7iter/s on 1 GPU; batched+adjoint vs plain at equal bad-1.0px %"
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak (published)
FLOPS_PER_TAP = 42         # algorithmic flops of one window tap (FMA = 2)
CANDIDATES = 18            # candidates per pixel per view per iteration, lower bound


TRAFFIC_FILE = "profiles/propagate_1242x375_counters.md"


def measured_traffic():
    """HBM bytes per propagate launch on this workload: rocprofv3 FETCH_SIZE and
    WRITE_SIZE per dispatch, each scaled by a 512 MiB copy's reading (the
    counter summary holds the commands and the calibration)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), TRAFFIC_FILE)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch propagate_kernel:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def batched_options(_lib, **over):
    """The batched method: red-black propagation, a coarse-to-fine start (30
    Huber-ROF steps per smoothing at the coarse levels, where the smoothed
    field only feeds the smooth-plane candidate), gradient plane-refinement
    steps after each view's propagation, graph replay."""
    kw = dict(sweep=_lib.SWEEP_REDBLACK, maxdispdiff=1.0, coarse_levels=2, coarse_iters=5,
              coarse_patch=6, coarse_pd_iters=30, grad_steps=2, use_graph=1)
    kw.update(over)
    return _lib.default_options(**kw)


# Configurations of the batched method tried, fastest first (coarse levels,
# iterations per coarse level, coarse half window, gradient steps per view
# per iteration, randomized refinements and random-neighbourhood candidates
# per pixel at full resolution (the coarse levels keep the reference's 2 and
# 6), full-resolution iterations); the comparison reports the fastest rung
# whose bad-1.0px is at or below the plain method's (every rung is listed in
# the output).
def _rung(levels, patch, grad, refine, neighbours, iterations):
    return {"coarse_levels": levels, "coarse_iters": 5, "coarse_patch": patch,
            "grad_steps": grad, "random_refine": refine, "coarse_random_refine": 2,
            "random_neighbours": neighbours, "coarse_random_neighbours": 6,
            "iterations": iterations}


BATCHED_LADDER = (
    _rung(1, 5, 1, 0, 0, 1),
    _rung(1, 6, 1, 0, 0, 1),
    _rung(1, 5, 1, 0, 6, 1),
    _rung(2, 6, 2, 0, 0, 1),
    _rung(2, 6, 2, 0, 6, 1),
    _rung(2, 6, 2, 2, 6, 1),
)


def rung_options(_lib, rung):
    return batched_options(_lib, **{k: v for k, v in rung.items() if k != "iterations"})


def describe_rung(rung):
    w = 2 * rung["coarse_patch"] + 1
    return (f"coarse-to-fine start ({rung['coarse_levels']} level(s) x {rung['coarse_iters']} "
            f"iterations, {w}x{w} window, 30 Huber-ROF steps), {rung['iterations']} "
            f"full-resolution red-black iteration(s) with {rung['grad_steps']} gradient "
            f"plane-refinement step(s) per view, {rung['random_refine']} randomized "
            f"refinement(s) and {rung['random_neighbours']} random-neighbourhood candidates "
            "per pixel, graph replay")


# stage codes of the tuning file (gray): 0 is the default
PROPAGATE_KERNELS = {0: "propagate_block_kernel<1792>",
                     1: "propagate_split_kernel<896, 1, false>",
                     2: "propagate_split_kernel<896, 1, true>",
                     3: "propagate_split3_kernel<576, 1>",
                     4: "propagate_block_kernel<1792>",
                     5: "propagate_block3_kernel<1152>"}


def propagate_kernel_name(ctx):
    """The window-cost kernels the autotuner chose for the full-resolution
    propagation launches (one per view and colour; joined when they differ)."""
    import tempfile
    names = []
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tuning.txt")
        ctx.tuning_save(path)
        for line in open(path):
            key, stage, chunk = line.split()
            if key.startswith("propagate|p17|"):
                name = (f"{PROPAGATE_KERNELS.get(int(stage), f'propagate_kernel<{stage}, 1>')}, "
                        f"{chunk} tile(s) per block")
                if name not in names:
                    names.append(name)
    return " / ".join(sorted(names)) or PROPAGATE_KERNELS[0]


def pick_rung(rungs, plain_bad):
    """Fastest rung at or below the plain method's error, else the most accurate."""
    ok = [r for r in rungs if r["bad_1px_percent"] <= plain_bad]
    if ok:
        return min(ok, key=lambda r: r["seconds"]), True
    return min(rungs, key=lambda r: r["bad_1px_percent"]), False


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("plain", "c4", "c5"), default="plain",
                    help="plain: 1242x375 pair (default); c4: 64 uneven pairs; "
                         "c5: 8 pairs of 2880x1988")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-method comparison (plain workload only)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy a MIB-sized device buffer once (a kernel of "
                         "known traffic for scaling hardware byte counters)")
    ap.add_argument("--cpu-rows", type=int, default=375,
                    help="rows of the pair in the CPU-baseline sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): pairs shard across GPUs "
              "with no communication, so the per-GPU figure is the measurement", file=sys.stderr)
        return 2
    if args.workload != "plain":
        return bench_pairs(args)
    return bench_plain(args)


def bench_plain(args) -> int:
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
    for k in ("propagate", "pd_steps", "consistency", "fill_state", "viewbuf", "init", "prep",
              "zero", "fill_depthmap"):
        ms, n = ctx.profile_read(k)
        kernels[k] = {"ms_per_step": round(ms / args.steps, 4), "launches_per_step": n // args.steps,
                      "avg_us": round(1e3 * ms / max(n, 1), 2)}
    prop_ms, prop_n = ctx.profile_read("propagate")
    avg_s = prop_ms * 1e-3 / prop_n
    px_per_launch = W * H / 2  # red-black: one colour per launch
    flops = px_per_launch * CANDIDATES * (2 * P + 1) ** 2 * FLOPS_PER_TAP
    achieved = flops / avg_s / 1e12
    roofline = {"bound": "valu", "kernel": propagate_kernel_name(ctx),
                "achieved": round(achieved, 3),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": measured_traffic(),
                "traffic_source": TRAFFIC_FILE,
                "avg_launch_ms": round(avg_s * 1e3, 4),
                "flops_per_launch": flops}

    if args.copy_calibration > 0:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    batched = None
    if not args.no_batched:
        batched = run_batched(args, L, R, P, dmax, pair, ms_per_step, bad)

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
               "sample": (f"the whole {W}x{H} pair" if rows == H else
                          f"{W}x{rows} band of the {W}x{H} pair") + ", 1 iteration, red-black, "
                         f"window 35x35, full pipeline (init, propagate, consistency, fill, "
                         f"250-step Huber-ROF x2); {t1 - t0:.1f} s on {cpu_model()}"}

    out = {
        "metric": METRIC,
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
        "batched_vs_plain": batched,
        "kernels": kernels,
    }
    print(json.dumps(out))
    return 0


def run_batched(args, L, R, P, dmax, pair, ms_per_step, bad):
    """The batched method on the plain workload's pair (a batch of one)."""
    import torch

    from cvt_amd import _lib, ops, synth
    rungs = []
    for rung in BATCHED_LADDER:
        bopts, it = rung_options(_lib, rung), rung["iterations"]
        for _ in range(2):  # first run tunes and runs eagerly, the second captures
            ops.depthmap(L, R, P, dmax, it, 4, 1.0, bopts)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            bmap = ops.depthmap(L, R, P, dmax, it, 4, 1.0, bopts)["dmap"]
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / args.steps
        rungs.append(dict(rung, seconds=round(sec, 5),
                          bad_1px_percent=round(synth.bad_pixels(bmap.cpu().numpy(), pair), 3)))
    pick, met = pick_rung(rungs, round(bad, 3))
    return {"method": describe_rung(pick),
            "ms_per_pair": round(pick["seconds"] * 1e3, 3),
            "bad_1px_percent": pick["bad_1px_percent"],
            "plain_ms_per_pair": round(ms_per_step, 3), "plain_bad_1px_percent": round(bad, 3),
            "speedup_mpix_per_s": round(ms_per_step / (pick["seconds"] * 1e3), 3),
            "equal_or_lower_error": met, "ladder": rungs}


def bench_pairs(args) -> int:
    """Plain (one call per pair) against batched (one call for all pairs)."""
    import numpy as np
    import torch

    from cvt_amd import _lib, ops, synth

    P = 17
    t0 = time.perf_counter()
    if args.workload == "c4":
        pairs = synth.uneven_batch(64, seed=4)
        plain_iters, desc = 4, ("64 uneven pairs (W 320-1920, H 240-1080, disparity 32-256, "
                                "4-24 planes, noise 0.01), window 35x35")
    else:
        pairs = [synth.make_pair(2880, 1988, 32, 400.0, 0.02, seed=5 + k) for k in range(8)]
        plain_iters, desc = 8, ("8 pairs of 2880x1988 (32 planes, disparity 0-400, noise 0.02), "
                                "window 35x35")
    print(f"made {len(pairs)} pairs in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    dev = torch.device("cuda")
    dev_pairs = [(torch.from_numpy(p.left).to(dev), torch.from_numpy(p.right).to(dev), p.dmax,
                  1.0) for p in pairs]
    mpx = sum(p.left.size for p in pairs) / 1e6

    def bad_all(maps):
        # pixel-weighted bad-1.0px over all pairs' non-occluded pixels
        n_bad = n_all = 0
        for m, p in zip(maps, pairs):
            b = synth.bad_pixels(m.cpu().numpy(), p)
            k = int(p.nonocc.sum())
            n_bad += b * k / 100.0
            n_all += k
        return 100.0 * n_bad / n_all

    popts = _lib.default_options(sweep=_lib.SWEEP_REDBLACK, maxdispdiff=1.0)
    # warm-up: one iteration of every pair tunes every shape class (the tuning
    # keys do not depend on the iteration count), so no tuning is timed
    for l, r, d, s in dev_pairs:
        ops.depthmap(l, r, P, d, 1, 4, s, popts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pmaps = []
    for k, (l, r, d, s) in enumerate(dev_pairs):
        pmaps.append(ops.depthmap(l, r, P, d, plain_iters, 4, s, popts)["dmap"])
        print(f"plain pair {k + 1}/{len(pairs)}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    plain_bad = bad_all(pmaps)
    del pmaps

    rungs = []
    for rung in BATCHED_LADDER:
        bopts, it = rung_options(_lib, rung), rung["iterations"]
        for w in range(2):  # first call tunes and runs eagerly, the second captures the graph
            ops.batch_depthmap(dev_pairs, P, it, bopts)
            torch.cuda.synchronize()
            print(f"batched ({describe_rung(rung)[:40]}...) warm-up {w + 1}/2", file=sys.stderr,
                  flush=True)
        g0 = ops.context().graph_stats()
        t0 = time.perf_counter()
        res = ops.batch_depthmap(dev_pairs, P, it, bopts)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        g1 = ops.context().graph_stats()
        rungs.append(dict(rung, seconds=round(sec, 3), mpix_per_s=round(mpx / sec, 4),
                          bad_1px_percent=round(bad_all([r["dmap"] for r in res]), 3),
                          graph_replays=g1[1] - g0[1]))
        del res
    pick, met = pick_rung(rungs, round(plain_bad, 3))
    batched_s = pick["seconds"]

    out = {
        "metric": METRIC,
        "value": round(mpx / batched_s, 4), "unit": "Mpix/s", "n_gpus": 1, "steps": 1,
        "warmup": 2, "ms_per_step": round(batched_s * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded slanted-plane Voronoi scenes, band-limited texture)",
        "config": {"workload": f"batched vs plain GPU PatchMatch, {desc}",
                   "pairs": len(pairs), "mpix": round(mpx, 3), "patch_half": P},
        "batched": {"method": "one pmh_batch_depthmap call over all pairs: " +
                              describe_rung(pick),
                    "seconds": pick["seconds"], "mpix_per_s": pick["mpix_per_s"],
                    "bad_1px_percent": pick["bad_1px_percent"],
                    "graph_replays": pick["graph_replays"],
                    "equal_or_lower_error": met, "ladder": rungs},
        "plain": {"method": f"one depthMap call per pair, {plain_iters} red-black iterations, "
                            "random refinement",
                  "seconds": round(plain_s, 3), "mpix_per_s": round(mpx / plain_s, 4),
                  "mpix_iter_per_s": round(mpx * plain_iters / plain_s, 4),
                  "bad_1px_percent": round(plain_bad, 3)},
        "speedup": round(plain_s / batched_s, 3),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
