#!/usr/bin/env python3
"""PatchMatch stereo benchmark: Mpix*iter/s of PMHuberStereo::depthMap on one MI355X.

Workload of `value` (default): one 1242x375 fp32 gray pair, 16 slanted planes,
disparity 0-192, noise sigma 0.01, window 35x35, 4 iterations of red-black
propagation + random refinement, L/R tolerance 1 px.  A step is one depthMap
call over the pair (inputs resident in HBM, disparity map resident in HBM at
the end).  The smoothing after the last iteration is not run: nothing reads
its output (the final check reads the plane fields), so the disparity map is
the reference sequence's.

The same JSON line carries
  - roofline: the propagate kernel (window cost, VALU FP32), timed live with
    HIP events on the context's stream inside the timed region (only the
    propagate launches are bracketed); its algorithmic flops count every
    candidate the algorithm costs per pixel (16 + the pixel's view
    candidates + 2 refinements, counted on the device) x 35^2 taps x 42 flop;
  - cpu_baseline: the oracle (the CPU restatement, OpenMP on this host's
    cores) on the whole pair for 1 iteration, and on the 450x375 pair
    (window 21x21, 3 iterations, the reference's Jacobi sweep) in full beside
    the GPU on that pair;
  - batched: the batched method -- fixed in advance, see BATCHED_METHOD -- on
    the 1242x375 pair, on 64 uneven pairs and on 8 pairs of 2880x1988, each
    beside the plain method at its bad-1.0px; on the 1242x375 pair also an
    ablation (batching and graph replay alone, gradient refinement without the
    coarse-to-fine start, the coarse-to-fine start without gradient steps);
  - multiview: the multi-view method on a reference and two source views of
    1242x375 (MV_WORKLOAD), with its propagate kernel's roofline and the oracle
    on the same workload;
  - adjoint: the differentiation workload -- the 1242x375 pair's window cost
    at a converged plane field, and its adjoint with respect to the planes
    alone and to the planes and both images (pmh_cost / pmh_cost_backward,
    window 35x35), each as a caller makes it, outputs allocated;
  - variants: depthMapInpaint on the gray pair and depthMap on an RGBA pair of
    the same scene, each with its propagate kernel's roofline and the oracle
    beside the GPU on the 450x375 pair.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

METRIC = "PatchMatch Mpix·iter/s on 1 GPU; batched+adjoint vs plain at equal bad-1.0px %"
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak (published)
# The published peak counts packed FP32 (two FMAs per lane per instruction).
# This build emits no packed FP32 operations, so its VALU ceiling is one FMA
# per lane per cycle: 256 CUs x 64 lanes x 2 flop x 2.4 GHz.
UNPACKED_FP32_TFLOPS = 78.6
FLOPS_PER_TAP = 42         # algorithmic flops of one window tap (FMA = 2)
# RGBA inputs: the colour term sums R, G and B (3 x (fma, difference,
# truncated magnitude) + 2 adds + the 0.05 scale: 18 flop instead of gray's
# 6) and the support weight's colour distance sums three channels (8 instead
# of 2), so a tap is 42 + 12 + 6 flop.
RGBA_FLOPS_PER_TAP = 60
LOWER_BOUND_CANDIDATES = 18  # candidates per pixel per pass without view candidates

TRAFFIC_FILE = "profiles/propagate_1242x375_counters.md"

# Multi-view workload: a reference view and two source views of 1242x375 of a
# rendered scene (slanted background at depth 10-12, 4 slanted rectangles at
# 4-8, depth range 3-14), window 9x9, 4 iterations of red-black passes with 2
# randomized refinements per pixel and pass.  Flops of one window tap: 5 per
# tap (plane-side test, weight) + 61 per source view (projection 6, 1 / Z
# and x', y' 3, fractions 2, five bilinear samples 30, truncated colour and
# gradient differences 11, sums 3, the tap's cost 3, accumulation 3); FMA = 2.
MV_WORKLOAD = {"width": 1242, "height": 375, "sources": 2, "seed": 7, "patch_half": 4,
               "iterations": 4, "refine": 2}
MV_FLOPS_PER_TAP, MV_FLOPS_PER_TAP_VIEW = 5, 61
MV_TRAFFIC_FILE = "profiles/mv_propagate_1242x375_counters.md"

# The batched method, fixed before the reported workloads were measured: of
# the configurations tried on held-out seeds (the 1242x375 pair with seed 12,
# 64 uneven pairs with seed 14, 4 pairs of 2880x1988 with seeds 105-108), the
# fastest (geometric mean of the three speedups) whose bad-1.0px was at least
# 3% below the plain method's on each.  Coarse-to-fine start over 3 levels of
# 2x2 box reduction, 6 iterations per level with a 9x9 window, 30 Huber-ROF
# steps and the reference's randomized refinements and random neighbours;
# then 1 full-resolution iteration with 2 gradient plane-refinement steps per
# view instead of the randomized refinements and random neighbours; every
# step a launch over all pairs; the whole run one replayed HIP graph.
BATCHED_METHOD = {"coarse_levels": 3, "coarse_iters": 6, "coarse_patch": 4,
                  "coarse_pd_iters": 30, "coarse_random_refine": 2,
                  "coarse_random_neighbours": 6, "coarse_joint_views": 1, "grad_steps": 2,
                  "random_refine": 0, "random_neighbours": 0, "iterations": 1}


def method_options(_lib, method):
    kw = {k: v for k, v in method.items() if k != "iterations"}
    return _lib.default_options(sweep=_lib.SWEEP_REDBLACK, maxdispdiff=1.0, use_graph=1, **kw)


def describe_method(m):
    w = 2 * m.get("coarse_patch", 0) + 1
    parts = []
    if m.get("coarse_levels", 0):
        joint = ", both views per launch" if m.get("coarse_joint_views", 0) else ""
        parts.append(f"coarse-to-fine start ({m['coarse_levels']} level(s) x "
                     f"{m['coarse_iters']} iterations, {w}x{w} window, "
                     f"{m['coarse_pd_iters']} Huber-ROF steps{joint})")
    parts.append(f"{m['iterations']} full-resolution red-black iteration(s) with "
                 f"{m.get('grad_steps', 0)} gradient plane-refinement step(s) per view, "
                 f"{m.get('random_refine', 2)} randomized refinement(s) and "
                 f"{m.get('random_neighbours', 6)} random-neighbourhood candidates per pixel")
    return ", ".join(parts) + ", all pairs per launch, graph replay"


def ablation(plain_iters):
    """Rows of the ablation on the plain workload's pair (iterations from the
    plain method's count where a row keeps the reference's schedule)."""
    return [
        ("batching and graph replay alone (the plain method's options)",
         {"iterations": plain_iters}),
        ("gradient refinement without the coarse-to-fine start (half the iterations, "
         "2 gradient steps per view)",
         {"iterations": max(1, plain_iters // 2), "grad_steps": 2}),
        ("the coarse-to-fine start without gradient steps",
         dict(BATCHED_METHOD, grad_steps=0)),
        ("the batched method with the reference's view order at the coarse levels",
         dict(BATCHED_METHOD, coarse_joint_views=0)),
        ("the batched method", dict(BATCHED_METHOD)),
    ]


def measured_traffic(path=TRAFFIC_FILE, key="propagate_kernel"):
    """HBM bytes per propagate launch on this workload: rocprofv3 FETCH_SIZE and
    WRITE_SIZE per dispatch, each scaled by a 512 MiB copy's reading (the
    counter summary holds the commands and the calibration)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), path)
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch {key}:"):
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


# stage codes of the tuning file (gray): 0 is the default
PROPAGATE_KERNELS = {0: "propagate_block_kernel<1792>",
                     1: "propagate_split_kernel<896, 1, false>",
                     2: "propagate_split_kernel<896, 1, true>",
                     3: "propagate_split3_kernel<576, 1>",
                     4: "propagate_block_kernel<1792>",
                     5: "propagate_block3_kernel<1152>"}
# RGBA inputs: the block-stage codes (4, 5) take the default shape (code 2)
PROPAGATE_KERNELS_RGBA = {0: "propagate_split_kernel<512, 4, true>",
                          1: "propagate_split_kernel<512, 4, false>",
                          2: "propagate_split_kernel<512, 4, true>",
                          3: "propagate_split3_kernel<352, 4>",
                          4: "propagate_split_kernel<512, 4, true>",
                          5: "propagate_split_kernel<512, 4, true>"}


def propagate_kernel_name(ctx, P, nc=1):
    """The window-cost kernels the autotuner chose for the propagate launches
    of half window P on inputs of nc channels (joined when they differ)."""
    import tempfile
    names = []
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tuning.txt")
        ctx.tuning_save(path)
        for line in open(path):
            key, stage, chunk = line.split()
            if key.startswith(f"propagate|p{P}|c{nc}|"):
                blocks = "persistent blocks" if chunk == "0" else f"{chunk} tile(s) per block"
                table = PROPAGATE_KERNELS if nc == 1 else PROPAGATE_KERNELS_RGBA
                name = (f"{table.get(int(stage), f'propagate_kernel<{stage}, {nc}>')}, "
                        f"{blocks}")
                if name not in names:
                    names.append(name)
    return " / ".join(sorted(names)) or PROPAGATE_KERNELS[0]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched method and the ablation")
    ap.add_argument("--no-pairs", action="store_true",
                    help="skip the 64-pair and 8 x 2880x1988 comparisons")
    ap.add_argument("--ablation-pairs", action="store_true",
                    help="also run the ablation rows on the 64-pair and 2880x1988 workloads")
    ap.add_argument("--no-multiview", action="store_true", help="skip the multi-view leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the depthMapInpaint and RGBA legs")
    ap.add_argument("--multiview-only", action="store_true",
                    help="run only the multi-view leg (GPU, no CPU baseline) and print its object")
    ap.add_argument("--pairs-only", choices=("c4", "c5"),
                    help="run only one batched comparison (64 uneven pairs or 8 x 2880x1988) "
                         "and print its object")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy a MIB-sized device buffer once (a kernel of "
                         "known traffic for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): pairs shard across GPUs "
              "with no communication, so the per-GPU figure is the measurement", file=sys.stderr)
        return 2
    if args.pairs_only:
        print(json.dumps(run_pairs(args, args.pairs_only)))
        return 0
    if args.multiview_only:
        res = run_multiview(args, cpu=False)
        copy_calibration(args)
        print(json.dumps(res))
        return 0
    return bench(args)


def copy_calibration(args):
    """Copies a MIB-sized device buffer once: a kernel of known traffic for
    scaling hardware byte counters."""
    if args.copy_calibration <= 0:
        return
    import torch
    n = args.copy_calibration * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def bench(args) -> int:
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
    # timed region: only the propagate launches are bracketed with events
    ctx.profile(True)
    ctx.profile_only("propagate")
    ctx.profile_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dmap = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prop_ms, prop_n = ctx.profile_read("propagate")
    cands = ctx.profile_candidates()
    ctx.profile(False)
    ctx.profile_only(None)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = W * H * iters / (ms_per_step * 1e-3) / 1e6
    bad = synth.bad_pixels(dmap.cpu().numpy(), pair)
    log(f"plain 1242x375: {ms_per_step:.2f} ms per depthMap, bad-1.0px {bad:.3f}%")

    taps = (2 * P + 1) ** 2
    avg_s = prop_ms * 1e-3 / prop_n
    flops = cands / prop_n * taps * FLOPS_PER_TAP
    flops_lb = W * H / 2 * LOWER_BOUND_CANDIDATES * taps * FLOPS_PER_TAP
    achieved = flops / avg_s / 1e12
    roofline = {"bound": "valu", "kernel": propagate_kernel_name(ctx, P),
                "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": measured_traffic(),
                "traffic_source": TRAFFIC_FILE, "avg_launch_ms": round(avg_s * 1e3, 4),
                "flops_per_launch": flops,
                "candidates_per_pixel": round(cands / prop_n / (W * H / 2), 3),
                "frac_18_candidates": round(flops_lb / avg_s / 1e12 / PEAK_FP32_TFLOPS, 4),
                "peak_without_packed_fp32": UNPACKED_FP32_TFLOPS,
                "frac_without_packed_fp32": round(achieved / UNPACKED_FP32_TFLOPS, 4)}

    # per-kernel table from one untimed call with every launch bracketed
    ctx.profile(True)
    ctx.profile_reset()
    step()
    torch.cuda.synchronize()
    kernels = {}
    for k in ("propagate", "pd_steps", "consistency", "fill_state", "viewbuf", "init", "prep",
              "zero", "fill_depthmap"):
        ms, n = ctx.profile_read(k)
        kernels[k] = {"ms": round(ms, 4), "launches": n, "avg_us": round(1e3 * ms / max(n, 1), 2)}
    ctx.profile(False)

    copy_calibration(args)

    batched = None
    if not args.no_batched:
        batched = {"method": describe_method(BATCHED_METHOD),
                   "c2_1242x375": run_pair_ablation(args, L, R, P, dmax, iters, pair,
                                                    ms_per_step, bad)}
        if not args.no_pairs:
            batched["c4_64_uneven_pairs"] = run_pairs(args, "c4")
            batched["c5_8x2880x1988"] = run_pairs(args, "c5")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pair, P, dmax)

    adjoint = None
    if not args.no_adjoint:
        adjoint = run_adjoint(L, R, P, dmax, pair)

    multiview = None
    if not args.no_multiview:
        multiview = run_multiview(args, cpu=not args.no_cpu)

    variants = None
    if not args.no_variants:
        variants = run_variants(args, L, R, P, dmax, iters, pair, cpu=not args.no_cpu)

    out = {
        "metric": METRIC,
        "value": round(value, 4), "unit": "Mpix*iter/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded slanted-plane Voronoi scene, band-limited texture)",
        "config": {"workload": "plain GPU PatchMatch, 1242x375 pair, 16 planes, disparity 0-192, "
                               "noise 0.01, window 35x35, 4 iterations red-black, L/R tol 1 px",
                   "width": W, "height": H, "patch_half": P, "iterations": iters,
                   "final_smoothing": "not run (its output is never read; the disparity map "
                                      "is the reference sequence's)",
                   "bad_1px_percent": round(bad, 3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "multiview": multiview,
        "adjoint": adjoint,
        "variants": variants,
        "kernels_one_call": kernels,
    }
    print(json.dumps(out))
    return 0


def time_batched(pairs_dev, P, method, reps=1):
    """Seconds of one pmh_batch_depthmap call with `method` over `pairs_dev`
    (two untimed calls first: the first tunes and runs eagerly, the second
    captures the graph), and the results of the last timed call."""
    import torch

    from cvt_amd import _lib, ops
    o = method_options(_lib, method)
    for _ in range(2):
        ops.batch_depthmap(pairs_dev, P, method["iterations"], o)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        res = ops.batch_depthmap(pairs_dev, P, method["iterations"], o)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps, res


def run_pair_ablation(args, L, R, P, dmax, iters, pair, plain_ms, plain_bad):
    """The ablation rows and the batched method on the plain workload's pair."""
    from cvt_amd import synth
    rows = []
    for what, method in ablation(iters):
        sec, res = time_batched([(L, R, dmax, 1.0)], P, method, reps=args.steps)
        b = synth.bad_pixels(res[0]["dmap"].cpu().numpy(), pair)
        rows.append({"row": what, "ms_per_pair": round(sec * 1e3, 3),
                     "bad_1px_percent": round(b, 3),
                     "speedup": round(plain_ms / (sec * 1e3), 3),
                     "equal_or_lower_error": round(b, 3) <= round(plain_bad, 3)})
        log(f"1242x375 {what}: {sec * 1e3:.2f} ms, bad-1.0px {b:.3f}%")
    fixed = rows[-1]
    return {"plain_ms_per_pair": round(plain_ms, 3), "plain_bad_1px_percent": round(plain_bad, 3),
            "ms_per_pair": fixed["ms_per_pair"], "bad_1px_percent": fixed["bad_1px_percent"],
            "speedup_mpix_per_s": fixed["speedup"],
            "equal_or_lower_error": fixed["equal_or_lower_error"], "ablation": rows}


def run_pairs(args, which):
    """Plain (one depthMap call per pair) against the batched method (one
    pmh_batch_depthmap call for all pairs), one timed repetition each."""
    import torch

    from cvt_amd import _lib, ops, synth

    P = 17
    t0 = time.perf_counter()
    if which == "c4":
        pairs = synth.uneven_batch(64, seed=4)
        plain_iters = 4
        desc = ("64 uneven pairs (W 320-1920, H 240-1080, disparity 32-256, 4-24 planes, "
                "noise 0.01), window 35x35, plain 4 iterations")
    else:
        pairs = synth.large_batch(8, seed=5)
        plain_iters = 8
        desc = ("8 pairs of 2880x1988 (32 planes, disparity 0-400, noise 0.02), window 35x35, "
                "plain 8 iterations")
    log(f"{which}: made {len(pairs)} pairs in {time.perf_counter() - t0:.1f} s")
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
    # warm-up: two iterations of every pair tune every shape class (the
    # propagate keys tell the first iteration from the later ones and do not
    # depend on the iteration count), so no tuning is timed
    for l, r, d, s in dev_pairs:
        ops.depthmap(l, r, P, d, 2, 4, s, popts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pmaps = [ops.depthmap(l, r, P, d, plain_iters, 4, s, popts)["dmap"]
             for l, r, d, s in dev_pairs]
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    plain_bad = bad_all(pmaps)
    del pmaps
    log(f"{which} plain: {plain_s:.2f} s, bad-1.0px {plain_bad:.3f}%")
    rows = [("the batched method", dict(BATCHED_METHOD))]
    if args.ablation_pairs:
        rows = ablation(plain_iters)
    out_rows = []
    for what, method in rows:
        sec, res = time_batched(dev_pairs, P, method)
        b = bad_all([x["dmap"] for x in res])
        del res
        out_rows.append({"row": what, "seconds": round(sec, 3), "mpix_per_s": round(mpx / sec, 3),
                         "bad_1px_percent": round(b, 3), "speedup": round(plain_s / sec, 3),
                         "equal_or_lower_error": round(b, 3) <= round(plain_bad, 3)})
        log(f"{which} {what}: {sec:.2f} s, bad-1.0px {b:.3f}%")
    fixed = out_rows[-1]
    res = {"workload": desc, "pairs": len(pairs), "mpix": round(mpx, 3),
           "plain_seconds": round(plain_s, 3), "plain_mpix_per_s": round(mpx / plain_s, 3),
           "plain_mpix_iter_per_s": round(mpx * plain_iters / plain_s, 3),
           "plain_bad_1px_percent": round(plain_bad, 3),
           "seconds": fixed["seconds"], "mpix_per_s": fixed["mpix_per_s"],
           "bad_1px_percent": fixed["bad_1px_percent"], "speedup": fixed["speedup"],
           "equal_or_lower_error": fixed["equal_or_lower_error"]}
    if args.ablation_pairs:
        res["ablation"] = out_rows
    del dev_pairs
    torch.cuda.empty_cache()
    return res


def run_adjoint(L, R, P, dmax, pair, reps=5):
    """The differentiation workload on the resident 1242x375 pair: the plane
    field of a 4-iteration plain run, then per repetition the window cost
    (pmh_cost), its adjoint with respect to the planes (pmh_cost_backward
    without images) and with respect to the planes and both images (dL/dI and
    dL/dG of each view); L = sum(cost).  Times are
    averages over `reps` synchronized calls after one warm-up call each."""
    import torch

    from cvt_amd import _lib, ops

    opts = _lib.default_options(sweep=_lib.SWEEP_REDBLACK, maxdispdiff=1.0)
    st = ops.depthmap(L, R, P, dmax, 4, 4, 1.0, opts, states=True)["left_states"]
    GL, GR = ops.gradxy(L), ops.gradxy(R)
    H, W = L.shape
    gout = torch.ones((H, W), dtype=torch.float32, device=L.device)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / reps

    fwd = timed(lambda: ops.cost(L, GL, R, GR, st, P))
    bwd_planes = timed(lambda: ops.cost_backward(L, GL, R, GR, st, gout, P, images=False))
    bwd_images = timed(lambda: ops.cost_backward(L, GL, R, GR, st, gout, P, images=True))
    mpx = W * H / 1e6
    taps = W * H * (2 * P + 1) ** 2
    log(f"adjoint 1242x375: cost {fwd:.2f} ms, backward (planes) {bwd_planes:.2f} ms, "
        f"backward (planes + images) {bwd_images:.2f} ms")
    return {"workload": "1242x375 pair, window 35x35, the left view's plane field after 4 "
                        "plain iterations; L = sum of the window costs",
            "cost_ms": round(fwd, 3), "backward_planes_ms": round(bwd_planes, 3),
            "backward_planes_images_ms": round(bwd_images, 3),
            "cost_mpix_per_s": round(mpx / (fwd * 1e-3), 1),
            "backward_planes_gtaps_per_s": round(taps / (bwd_planes * 1e-3) / 1e9, 1),
            "backward_planes_images_gtaps_per_s": round(taps / (bwd_images * 1e-3) / 1e9, 1)}


def run_multiview(args, cpu=True):
    """The multi-view method (pmh_mv_depthmap) on MV_WORKLOAD: Mpix*iter/s of
    whole runs (scene preparation excluded, inputs resident in HBM), the
    propagate kernel's roofline from its launches timed live with HIP events,
    depth accuracy against the rendered truth, and the oracle on the same
    workload on this host's cores."""
    import numpy as np
    import torch

    from cvt_amd import _lib, ops, synth
    m = MV_WORKLOAD
    W, H, P, iters, nref = m["width"], m["height"], m["patch_half"], m["iterations"], m["refine"]
    s = synth.make_multiview(W, H, m["sources"], seed=m["seed"])

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()

    scene = ops.MVScene(dev(s.ref), s.K, s.T, [dev(im) for im in s.srcs], s.Ks, s.Ts)
    ctx = ops.context()

    def run():
        return scene.depthmap(P, s.dmin, s.dmax, iters, _lib.SWEEP_REDBLACK, nref)

    for _ in range(max(1, args.warmup)):
        run()
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_only("mv_propagate")
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        depth = run()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    prop_ms, prop_n = ctx.profile_read("mv_propagate")
    ctx.profile(False)
    ctx.profile_only(None)
    dn = depth.cpu().numpy()
    err5 = synth.depth_error_rate(dn, s, 0.05)
    err1 = synth.depth_error_rate(dn, s, 0.01)
    log(f"multi-view 1242x375: {sec * 1e3:.2f} ms per run, depth off by >5% on {err5:.3f}%")
    avg_s = prop_ms * 1e-3 / prop_n
    cands = 8 + 3 + 1 + nref
    taps = (2 * P + 1) ** 2
    flops = W * H / 2 * cands * taps * (MV_FLOPS_PER_TAP + MV_FLOPS_PER_TAP_VIEW * m["sources"])
    achieved = flops / avg_s / 1e12
    res = {"workload": (f"multi-view PatchMatch, reference + {m['sources']} source views of "
                        f"{W}x{H} (rendered slanted-plane scene, depth 3-14), window "
                        f"{2 * P + 1}x{2 * P + 1}, {iters} iterations red-black, {nref} "
                        "randomized refinements per pass"),
           "value": round(W * H * iters / sec / 1e6, 4), "unit": "Mpix*iter/s",
           "ms_per_run": round(sec * 1e3, 3), "runs": args.steps,
           "depth_off_5pct_percent": round(err5, 3), "depth_off_1pct_percent": round(err1, 3),
           "roofline": {"bound": "valu", "kernel": f"mv_propagate_block_kernel<{m['sources']}>",
                        "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                        "traffic": measured_traffic(MV_TRAFFIC_FILE, "mv_propagate_kernel"),
                        "traffic_source": MV_TRAFFIC_FILE,
                        "avg_launch_ms": round(avg_s * 1e3, 4), "flops_per_launch": flops,
                        "candidates_per_pixel": cands,
                        "peak_without_packed_fp32": UNPACKED_FP32_TFLOPS,
                        "frac_without_packed_fp32": round(achieved / UNPACKED_FP32_TFLOPS, 4)}}
    if cpu:
        import oracle
        v = oracle.MVViews(s.ref, s.K, s.T, list(s.srcs), s.Ks, s.Ts)
        t0 = time.perf_counter()
        od, _ = oracle.mv_depthmap(v, P, s.dmin, s.dmax, iters, True, nref)
        c_s = time.perf_counter() - t0
        log(f"cpu multi-view 1242x375: {c_s:.1f} s")
        agree = float(np.mean(np.abs(dn - od) <= 0.01 * np.maximum(od, 1e-6)))
        res["cpu_baseline"] = {"value": round(W * H * iters / c_s / 1e6, 5),
                               "unit": "Mpix*iter/s", "cores": oracle.num_threads(),
                               "kind": "port", "sample": (f"the whole workload, {c_s:.1f} s on "
                                                          f"{cpu_model()}"),
                               "gpu_over_cpu": round(c_s / sec, 1),
                               "depth_within_1pct_of_oracle": round(agree, 4)}
    del scene
    torch.cuda.empty_cache()
    return res


def run_variants(args, L, R, P, dmax, iters, pair, cpu=True):
    """The two other input paths of the plain method on the 1242x375 scene,
    each timed like `value` (resident inputs, `args.steps` synchronized calls
    after the warm-up) with its propagate kernel's roofline from launches
    timed live with HIP events (candidates counted on the device):
      - depthMapInpaint (PMHuberStereo.cpp:293-507) on the gray pair: the
        reference's inpainting constants (occlusion map 0.5 px / 5 degrees,
        100 masked Huber-ROF steps, final check 1 px / 5 degrees), red-black;
      - depthMap on an RGBA pair of the same scene (three independent colour
        textures, alpha 1), red-black, L/R tolerance 1 px.
    With `cpu`, each also runs on the 450x375 pair (window 21x21, 3
    iterations, the reference's Jacobi sweep) on the GPU and in the oracle."""
    import numpy as np
    import torch

    from cvt_amd import _lib, ops, synth
    ctx = ops.context()
    W, H = pair.width, pair.height
    cp = synth.make_pair(W, H, 16, dmax, 0.01, seed=2, colour=True)
    L4, R4 = torch.from_numpy(cp.left).cuda(), torch.from_numpy(cp.right).cuda()
    iopts = _lib.default_inpaint_options(sweep=_lib.SWEEP_REDBLACK)
    ropts = _lib.default_options(sweep=_lib.SWEEP_REDBLACK, maxdispdiff=1.0)
    legs = {
        "inpaint_1242x375": ("depthMapInpaint, 1242x375 gray pair, window 35x35, 4 iterations "
                             "red-black, occlusion map 0.5 px / 5 deg, 100 masked Huber-ROF "
                             "steps", pair, FLOPS_PER_TAP,
                             lambda: ops.depthmap_inpaint(L, R, P, dmax, iters, 4, iopts)["dmap"]),
        "rgba_1242x375": ("depthMap, 1242x375 RGBA pair (same scene, colour texture), window "
                          "35x35, 4 iterations red-black, L/R tol 1 px", cp, RGBA_FLOPS_PER_TAP,
                          lambda: ops.depthmap(L4, R4, P, dmax, iters, 4, 1.0, ropts)["dmap"]),
    }
    out = {}
    taps = (2 * P + 1) ** 2
    for key, (desc, pr, fpt, fn) in legs.items():
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        ctx.profile(True)
        ctx.profile_only("propagate")
        ctx.profile_reset()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            d = fn()
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / args.steps
        prop_ms, prop_n = ctx.profile_read("propagate")
        cands = ctx.profile_candidates()
        ctx.profile(False)
        ctx.profile_only(None)
        avg_s = prop_ms * 1e-3 / prop_n
        flops = cands / prop_n * taps * fpt
        achieved = flops / avg_s / 1e12
        bad = synth.bad_pixels(d.cpu().numpy(), pr)
        log(f"{key}: {sec * 1e3:.2f} ms per call, bad-1.0px {bad:.3f}%, propagate "
            f"{avg_s * 1e3:.3f} ms per launch")
        out[key] = {"workload": desc, "value": round(W * H * iters / sec / 1e6, 4),
                    "unit": "Mpix*iter/s", "ms_per_call": round(sec * 1e3, 3),
                    "bad_1px_percent": round(bad, 3),
                    "roofline": {"bound": "valu",
                                 "kernel": propagate_kernel_name(ctx, P, 4 if pr is cp else 1),
                                 "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                                 "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                                 "avg_launch_ms": round(avg_s * 1e3, 4),
                                 "flops_per_tap": fpt,
                                 "candidates_per_pixel": round(cands / prop_n / (W * H / 2), 3),
                                 "frac_without_packed_fp32":
                                     round(achieved / UNPACKED_FP32_TFLOPS, 4)}}
    del L4, R4
    if cpu:
        import oracle
        P1, d1, it1 = 10, 64.0, 3
        g1 = synth.pair_450x375()
        c1 = synth.make_pair(450, 375, 8, d1, 0.01, seed=1, colour=True)
        small = {
            "inpaint_1242x375": (g1, lambda l, r: ops.depthmap_inpaint(
                l, r, P1, d1, it1, 4, _lib.default_inpaint_options())["dmap"],
                lambda: oracle.depthmap_inpaint(g1.left, g1.right, P1, d1, it1)["dmap"]),
            "rgba_1242x375": (c1, lambda l, r: ops.depthmap(
                l, r, P1, d1, it1, 4, 1.0, _lib.default_options())["dmap"],
                lambda: oracle.depthmap(c1.left, c1.right, P1, d1, it1, 1.0)["dmap"]),
        }
        for key, (pr, gfn, cfn) in small.items():
            l1, r1 = torch.from_numpy(pr.left).cuda(), torch.from_numpy(pr.right).cuda()
            gfn(l1, r1)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                gd = gfn(l1, r1)
            torch.cuda.synchronize()
            g_s = (time.perf_counter() - t0) / 5
            t0 = time.perf_counter()
            od = cfn()
            c_s = time.perf_counter() - t0
            agree = float(np.mean(np.abs(gd.cpu().numpy() - od) <= 0.5))
            log(f"cpu {key} on 450x375: {c_s:.1f} s (GPU {g_s * 1e3:.2f} ms)")
            mpi = 450 * 375 * it1 / 1e6
            out[key]["cpu_baseline"] = {
                "value": round(mpi / c_s, 5), "unit": "Mpix*iter/s",
                "cores": oracle.num_threads(), "kind": "port",
                "sample": (f"the same method on the 450x375 pair (window 21x21, 3 iterations, "
                           f"the reference's Jacobi sweep and tolerances), {c_s:.1f} s on "
                           f"{cpu_model()}"),
                "gpu_seconds_same_sample": round(g_s, 5), "gpu_over_cpu": round(c_s / g_s, 1),
                "within_half_px_of_oracle": round(agree, 4)}
    torch.cuda.empty_cache()
    return out


def cpu_baseline(pair, P, dmax):
    """The oracle on this host's cores: the whole 1242x375 pair for 1 iteration
    (`value`'s workload; initialisation and pre-processing included, not
    amortised), and the 450x375 pair in full beside the GPU on that pair."""
    import numpy as np
    import torch

    import oracle
    from cvt_amd import _lib, ops, synth
    W, H = pair.width, pair.height
    t0 = time.perf_counter()
    oracle.depthmap(pair.left, pair.right, P, dmax, 1, 1.0, maxdispdiff=1.0, sweep=1)
    c2_s = time.perf_counter() - t0
    log(f"cpu 1242x375 1 iteration: {c2_s:.1f} s")
    c1 = synth.pair_450x375()
    P1, d1, it1 = 10, 64.0, 3
    t0 = time.perf_counter()
    od = oracle.depthmap(c1.left, c1.right, P1, d1, it1, 1.0)["dmap"]
    c1_cpu = time.perf_counter() - t0
    log(f"cpu 450x375 3 iterations: {c1_cpu:.1f} s")
    L1 = torch.from_numpy(c1.left).cuda()
    R1 = torch.from_numpy(c1.right).cuda()
    o1 = _lib.default_options()  # the reference's Jacobi sweep and 0.5 px tolerance
    ops.depthmap(L1, R1, P1, d1, it1, 4, 1.0, o1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        gd = ops.depthmap(L1, R1, P1, d1, it1, 4, 1.0, o1)["dmap"]
    torch.cuda.synchronize()
    c1_gpu = (time.perf_counter() - t0) / 5
    agree = float(np.mean(np.abs(gd.cpu().numpy() - od) <= 0.5))
    mpi1 = c1.left.size * it1 / 1e6
    return {"value": round(W * H / c2_s / 1e6, 5), "unit": "Mpix*iter/s",
            "cores": oracle.num_threads(), "kind": "port",
            "sample": (f"the whole {W}x{H} pair, 1 iteration, red-black, window 35x35, full "
                       f"pipeline including initialisation and pre-processing (init, "
                       f"propagate, consistency, fill, 250-step Huber-ROF x2); {c2_s:.1f} s on "
                       f"{cpu_model()}"),
            "c1_450x375_full": {"workload": "450x375 pair, 8 planes, disparity 0-64, window "
                                           "21x21, 3 iterations, Jacobi sweep, L/R tol 0.5 px",
                                "cpu_seconds": round(c1_cpu, 3),
                                "cpu_mpix_iter_per_s": round(mpi1 / c1_cpu, 5),
                                "gpu_seconds": round(c1_gpu, 5),
                                "gpu_mpix_iter_per_s": round(mpi1 / c1_gpu, 4),
                                "gpu_over_cpu": round(c1_cpu / c1_gpu, 1),
                                "within_half_px_of_oracle": round(agree, 4)}}


if __name__ == "__main__":
    sys.exit(main())
