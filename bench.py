#!/usr/bin/env python3
"""Training-step throughput of the MI355X 3DGS trainer on one GPU.

Workload: 1M gaussians, SH degree 3, fp32, 1280x720, one view per step
(fwd + bwd + SSIM + Adam, densification every 100 steps), targets rendered
from a seeded 1M-gaussian ground-truth scene with 64 orbit cameras, training
from an independently seeded 1M-gaussian scene.  Synthetic data, inputs
resident in HBM before the timed region.

Prints one JSON line: the metric, per-kernel HIP-event times over the timed
region, the roofline figure of the dominant kernel and of the whole step, and
a CPU baseline (the oracle's restatement of the same step on the host).
"""
import argparse
import importlib.util
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))


def _load(name, rel):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, rel))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


HBM_PEAK = 8.0e12
SH_COEFFS = {0: 0, 1: 3, 2: 8, 3: 15}


def kernel_bytes(name, n_slots, n_visible, pairs, pixels, coeffs):
    """Algorithmic HBM bytes of one launch of each kernel family (fp32), as
    counted in the design document's roofline table."""
    sh = 12 * coeffs
    if name == "project":
        return n_slots * (64 + 112) + n_visible * sh
    if name == "forward":
        return pairs * 52 + pixels * 20
    if name == "backward":
        return pairs * 52 + pixels * 20 + n_visible * 36
    if name == "ssim":
        return pixels * (16 + 4 + 36)
    if name == "loss_grad":
        return pixels * (36 + 16 + 4 + 12)
    if name == "adam":
        return n_slots * (3 * 64 * 2 + 3 * sh * 2 + 36 + 4 + 32)
    if name == "emit":
        return n_visible * (48 + 16 + 8) + pairs * 8
    if name in ("tile_sort",):
        return pairs * 8 * 2
    if name == "depth_sort":
        return n_slots * 8 * 2
    if name == "scan":
        return n_slots * 12
    if name == "ranges":
        return pairs * 8
    return None


def step_bytes(n, pairs, pixels, degree):
    """Whole-step algorithmic bytes: B_G N + 132 P + 76 W H, B_G = 1832 B at
    SH degree 3 (572 B at degree 0)."""
    D = 4 * (11 + 3 * (degree + 1) ** 2)
    BG = D + 44 + 48 + 36 + 6 * D + 36 + 16
    return BG * n + 132 * pairs + 76 * pixels


def cpu_baseline(gs, orc, params, sh, cam, target, degree, budget_s):
    """The oracle's CPU restatement of one training step (project, bin,
    composite, loss + SSIM gradient, tile backward, projection adjoint, Adam),
    single-threaded, on the same 1M-gaussian 1280x720 view."""
    w, h = cam.width, cam.height
    t0 = time.perf_counter()
    rec, foot, tiles, keys = orc.project32(params, sh, cam, degree)
    pt, pg, ranges = orc.bin32(rec, foot, tiles, keys, w, h)
    img, last = orc.forward32(rec, pg, ranges, w, h)
    x = img[:, :3].astype(np.float64) + img[:, 3:4].astype(np.float64) * 0.0
    y = target.reshape(-1, 4)[:, :3] / 255.0
    _, dldc = orc.loss_image64(x, y, w, h, 0.2, grad=True)
    d4 = np.zeros((w * h, 4), np.float32)
    d4[:, :3] = dldc
    g2 = orc.backward32(rec, pg, ranges, img, last, d4, (0.0, 0.0, 0.0), w, h)
    gp, gsh = orc.adjoint64(params, sh, cam, degree, g2)
    lr = np.full(params.shape, 5e-3, np.float32)
    orc.adam(params.reshape(-1), gp.astype(np.float32).reshape(-1), np.zeros(params.size, np.float32),
             np.zeros(params.size, np.float32), lr.reshape(-1))
    if gsh.size:
        orc.adam(sh.reshape(-1), gsh.astype(np.float32).reshape(-1), np.zeros(sh.size, np.float32),
                 np.zeros(sh.size, np.float32), np.full(sh.size, 5e-3, np.float32))
    dt = time.perf_counter() - t0
    return dt, int(len(pt))


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--gaussians", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=720)
    ap.add_argument("--sh-degree", type=int, default=3)
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--densify-interval", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build (multi-GPU "
              "sharding is a design note, not an implementation)", file=sys.stderr)
        sys.exit(2)

    gs = _load("gs_hip", "on-divice-gsplate_amd/gs_hip.py")
    device = gs.init(0)
    log(f"device {device}")
    N, W, H, deg = args.gaussians, args.width, args.height, args.sh_degree
    coeffs = SH_COEFFS[deg]
    cams = gs.orbit_cameras(args.views, W, H)

    log(f"ground truth: {N} gaussians, {args.views} views at {W}x{H}")
    gt_params, gt_sh = gs.synthetic_scene(N, deg, 1001)
    targets = []
    with gs.Workspace(N, deg, W, H) as gt:
        gt.load(gt_params, gt_sh)
        for cam in cams:
            img = gt.render(gs.make_frame(cam, N, deg))
            img[..., 3] = 255
            targets.append(img)
    target_bufs = [gs.DeviceBuffer(t) for t in targets]
    del gt_params, gt_sh

    log("training scene")
    params, sh = gs.synthetic_scene(N, deg, 2002)
    capacity = 2 * N
    ws = gs.Workspace(capacity, deg, W, H)
    ws.load(params, sh)
    tp = gs.TrainParams()
    order = np.random.default_rng(3003).permutation(args.views)
    state = {"slots": N, "densified": 0, "pairs": 0}

    def step(t):
        v = int(order[t % args.views])
        ws.set_target_device(target_bufs[v].ptr)
        frame = gs.make_frame(cams[v], state["slots"], deg, step=t, tp=tp)
        pairs, _ = ws.train_step(frame, want_losses=False)
        state["pairs"] += pairs
        if args.densify_interval > 0 and (t + 1) % args.densify_interval == 0:
            state["slots"], _ = ws.densify(frame)
            state["densified"] += 1
        return pairs

    for t in range(args.warmup):
        step(t)
    gs.synchronize()
    log(f"warmup done ({args.warmup} steps); timing {args.steps} steps")
    ws.timings(enable=True)
    state["pairs"] = 0
    slots_before = state["slots"]
    t0 = time.perf_counter()
    for t in range(args.warmup, args.warmup + args.steps):
        step(t)
        if (t - args.warmup + 1) % 50 == 0:
            log(f"step {t - args.warmup + 1}/{args.steps}")
    gs.synchronize()
    elapsed = time.perf_counter() - t0
    events = ws.timings(enable=False)
    iters = args.steps / elapsed
    mean_pairs = state["pairs"] / args.steps
    slots_mean = 0.5 * (slots_before + state["slots"])

    per_kernel = {}
    for name, ms in events:
        d = per_kernel.setdefault(name, [0.0, 0])
        d[0] += ms
        d[1] += 1
    kernels = {k: {"ms_avg": v[0] / v[1], "launches": v[1], "ms_per_step": v[0] / args.steps}
               for k, v in per_kernel.items()}
    if any(v["launches"] < args.steps for k, v in kernels.items() if k != "densify"):
        log("warning: fewer timed launches than steps for some kernels")
    tiles_touched = ws.download(gs.BUF_TILES, np.uint32, state["slots"])
    n_visible = int((tiles_touched > 0).sum())
    for k, v in kernels.items():
        b = kernel_bytes(k, slots_mean, n_visible, mean_pairs, W * H, coeffs)
        if b is not None:
            v["algorithmic_bytes"] = int(b)
            v["GBps"] = b / (v["ms_avg"] * 1e-3) / 1e9
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_step"]) if kernels else None
    roofline = None
    if dominant is not None and "GBps" in kernels[dominant]:
        ach = kernels[dominant]["GBps"]
        roofline = {"kernel": dominant, "bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK / 1e9,
                    "unit": "GB/s", "frac": round(ach * 1e9 / HBM_PEAK, 5), "traffic": None,
                    "ms_avg": round(kernels[dominant]["ms_avg"], 4)}
    sb = step_bytes(slots_mean, mean_pairs, W * H, deg)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline: the oracle's single-threaded step on one view")
        orc = _load("gs_oracle", "oracle/oracle.py")
        p0, s0 = gs.synthetic_scene(N, deg, 2002)
        dt, cpu_pairs = cpu_baseline(gs, orc, p0, s0, cams[int(order[0])], targets[int(order[0])], deg, 30.0)
        cpu = {"value": round(1.0 / dt, 5), "unit": "iter/s", "cores": 1, "kind": "port",
               "sample": f"one full training step (1 view, {N} gaussians, SH{deg}, {W}x{H}, {cpu_pairs} pairs) "
                         f"through the oracle's C restatement in {dt:.1f} s on {platform.processor() or 'host'} "
                         f"({os.cpu_count()} cores present, 1 used)"}

    out = {
        "metric": "train iter/s (fwd+bwd+SSIM+Adam) @1M Gaussians 1280x720",
        "value": round(iters, 3),
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 / iters, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded 1M-gaussian GT scene, 64 orbit views rendered by this trainer)",
        "config": {"workload": "plain single-view training step, 1M gaussians SH3 fp32, 1280x720, 16x16 tiles, "
                               "0.8 L1 + 0.2 D-SSIM, Adam, densify every 100 steps",
                   "gaussians_start": N, "gaussians_end": state["slots"], "width": W, "height": H,
                   "sh_degree": deg, "mean_pairs": int(mean_pairs), "densify_events": state["densified"],
                   "device": device, "mode": "plain"},
        "mpix_per_s": round(iters * W * H / 1e6, 2),
        "step_hbm": {"bytes_per_iter": int(sb), "achieved_GBps": round(sb * iters / 1e9, 2),
                     "frac": round(sb * iters / HBM_PEAK, 5)},
        "roofline": roofline,
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_step"])},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    for b in target_bufs:
        b.free()
    ws.close()


if __name__ == "__main__":
    main()
