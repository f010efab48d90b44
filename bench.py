#!/usr/bin/env python3
"""Training-step throughput of the MI355X 3DGS trainer on one GPU.

Workload: 1M gaussians, SH degree 3, fp32, 1280x720, fwd + bwd + SSIM + Adam
with densification every 100 view-iterations; targets rendered from a seeded
1M-gaussian ground-truth scene at 64 seeded jittered cameras (radius 2.5-3.5,
height -0.5-1.5), training from an independently seeded 1M-gaussian scene.
Measured twice on the same views: plain (one view per optimizer step) and
batched (8 views per optimizer step: every stage of all 8 views in one launch,
one Adam on the mean gradient).  Each mode's launch shapes are picked first by
the library's autotuner on that mode's first views.  Synthetic data, inputs
resident in HBM before the timed region.

Prints one JSON line: the metric, per-kernel HIP-event times over the timed
region, the roofline figure of the dominant kernel and of the whole step, the
held-out PSNR of plain and batched training at equal view-iterations (a
curve from a common start), and a CPU baseline (the oracle's restatement of
the same step on the host).
"""
import argparse
import importlib.util
import json
import math
import os
import platform
import re
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


def kernel_bytes(name, n_slots, n_visible, pairs, pixels, coeffs, views):
    """Algorithmic HBM bytes of one launch of each kernel family (fp32) over
    `views` views, from the per-unit figures of the step's byte model
    (B_G per gaussian, B_P = 132 per pair, B_px = 76 per pixel): projection
    D + 44 per gaussian (parameters in; record and count out); emission 48
    per visible gaussian (re-read + offset) + 12 per pair (key + value out);
    one tile-sort pass 24 per pair (read + write); ranges 8 per pair; forward
    44 per pair (index + 40 B record) + 28 per pixel; SSIM 16 per pixel
    (render + target in) and the loss gradient 12 (dL/dC out); backward 44
    per pair + 20 per pixel + 36 per visible gaussian (the 2D gradient sums);
    adjoint + Adam on the 16-float slots 6 x 64 + 32 + 48 (state,
    statistics, the three SH groups' direction terms) per slot and 44 per
    slot per view (tiles, 2D gradient, misc); SH rest Adam 6 x 12 coeffs + 64
    per slot (state, position, direction terms) and 16 per slot per view.  `pairs` and `n_visible` are per view; D = 4 (11 + 3 K) bytes of
    parameters per gaussian with K = coeffs + 1 SH coefficients per channel."""
    D = 4 * (11 + 3 * (coeffs + 1))
    if name == "project":
        return views * n_slots * (D + 44)
    if name == "emit":
        return views * (n_visible * 48 + pairs * 12)
    if name == "tile_sort":
        return views * pairs * 24
    if name == "ranges":
        return views * pairs * 8
    if name == "forward":
        return views * (pairs * 44 + pixels * 28)
    if name == "ssim":
        return views * pixels * 16
    if name == "loss_grad":
        return views * pixels * 12
    if name == "backward":
        return views * (pairs * 44 + pixels * 20 + n_visible * 36)
    if name == "adam":  # 16-float state read + written, statistics, SH direction terms; per view tiles, GRAD2D, misc
        return n_slots * (6 * 64 + 32 + (48 if coeffs else 0)) + views * n_slots * 44
    if name == "adam_sh":  # SH rest state read + written, position, direction terms out; per view colour gradient, misc
        return n_slots * (6 * 12 * coeffs + 16 + 48) + views * n_slots * 16
    if name == "depth_sort":  # one pass over (key, value): not in the step model, which counts the pairs' sort
        return views * n_slots * 16
    if name == "scan":
        return views * n_slots * 8
    return None


def step_bytes(n, pairs, pixels, degree):
    """Whole-step algorithmic bytes: B_G N + 132 P + 76 W H, B_G = 1832 B at
    SH degree 3 (572 B at degree 0)."""
    D = 4 * (11 + 3 * (degree + 1) ** 2)
    BG = D + 44 + 48 + 36 + 6 * D + 36 + 16
    return BG * n + 132 * pairs + 76 * pixels


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "host"


def cpu_threads():
    """The OpenMP threads the oracle runs on: OMP_NUM_THREADS if set, else every host core."""
    try:
        return int(os.environ.get("OMP_NUM_THREADS", "")) or os.cpu_count()
    except ValueError:
        return os.cpu_count()


def cpu_baseline(gs, orc, params, sh, cams, targets, order, degree, steps):
    """The oracle's CPU restatement of the training step (fp32 projection,
    binning, compositing and tile backward; fp64 loss, SSIM and projection
    adjoint; fp32 Adam), multithreaded over the host cores, for `steps` plain
    steps of the benchmark's own scene and views.  Returns (seconds per step, mean pairs)."""
    n = params.shape[0]
    lr = orc.adam_rates(n, 1.6e-4, 5e-3, 5e-3, 5e-3, 5e-3)
    cams_o = [cams[int(v)] for v in order[:steps]]
    tgts_o = [targets[int(v)] for v in order[:steps]]
    t0 = time.perf_counter()
    orc.train(params, sh, degree, cams_o, tgts_o, steps, lr, 5e-3, 0.2)
    return (time.perf_counter() - t0) / steps


def cpu_small_scene(gs, orc):
    """The small-scene configuration run in full on the CPU restatement
    (tests/golden/make_c1_trajectory.py): 20k gaussians SH0, 24 orbit views at
    360x640, 300 steps.  Returns (seconds, PSNR before, PSNR after)."""
    W, H, N = 360, 640, 20_000
    cams = gs.orbit_cameras(24, W, H)
    gt_p, gt_sh = gs.synthetic_scene(N, 0, 1001)
    targets = orc.render_targets64(gt_p, gt_sh, 0, cams)
    p0, s0 = gs.synthetic_scene(N, 0, 2002)
    lr = orc.adam_rates(N, 1.6e-4, 5e-3, 5e-3, 5e-3, 5e-3)
    t0 = time.perf_counter()
    p, s = orc.train(p0, s0, 0, cams, targets, 300, lr, 5e-3, 0.2)
    dt = time.perf_counter() - t0
    return dt, orc.mean_psnr(p0, s0, 0, cams, targets), orc.mean_psnr(p, s, 0, cams, targets)


def run(gs, ws, cams, target_bufs, order, per_step, views_total, warmup_views, deg, tp, densify_every, state,
        timed):
    """Trains for `views_total` view-iterations, `per_step` views per
    optimizer step (1: the plain single-view step).  Returns elapsed seconds."""
    steps = views_total // per_step
    t0 = None
    for s in range(-(warmup_views // per_step), steps):
        if s == 0:
            gs.synchronize()
            if timed:
                ws.timings(enable=True)
            state["pairs"] = 0
            state["slots_start"] = state["slots"]
            t0 = time.perf_counter()
        base = state["views"]
        vs = [int(order[(base + k) % len(order)]) for k in range(per_step)]
        # The step's frames differ only in their camera: one is built, the others copied.
        f0 = gs.make_frame(cams[vs[0]], state["slots"], deg, step=state["step"], tp=tp, flags=state["flags"],
                           batch_scaling=state["batch_scaling"])
        frames = [f0]
        for v in vs[1:]:
            f = gs.Frame.from_buffer_copy(f0)
            f.cam = cams[v]
            frames.append(f)
        if per_step == 1:
            ws.set_target_device(target_bufs[vs[0]].ptr)
            pairs, _ = ws.train_step(frames[0], want_losses=False)
            pairs = [pairs]
        else:
            pairs, _ = ws.train_views(frames, [target_bufs[v].ptr for v in vs], want_losses=False)
        if s >= 0:
            state["pairs"] += sum(pairs)
        state["views"] += per_step
        state["step"] += 1
        if densify_every > 0 and state["views"] // densify_every != base // densify_every:
            state["slots"], _ = ws.densify(frames[0])
            state["densified"] += 1
        if s >= 0 and (s + 1) % max(1, steps // 4) == 0:
            log(f"  {per_step}-view steps: {s + 1}/{steps}")
    gs.synchronize()
    return time.perf_counter() - t0, steps


def mean_tile_pairs(gs, ws, cams, target_bufs, order, per_step, state, deg, tp):
    """(16x16 tile, gaussian) overlaps per view, the pair count of the byte
    model, whatever bin size the run used: every training camera once,
    NO_UPDATE steps with 16 px bins and no graph, then the run's launch shape
    back.  Also returns the pairs per view at the run's own bin size over the
    same cameras."""
    shape = ws.tuning()
    views = [int(v) for v in dict.fromkeys(int(v) for v in order)]
    counts = {}
    for shift in sorted({0, shape.bin_shift}):
        ws.set_tuning(graph=0, bin_shift=shift)
        total = 0
        for k in range(0, len(views), per_step):
            vs = views[k:k + per_step]
            frames = [gs.make_frame(cams[v], state["slots"], deg, step=state["step"], tp=tp,
                                    flags=state["flags"] | gs.FLAG_NO_UPDATE) for v in vs]
            if len(vs) == 1:
                ws.set_target_device(target_bufs[vs[0]].ptr)
                pairs, _ = ws.train_step(frames[0], want_losses=False)
                total += pairs
            else:
                pairs, _ = ws.train_views(frames, [target_bufs[v].ptr for v in vs], want_losses=False)
                total += sum(pairs)
        counts[shift] = total / len(views)
    ws.set_tuning(graph=shape.graph, bin_shift=shape.bin_shift)
    return counts[0], counts[shape.bin_shift]


TRAFFIC_PROFILE = os.path.join("profiles", "hbm_traffic_8views_1M_1280x720_pmc.md")
SQ_PROFILE = os.path.join("profiles", "sq_counters_8views_1M_1280x720_pmc.md")
VALU_PEAK = 1024 * 2.4e9 / 4  # wave64 VALU instructions per second: 1024 SIMDs, 4 cycles each


def profile_shape(text):
    """The launch shape a profile summary was taken at, from the `--shape G,D,T,GRAPH[,BIN[,STORED]]`
    of its command: (raster groups per CU, depth digit bits, tile digit bits, bin shift, stored
    state); graph replay leaves the kernels as they are and is not part of it."""
    m = re.search(r"--shape ([0-9]+),([0-9]+),([0-9]+),[0-9]+(?:,([0-9]+))?(?:,([0-9]+))?", text)
    if m is None:
        return None
    return int(m.group(1)), int(m.group(2)), int(m.group(3)), int(m.group(4) or 0), int(m.group(5) or 0)


def profile_row(path, kernel, columns, tuning):
    """The first `columns` numbers of `kernel`'s row in the profile summary at
    `path` when that summary was taken at the run's pair-bin size, else None.
    The rasterizers' work depends on the bin size; the sort digit widths and
    the grid mode (one workgroup per tile or persistent) leave it as it is."""
    try:
        text = open(os.path.join(ROOT, path)).read()
    except OSError:
        return None
    shape = profile_shape(text)
    # The stored-state backward runs with 16 px bins only (coarser bins re-walk).
    stored = int(bool(tuning and tuning.get("stored_state", 0)) and tuning["bin_shift"] == 0)
    if tuning is None or shape is None or shape[3] != tuning["bin_shift"] or int(shape[4] and shape[3] == 0) != stored:
        return None
    if kernel == "backward" and stored:
        kernel = "backward_stored"
    m = re.search(r"^\| k_%s \|((?: [0-9.e+]+ \|){%d})" % (re.escape(kernel), columns), text, re.M)
    return [float(x) for x in m.group(1).split("|")[:-1]] if m else None


def profiled_valu(kernel, tuning):
    """VALU instructions of one 8-view launch of `kernel` from the SQ counter
    summary of this workload at the run's launch shape, or None."""
    row = profile_row(SQ_PROFILE, kernel, 3, tuning)
    return row[2] if row else None


def profiled_traffic(kernel, tuning):
    """HBM bytes of one 8-view launch of `kernel` (calibrated FETCH_SIZE + WRITE_SIZE) from the
    rocprofv3 summary of this workload at the run's launch shape, or None."""
    row = profile_row(TRAFFIC_PROFILE, kernel, 4, tuning)
    return int(row[3] * 1e6) if row else None


RESOLUTIONS = ((640, 360), (1280, 720), (1920, 1080))


def multi_scene_workload(seed_base=5000, total=6_000_000, tol=0.10):
    """16 scenes, gaussian counts floor(exp(U[ln 5e4, ln 3e6])) and a
    resolution from 640x360 / 1280x720 / 1920x1080 each, drawn from the first
    seed >= seed_base whose counts total `total` within `tol`."""
    for seed in range(seed_base, seed_base + 10000):
        rng = np.random.default_rng(seed)
        counts = np.floor(np.exp(rng.uniform(math.log(5e4), math.log(3e6), 16))).astype(int)
        res = [RESOLUTIONS[int(i)] for i in rng.integers(0, 3, 16)]
        if abs(counts.sum() - total) <= tol * total:
            return seed, [int(c) for c in counts], res
    raise RuntimeError("no seed gives the multi-scene total")


def device_used_bytes():
    """Bytes of device memory in use (hipMemGetInfo: total - free)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, tot = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(tot)) != 0:
        return None
    return tot.value - free.value


def multi_scene(gs, args, device):
    """16 independent scenes (about 6M gaussians, SH3 fp32, mixed resolutions),
    2 views per scene per optimizer step: every stage of all 32 views as one
    launch through a gs_hip_batch, against the same scenes stepped one after
    another (each scene's own 2-view step).  Prints one JSON line."""
    deg = 3
    seed, counts, res = multi_scene_workload()
    log(f"multi-scene: seed {seed}, {sum(counts)} gaussians in 16 scenes, counts {counts}")
    used0 = device_used_bytes()
    tp = gs.TrainParams()
    scenes, targets, cams = [], [], []
    for s, (n, (w, h)) in enumerate(zip(counts, res)):
        cs = gs.jitter_cameras(8, w, h, 3003 + s)
        gt_p, gt_sh = gs.synthetic_scene(n, deg, 1001 + 100 * s)
        with gs.Workspace(n, deg, w, h) as gt:
            gt.load(gt_p, gt_sh)
            imgs = []
            for cam in cs:
                img = gt.render(gs.make_frame(cam, n, deg))
                img[..., 3] = 255
                imgs.append(gs.DeviceBuffer(img))
        del gt_p, gt_sh
        p, sh = gs.synthetic_scene(n, deg, 2002 + 100 * s)
        ws = gs.Workspace(n, deg, w, h)
        ws.set_views(2)
        ws.load(p, sh)
        scenes.append(ws)
        targets.append(imgs)
        cams.append(cs)
    gs.synchronize()

    def frames_of(step, flags=0):
        fps, tps = [], []
        for s, (n, cs, ts) in enumerate(zip(counts, cams, targets)):
            vs = [(2 * step) % 8, (2 * step + 1) % 8]
            fps.append([gs.make_frame(cs[v], n, deg, step=step, tp=tp, flags=flags | gs.FLAG_DEPTH16) for v in vs])
            tps.append([ts[v].ptr for v in vs])
        return fps, tps

    batch = gs.Batch(scenes)
    steps, warmup = args.steps, args.warmup
    pixels = sum(2 * w * h for (w, h) in res)
    step_no = [0]

    def choose_shape(stored):
        """The fastest of a few launch shapes, each timed on two untimed (NO_UPDATE) batch steps."""
        best, best_ms = None, None
        for shape in ((0, 8, 8, 0, 0), (0, 8, 6, 1, 0), (0, 6, 6, 1, 0), (2, 8, 6, 1, 0), (0, 6, 6, 1, 1),
                      (0, 6, 6, 1, 2)):
            if stored and shape[4] != 0:
                continue  # the stored-state backward runs with the 16 px tiles' own lists only
            batch.set_tuning(raster_groups_per_cu=shape[0], depth_digit_bits=shape[1], tile_digit_bits=shape[2],
                             graph=shape[3], bin_shift=shape[4], stored_state=stored)
            fps, tps = frames_of(0, flags=gs.FLAG_NO_UPDATE)
            batch.train(fps, tps, want_losses=False)
            gs.synchronize()
            t0 = time.perf_counter()
            for _ in range(2):
                batch.train(fps, tps, want_losses=False)
            gs.synchronize()
            ms = 1000 * (time.perf_counter() - t0) / 2
            if best_ms is None or ms < best_ms:
                best, best_ms = shape, ms
        batch.set_tuning(raster_groups_per_cu=best[0], depth_digit_bits=best[1], tile_digit_bits=best[2],
                         graph=best[3], bin_shift=best[4], stored_state=stored)
        log(f"multi-scene ({'stored-state' if stored else 'recompute'} backward): launch shape {best} "
            f"({best_ms:.2f} ms per step while choosing)")
        return best

    def timed_batch(label):
        """`warmup` + `steps` training steps of the batch; seconds per step, pairs, kernel events."""
        pairs_total = 0
        for s in range(-warmup, steps):
            if s == 0:
                gs.synchronize()
                batch.timings(enable=True)
                t0 = time.perf_counter()
            fps, tps = frames_of(step_no[0])
            step_no[0] += 1
            pairs, _ = batch.train(fps, tps, want_losses=False)
            if s >= 0:
                pairs_total += sum(pairs)
            if s >= 0 and (s + 1) % max(1, steps // 4) == 0:
                log(f"  {label} batch steps: {s + 1}/{steps}")
        gs.synchronize()
        return (time.perf_counter() - t0) / steps, pairs_total, batch.timings(enable=False)

    # The recompute backward (the configuration's), then the stored-state
    # backward on the same scenes: its per-view bucket states are allocated on
    # its first step, so the memory in use after each run is that variant's.
    best = choose_shape(0)
    batch_s, pairs_total, events = timed_batch("recompute")
    peak = device_used_bytes()
    best_stored = choose_shape(1)
    stored_s, _, stored_events = timed_batch("stored-state")
    stored_peak = device_used_bytes()
    # A stored-list backward would keep every pixel's contributors (gaussian index + alpha, 8 B each) from
    # the forward; the recompute backward keeps LAST (4 B) and T (in the image) per pixel.  Counted from
    # LAST after the stored-state run, whose 16 px tile lists make LAST a position among the tile's own
    # splats (in a coarse bin's list it would count the bin's other tiles' splats too).
    stored_list, recompute_state = 0, 0
    for ws, n in zip(scenes, counts):
        for v in range(2):
            last = ws.download(gs.BUF_LAST, np.uint32, view=v)
            stored_list += 8 * int(last.astype(np.int64).sum())
            recompute_state += 8 * last.size
    batch.close()
    # The same scenes, one scene's 2-view step after another (each scene's own launches).
    for ws in scenes:
        ws.set_tuning(raster_groups_per_cu=best[0], depth_digit_bits=best[1], tile_digit_bits=best[2], graph=best[3],
                      bin_shift=best[4], stored_state=0)
    for s in range(-warmup, steps):
        if s == 0:
            gs.synchronize()
            t0 = time.perf_counter()
        fps, tps = frames_of(steps + warmup + s + warmup)
        for ws, f, t in zip(scenes, fps, tps):
            ws.train_views(f, t, want_losses=False)
    gs.synchronize()
    seq_s = (time.perf_counter() - t0) / steps
    kern, stored_kern = {}, {}
    for name, ms in events:
        kern[name] = kern.get(name, 0.0) + ms / steps
    for name, ms in stored_events:
        stored_kern[name] = stored_kern.get(name, 0.0) + ms / steps
    views = 32
    out = {
        "metric": "multi-scene train iter/s (fwd+bwd+SSIM+Adam): 16 scenes, ~6M gaussians, 2 views per scene per step",
        "value": round(views / batch_s, 3),
        "unit": "iter/s (view-iterations)",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": round(1000 * batch_s, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": f"synthetic (16 seeded scenes, workload seed {seed}; 8 jittered views each rendered by this trainer)",
        "config": {"workload": "16 independent scenes trained concurrently: one launch per stage over all 32 views "
                               "(gs_hip_batch), each scene's Adam on the mean of its 2 views, recompute backward",
                   "scenes": 16, "gaussians": counts, "gaussians_total": sum(counts),
                   "resolutions": [f"{w}x{h}" for (w, h) in res], "views_per_scene": 2, "sh_degree": deg,
                   "tuning": {"raster_groups_per_cu": best[0], "depth_digit_bits": best[1],
                              "tile_digit_bits": best[2], "graph": best[3], "bin_shift": best[4]}, "device": device},
        "sequential_scenes": {"iters_per_s": round(views / seq_s, 3), "ms_per_step": round(1000 * seq_s, 4)},
        "speedup_batch_vs_sequential": round(seq_s / batch_s, 3),
        "mpix_per_s": round(pixels / batch_s / 1e6, 2),
        "mean_pairs_per_view": int(pairs_total / (steps * views)),
        "memory": {"device_bytes_in_use": peak, "device_bytes_before": used0,
                   "recompute_per_pixel_state_bytes": recompute_state,
                   "stored_list_per_pixel_bytes_counted": stored_list},
        "stored_state": {
            "note": "the same batch with the stored-state backward (the forward stores each pixel's (C, T) every "
                    "64 splats of its tile's walk; one wave per 64-splat bucket sweeps the pixels front to back)",
            "iters_per_s": round(views / stored_s, 3), "ms_per_step": round(1000 * stored_s, 4),
            "recompute_speedup": round(stored_s / batch_s, 3),
            "device_bytes_in_use": stored_peak,
            "bucket_state_bytes": None if stored_peak is None or peak is None else stored_peak - peak,
            "tuning": {"raster_groups_per_cu": best_stored[0], "depth_digit_bits": best_stored[1],
                       "tile_digit_bits": best_stored[2], "graph": best_stored[3], "bin_shift": best_stored[4]},
            "kernels_ms_per_step": {k: round(v, 4) for k, v in sorted(stored_kern.items(), key=lambda kv: -kv[1])}},
        "kernels_ms_per_step": {k: round(v, 4) for k, v in sorted(kern.items(), key=lambda kv: -kv[1])},
    }
    print(json.dumps(out), flush=True)
    for ws in scenes:
        ws.close()
    for ts in targets:
        for t in ts:
            t.free()


def rnd(v, digits=4):
    return None if v is None else round(v, digits)


def heldout_psnr(gs, ws, cams, targets, slots, deg):
    """Mean PSNR (dB) over held-out views of the workspace's model (float
    render over black against the RGBA8 target's colour)."""
    out = []
    for cam, t in zip(cams, targets):
        ws.render(gs.make_frame(cam, slots, deg))
        img = ws.download(gs.BUF_IMAGE, np.float32).reshape(-1, 4)[:, :3].astype(np.float64)
        y = t.reshape(-1, 4)[:, :3] / 255.0
        out.append(-10.0 * math.log10(max(float(np.mean((img - y) ** 2)), 1e-12)))
    return float(np.mean(out))


def depth_key_difference(gs, params, sh, cam, target_buf, deg):
    """One view's step (model untouched) under the 16-bit log-depth pair
    order and under the float-bits order: the image and 2D-gradient
    differences the order makes (ties broken differently between splats of
    equal 16-bit key)."""
    n = params.shape[0]
    out = {}
    with gs.Workspace(n, deg, cam.width, cam.height) as ws:
        ws.load(params, sh)
        for name, flags in (("depth16", gs.FLAG_DEPTH16), ("float", 0)):
            ws.train_views([gs.make_frame(cam, n, deg, flags=flags | gs.FLAG_NO_UPDATE)], [target_buf.ptr],
                           want_losses=False)
            out[name] = (ws.download(gs.BUF_IMAGE, np.float32).reshape(-1, 4),
                         ws.download(gs.BUF_GRAD2D, np.float32, 9 * n).reshape(n, 9))
    (i16, g16), (i32, g32) = out["depth16"], out["float"]
    scale = np.abs(g32).max(axis=0) + 1e-30
    return {"image_max_abs": float(np.abs(i16 - i32).max()), "image_mean_abs": float(np.abs(i16 - i32).mean()),
            "pixels_differing": int((np.abs(i16 - i32).max(axis=1) > 0).sum()),
            "grad2d_max_rel": float((np.abs(g16 - g32).max(axis=0) / scale).max())}


def quality_curve(gs, cams, target_bufs, orders, params, sh, held_cams, held_targets, deg, tp, flags, batch,
                  densify_every, checkpoints):
    """Held-out PSNR of plain (one view per optimizer step) and batched
    (`batch` views per step; Adam's hyperparameters as given, and scaled for
    the mean of `batch` views) training from the same start with the same
    densification schedule, at each of `checkpoints` view-iterations, with
    the training time (held-out renders excluded) to reach it; the mean and
    the range over the view sequences `orders` (a single run's curve moves by
    up to a dB when any rounding changes: training is chaotic).  Default
    launch shapes, eager steps, the deterministic backward (a rerun gives the
    same curve)."""
    n = params.shape[0]
    w, h = cams[0].width, cams[0].height
    curve = {"view_iterations": list(checkpoints), "orders": len(orders)}
    for mode, per_step, scaling in (("plain", 1, 1), ("batched", batch, 1), ("batched_lr_scaled", batch, batch)):
        runs_db, runs_s = [], []
        for order in orders:
            ws = gs.Workspace(2 * n, deg, w, h)
            ws.set_views(per_step)
            ws.load(params, sh)
            state = {"slots": n, "densified": 0, "pairs": 0, "views": 0, "step": 0,
                     "flags": flags | gs.FLAG_DETERMINISTIC, "batch_scaling": scaling}
            psnr, seconds, done, total = [], [], 0, 0.0
            for c in checkpoints:
                elapsed, _ = run(gs, ws, cams, target_bufs, order, per_step, c - done, 0, deg, tp, densify_every,
                                 state, timed=False)
                total += elapsed
                done = c
                psnr.append(heldout_psnr(gs, ws, held_cams, held_targets, state["slots"], deg))
                seconds.append(total)
            ws.close()
            runs_db.append(psnr)
            runs_s.append(seconds)
        db = np.asarray(runs_db)
        curve[mode + "_db"] = [round(float(x), 3) for x in db.mean(axis=0)]
        curve[mode + "_db_min"] = [round(float(x), 3) for x in db.min(axis=0)]
        curve[mode + "_db_max"] = [round(float(x), 3) for x in db.max(axis=0)]
        curve[mode + "_s"] = [round(float(x), 4) for x in np.asarray(runs_s).mean(axis=0)]
        log(f"quality: {mode} {curve[mode + '_db']} dB (range {curve[mode + '_db_min']} .. {curve[mode + '_db_max']}) "
            f"at {list(checkpoints)} view-iterations ({curve[mode + '_s'][-1]:.2f} s)")
    return curve


def small_scene_gpu(gs):
    """The small-scene configuration in full on the GPU (the CPU baseline's
    workload): 20k gaussians SH0, 24 orbit views at 360x640, 300 plain steps
    (one view per optimizer step, in view order), 0.8 L1 + 0.2 D-SSIM, Adam
    at 1.6e-4 / 5e-3, no densification; targets rendered on the GPU from the
    seeded ground truth.  Returns its iterations/s and mean PSNR over the 24
    views before and after."""
    W, H, N, deg = 360, 640, 20_000, 0
    cams = gs.orbit_cameras(24, W, H)
    gt_p, gt_sh = gs.synthetic_scene(N, deg, 1001)
    with gs.Workspace(N, deg, W, H) as gt:
        gt.load(gt_p, gt_sh)
        targets = []
        for cam in cams:
            img = gt.render(gs.make_frame(cam, N, deg))
            img[..., 3] = 255
            targets.append(img)
    bufs = [gs.DeviceBuffer(t) for t in targets]
    p0, s0 = gs.synthetic_scene(N, deg, 2002)
    tp = gs.TrainParams()

    def mean_psnr(ws):
        out = []
        for cam, t in zip(cams, targets):
            ws.render(gs.make_frame(cam, N, deg))
            img = ws.download(gs.BUF_IMAGE, np.float32).reshape(-1, 4)[:, :3].astype(np.float64)
            out.append(-10.0 * math.log10(max(float(np.mean((img - t.reshape(-1, 4)[:, :3] / 255.0) ** 2)), 1e-12)))
        return float(np.mean(out))

    with gs.Workspace(N, deg, W, H) as ws:
        ws.load(p0, s0)
        before = mean_psnr(ws)
        frames = [gs.make_frame(cams[s % 24], N, deg, step=s, tp=tp) for s in range(300)]
        gs.synchronize()
        t0 = time.perf_counter()
        for s in range(300):
            ws.set_target_device(bufs[s % 24].ptr)
            ws.train_step(frames[s], want_losses=False)
        gs.synchronize()
        dt = time.perf_counter() - t0
        after = mean_psnr(ws)
    for b in bufs:
        b.free()
    return dt, before, after


def kernel_check(gs, repeats=20):
    """The single-view forward + backward kernel check: 200k gaussians of the
    benchmark distributions at SH degree 3, one camera on the radius-3 orbit
    at 1280x720, the target a render of an independently seeded 200k scene.
    Full training steps with FLAG_NO_UPDATE (the model stays put, so every
    repeat sees the same pairs) timed kernel by kernel with HIP events; the
    rasterizers' Mpix/s and algorithmic GB/s from the step byte model."""
    n, deg, w, h = 200_000, 3, 1280, 720
    cam = gs.orbit_cameras(1, w, h)[0]
    gt_params, gt_sh = gs.synthetic_scene(n, deg, 7001)
    with gs.Workspace(n, deg, w, h) as gt:
        gt.load(gt_params, gt_sh)
        target = gt.render(gs.make_frame(cam, n, deg))
    target[..., 3] = 255
    params, sh = gs.synthetic_scene(n, deg, 7002)
    out = {"workload": "single view, 200k gaussians SH3 fp32, 1280x720, radius-3 orbit camera, 16x16 tiles, "
                       "float-bits depth key; NO_UPDATE training steps", "repeats": repeats}
    with gs.Workspace(n, deg, w, h) as ws:
        ws.load(params, sh)
        ws.set_target(target)
        frame = gs.make_frame(cam, n, deg, flags=gs.FLAG_NO_UPDATE)
        for _ in range(3):
            pairs, _ = ws.train_step(frame, want_losses=False)
        gs.synchronize()
        ws.timings(enable=True)
        for _ in range(repeats):
            ws.train_step(frame, want_losses=False)
        gs.synchronize()
        events = ws.timings(enable=False)
        visible = int((ws.download(gs.BUF_TILES, np.uint32, n) > 0).sum())
    table = kernel_table(events, repeats, n, visible, pairs, 1, w * h, SH_COEFFS[deg])
    out.update({"pairs": int(pairs), "visible_gaussians": visible})
    for k in ("forward", "backward"):
        v = table[k]
        out[k] = {"ms": round(v["ms_avg"], 4), "mpix_per_s": round(w * h / (v["ms_avg"] * 1e-3) / 1e6, 1),
                  "algorithmic_bytes": v["algorithmic_bytes"], "GBps": round(v["GBps"], 1),
                  "frac": round(v["GBps"] * 1e9 / HBM_PEAK, 4)}
    return out


def kernel_table(events, steps, slots_mean, n_visible, mean_pairs_per_view, per_step, pixels, coeffs):
    per_kernel = {}
    for name, ms in events:
        d = per_kernel.setdefault(name, [0.0, 0])
        d[0] += ms
        d[1] += 1
    kernels = {k: {"ms_avg": v[0] / v[1], "launches": v[1], "ms_per_step": v[0] / steps} for k, v in per_kernel.items()}
    for k, v in kernels.items():
        b = kernel_bytes(k, slots_mean, n_visible, mean_pairs_per_view, pixels, coeffs, per_step)
        if b is not None:
            v["algorithmic_bytes"] = int(b)
            v["GBps"] = b / (v["ms_avg"] * 1e-3) / 1e9
    return kernels


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed batched optimizer steps")
    ap.add_argument("--warmup", type=int, default=5, help="untimed batched optimizer steps")
    ap.add_argument("--batch", type=int, default=8, help="views per batched optimizer step")
    ap.add_argument("--gaussians", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=720)
    ap.add_argument("--sh-degree", type=int, default=3)
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--densify-every", type=int, default=100, help="view-iterations between densifications")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=8, help="oracle training steps timed for the CPU baseline")
    ap.add_argument("--no-cpu-small-scene", action="store_true",
                    help="skip the CPU baseline's full small-scene run (300 steps at 20k gaussians)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-kernel-check", action="store_true",
                    help="skip the single-view 200k SH3 forward/backward kernel check")
    ap.add_argument("--no-autotune", action="store_true", help="default launch shapes instead of the autotuner's")
    ap.add_argument("--shape", type=str, default=None, metavar="G,D,T,GRAPH[,BIN[,STORED]]",
                    help="fixed launch shape instead of autotuning: raster groups per CU (0 = one per tile), depth "
                         "and tile radix digit bits (tile 0: one counting pass), graph replay 0/1, pair bin shift "
                         "0..2 (bins of 16 << shift pixels; default 0), backward 0/1 (1: stored state; default 0)")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the timed steps after the targets are made (no held-out renders, 16 px pair probe "
                         "or depth-key comparison): for profiler runs")
    ap.add_argument("--no-quality", action="store_true",
                    help="skip the held-out PSNR curve of plain vs batched training at equal view-iterations")
    ap.add_argument("--quality-views", type=str, default="200,400,800,1600,3200",
                    help="view-iteration checkpoints of the quality curve")
    ap.add_argument("--batch-scaling", choices=("none", "sqrt"), default="none",
                    help="batched steps: Adam's hyperparameters as given (none) or scaled for the mean of the "
                         "step's views (sqrt: rates x sqrt(views), betas to the power views)")
    ap.add_argument("--depth-key", type=int, default=16, choices=(16, 32),
                    help="pair order within a tile: the reference's 16-bit log-depth key (train_project.comp:110) "
                         "or the float bits of the depth")
    ap.add_argument("--multi-scene", action="store_true",
                    help="instead: 16 scenes (~6M gaussians, mixed resolutions), 2 views each per step, as one "
                         "batch against scene after scene")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="after the bench, copy MIB MiB device to device twice (a known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counts to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build (multi-GPU "
              "sharding is a design note, not an implementation)", file=sys.stderr)
        sys.exit(2)

    gs = _load("gs_hip", "on-divice-gsplate_amd/gs_hip.py")
    device = gs.init(0)
    log(f"device {device}")
    if args.multi_scene:
        multi_scene(gs, args, device)
        return
    N, W, H, deg = args.gaussians, args.width, args.height, args.sh_degree
    coeffs = SH_COEFFS[deg]
    cams = gs.jitter_cameras(args.views, W, H, 3003)

    log(f"ground truth: {N} gaussians, {args.views} jittered views at {W}x{H}")
    gt_params, gt_sh = gs.synthetic_scene(N, deg, 1001)
    targets = []
    with gs.Workspace(N, deg, W, H) as gt:
        gt.load(gt_params, gt_sh)
        for cam in cams:
            img = gt.render(gs.make_frame(cam, N, deg))
            img[..., 3] = 255
            targets.append(img)
        # Held-out views (not trained on): the models' quality after equal view-iterations.
        held_cams = gs.jitter_cameras(8, W, H, 4004)
        held_targets = []
        for cam in held_cams:
            img = gt.render(gs.make_frame(cam, N, deg))
            held_targets.append(img)
    target_bufs = [gs.DeviceBuffer(t) for t in targets]
    del gt_params, gt_sh

    params, sh = gs.synthetic_scene(N, deg, 2002)
    order = np.random.default_rng(3003).permutation(args.views)
    key_flags = gs.FLAG_DEPTH16 if args.depth_key == 16 else 0
    tp = gs.TrainParams()
    views_total = args.steps * args.batch
    warmup_views = args.warmup * args.batch
    results = {}
    modes = [("plain", 1, key_flags), ("batched", args.batch, key_flags)]
    if not args.no_plain and args.depth_key == 16:
        modes.append(("plain_float_key", 1, 0))  # the configuration's own (tile << 32) | float-bits(z) order
    for mode, per_step, mode_flags in modes:
        if mode.startswith("plain") and args.no_plain:
            continue
        log(f"{mode}: {views_total} view-iterations, {per_step} view(s) per optimizer step")
        ws = gs.Workspace(2 * N, deg, W, H)
        ws.set_views(per_step)
        ws.load(params, sh)
        tuning, tune_ms = None, None
        if args.shape:
            g, d, t, gr, *rest = (int(x) for x in args.shape.split(","))
            tuning = ws.set_tuning(raster_groups_per_cu=g, depth_digit_bits=d, tile_digit_bits=t, graph=gr,
                                   bin_shift=rest[0] if rest else 0,
                                   stored_state=rest[1] if len(rest) > 1 else 0).as_dict()
            log(f"{mode}: launch shape {tuning} (given)")
        elif not args.no_autotune:
            vs = [int(order[k % len(order)]) for k in range(per_step)]
            frames = [gs.make_frame(cams[v], N, deg, step=0, tp=tp, flags=mode_flags) for v in vs]
            best, tune_ms = ws.autotune(frames, [target_bufs[v].ptr for v in vs], repeats=3)
            tuning = best.as_dict()
            log(f"{mode}: autotuned launch shape {tuning} ({tune_ms:.3f} ms per step while tuning)")
        if tuning is None:
            tuning = ws.tuning().as_dict()
        graph = bool(tuning["graph"])
        state = {"slots": N, "densified": 0, "pairs": 0, "views": 0, "step": 0, "flags": mode_flags,
                 "batch_scaling": per_step if args.batch_scaling == "sqrt" else 1}
        # The timed run records no per-kernel events (creating and recording
        # them is host work inside every step, and graph replay needs the
        # eager launches they bracket); a short eager pass after it times the
        # kernels.
        elapsed, steps = run(gs, ws, cams, target_bufs, order, per_step, views_total, warmup_views, deg, tp,
                             args.densify_every, state, timed=False)
        # The model's quality now, before the kernel-timing pass below trains it further.
        heldout = None if args.no_extras else heldout_psnr(gs, ws, held_cams, held_targets, state["slots"], deg)
        if heldout is not None:
            log(f"{mode}: held-out PSNR {heldout:.3f} dB after {state['views']} view-iterations")
        probe = dict(state)
        probe_views = max(per_step, (views_total // 5 // per_step) * per_step)
        run(gs, ws, cams, target_bufs, order, per_step, probe_views, 0, deg, tp, 0, probe, timed=True)
        kernel_steps = probe_views // per_step
        events = ws.timings(enable=False)
        views_per_s = steps * per_step / elapsed
        mean_bin_pairs = state["pairs"] / (steps * per_step)
        # The byte model counts 16x16 tile pairs; with coarser bins the run's
        # own pair count is lower than that, by the bins' sharing of entries.
        if args.no_extras:  # the run's own pairs (16 px tile pairs only at bin shift 0)
            tile_pairs = probe_bin_pairs = state["pairs"] / (steps * per_step)
        else:
            tile_pairs, probe_bin_pairs = mean_tile_pairs(gs, ws, cams, target_bufs, order, per_step, state, deg,
                                                          tp)
        mean_pairs = tile_pairs
        slots_mean = 0.5 * (state["slots_start"] + state["slots"])
        tiles_touched = ws.download(gs.BUF_TILES, np.uint32, state["slots"])
        n_visible = int((tiles_touched > 0).sum())
        results[mode] = {
            "views_per_s": views_per_s, "optimizer_steps_per_s": steps / elapsed, "ms_per_step": 1000 * elapsed / steps,
            "mean_pairs": mean_pairs, "mean_bin_pairs": mean_bin_pairs, "probe_bin_pairs": probe_bin_pairs,
            "slots_start": state["slots_start"], "slots_end": state["slots"],
            "densify_events": state["densified"], "slots_mean": slots_mean, "n_visible": n_visible,
            "kernels": kernel_table(events, kernel_steps, slots_mean, n_visible, mean_pairs, per_step, W * H, coeffs),
            "tuning": tuning, "tune_ms": tune_ms, "overflow_steps": ws.overflows(),
            "kernel_timing": "eager pass after the timed run",
        }
        log(f"{mode}: {views_per_s:.1f} view-iterations/s ({1000 * elapsed / steps:.3f} ms per optimizer step)")
        results[mode]["heldout_psnr"] = heldout
        ws.close()

    main_mode = "batched"
    r = results[main_mode]
    kernels = r["kernels"]
    per_view = {k: v["ms_per_step"] / args.batch for k, v in kernels.items()}
    dominant = max(per_view, key=per_view.get) if per_view else None
    roofline = None
    default_workload = (args.batch == 8 and N == 1_000_000 and (W, H) == (1280, 720) and deg == 3 and
                        args.depth_key == 16)
    if dominant is not None and "GBps" in kernels[dominant]:
        ach = kernels[dominant]["GBps"]
        traffic = profiled_traffic(dominant, r["tuning"]) if default_workload else None
        roofline = {"kernel": dominant, "bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK / 1e9,
                    "unit": "GB/s", "frac": round(ach * 1e9 / HBM_PEAK, 5), "traffic": traffic,
                    "traffic_source": TRAFFIC_PROFILE if traffic is not None else None,
                    "algorithmic_bytes": kernels[dominant].get("algorithmic_bytes"),
                    "bytes_model": "44 B per pair + 20 B per pixel + 36 B per visible gaussian, per view",
                    "ms_avg": round(kernels[dominant]["ms_avg"], 4)}
    # The rasterizers' VALU issue rate: wave64 VALU instructions per launch (the
    # committed SQ counter profile of this workload) over the launch time,
    # against 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction.
    valu = {}
    for k in ("backward", "forward"):
        insts = profiled_valu(k, r["tuning"]) if default_workload else None
        if insts is not None and k in kernels:
            rate = insts / (kernels[k]["ms_avg"] * 1e-3)
            valu[k] = {"insts_per_launch": insts, "achieved": round(rate / 1e12, 4), "peak": VALU_PEAK / 1e12,
                       "unit": "T wave-instructions/s", "frac": round(rate / VALU_PEAK, 4)}
    # Whole-step bytes per view-iteration; the batched step reads and writes the
    # optimizer state once per `batch` views.
    D = 4 * (11 + 3 * (deg + 1) ** 2)
    adam_bytes = 6 * D + 36
    sb_plain = step_bytes(r["slots_mean"], r["mean_pairs"], W * H, deg)
    sb = sb_plain - adam_bytes * r["slots_mean"] * (1.0 - 1.0 / args.batch)

    cpu = None
    if not args.no_cpu_baseline:
        orc = _load("gs_oracle", "oracle/oracle.py")
        threads = cpu_threads()
        log(f"cpu baseline: the oracle's training step, {args.cpu_steps} steps of this scene on {threads} threads")
        dt = cpu_baseline(gs, orc, params.copy(), sh.copy(), cams, targets, order, deg, args.cpu_steps)
        cpu = {"value": round(1.0 / dt, 5), "unit": "iter/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} plain training steps (1 view each: project, bin, composite, loss + SSIM "
                         f"gradient, tile backward, projection adjoint, Adam) of this {N}-gaussian SH{deg} {W}x{H} "
                         f"scene through the oracle's C restatement (OpenMP) on {cpu_model()}, {threads} threads "
                         f"({os.cpu_count()} present): {dt:.2f} s per step; own CPU path, not the reference's "
                         "Vulkan-on-lavapipe"}
        if not args.no_cpu_small_scene:
            log("cpu baseline: the small-scene configuration in full (20k SH0 360x640, 300 steps)")
            ds, p0, p1 = cpu_small_scene(gs, orc)
            cpu["small_scene_full"] = {"seconds": round(ds, 2), "iters_per_s": round(300 / ds, 3), "steps": 300,
                                       "psnr_start": round(p0, 4), "psnr_end": round(p1, 4),
                                       "workload": "20k gaussians SH0, 24 orbit views at 360x640, 0.8 L1 + 0.2 "
                                                   "D-SSIM, Adam 1.6e-4 / 5e-3, no densification"}
        log(f"cpu baseline: {cpu['value']} iter/s at the benchmark scene")
    if args.depth_key == 16 and not args.no_extras:
        log("depth-key difference on one view")
        key_diff = depth_key_difference(gs, params, sh, cams[int(order[0])], target_bufs[int(order[0])], deg)
    else:
        key_diff = None

    check = None
    quality = None
    if not args.no_quality:
        checkpoints = [int(x) for x in args.quality_views.split(",") if x]
        orders = [order] + [np.random.default_rng(seed).permutation(args.views) for seed in (3004, 3005)]
        quality = quality_curve(gs, cams, target_bufs, orders, params, sh, held_cams, held_targets, deg, tp, key_flags,
                                args.batch, args.densify_every, checkpoints)
        quality["note"] = ("held-out PSNR (8 views not trained on) after equal view-iterations, mean and range over "
                           "3 view orders, same start and densification every "
                           f"{args.densify_every} view-iterations; plain takes one Adam "
                           f"step per view, batched one per {args.batch} views on their mean gradient (lr_scaled: "
                           f"rates x sqrt({args.batch}), betas ^{args.batch}); seconds are training time to each "
                           "checkpoint (eager, default launch shape, deterministic backward)")
    small = None
    if not args.no_kernel_check:
        log("small scene on the GPU: 20k SH0 360x640, 24 views, 300 plain steps")
        ds, p0, p1 = small_scene_gpu(gs)
        small = {"workload": "20k gaussians SH0, 24 orbit views at 360x640, 300 plain steps (one view per Adam "
                             "step), 0.8 L1 + 0.2 D-SSIM, Adam 1.6e-4 / 5e-3, no densification; default launch shape",
                 "seconds": round(ds, 4), "iters_per_s": round(300 / ds, 1), "psnr_start": round(p0, 4),
                 "psnr_end": round(p1, 4)}
        log(f"small scene on the GPU: {300 / ds:.0f} it/s, {p0:.2f} -> {p1:.2f} dB")
        log("kernel check: single view, 200k gaussians SH3, 1280x720")
        check = kernel_check(gs)
        log(f"kernel check: forward {check['forward']['mpix_per_s']} Mpix/s, backward "
            f"{check['backward']['mpix_per_s']} Mpix/s")

    plain = results.get("plain")
    out = {
        "metric": "train iter/s (fwd+bwd+SSIM+Adam) @1M Gaussians 1280x720; batched vs plain, %HBM",
        "value": round(r["views_per_s"], 3),
        "unit": "iter/s (view-iterations)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded 1M-gaussian GT scene; 64 jittered views rendered by this trainer)",
        "config": {"workload": f"batched training: {args.batch} views per optimizer step (project, bin, forward, "
                               "SSIM loss and recompute backward of all views, one launch per stage; one adjoint + "
                               "Adam on the mean gradient), 1M gaussians SH3 fp32, 1280x720, 16x16 tiles, "
                               f"{args.depth_key}-bit depth key, 0.8 L1 + 0.2 D-SSIM, densify every 100 "
                               "view-iterations",
                   "views_per_step": args.batch, "gaussians_start": N, "gaussians_end": r["slots_end"],
                   "width": W, "height": H, "sh_degree": deg, "mean_pairs_per_view": int(r["mean_pairs"]),
                   "mean_bin_pairs_per_view": int(r["mean_bin_pairs"]),
                   "depth_key_bits": args.depth_key,
                   "densify_events": r["densify_events"], "device": device, "tuning": r["tuning"],
                   "kernel_timing": r["kernel_timing"], "overflow_steps": r["overflow_steps"]},
        "plain": None if plain is None else {
            "iters_per_s": round(plain["views_per_s"], 3), "ms_per_step": round(plain["ms_per_step"], 4),
            "mean_pairs": int(plain["mean_pairs"]), "densify_events": plain["densify_events"],
            "tuning": plain["tuning"],
            "kernels": {k: round(v["ms_avg"], 4) for k, v in plain["kernels"].items()}},
        "speedup_batched_vs_plain": None if plain is None else round(r["views_per_s"] / plain["views_per_s"], 3),
        "optimizer_steps_per_s": round(r["optimizer_steps_per_s"], 3),
        "mpix_per_s": round(r["views_per_s"] * W * H / 1e6, 2),
        "step_hbm": {"bytes_per_view_iter": int(sb), "achieved_GBps": round(sb * r["views_per_s"] / 1e9, 2),
                     "frac": round(sb * r["views_per_s"] / HBM_PEAK, 5),
                     "plain_frac": None if plain is None else round(
                         step_bytes(plain["slots_mean"], plain["mean_pairs"], W * H, deg) * plain["views_per_s"] /
                         HBM_PEAK, 5)},
        "roofline": roofline,
        "valu": valu or None,
        "plain_float_key": None if "plain_float_key" not in results else {
            "iters_per_s": round(results["plain_float_key"]["views_per_s"], 3),
            "ms_per_step": round(results["plain_float_key"]["ms_per_step"], 4),
            "heldout_psnr": rnd(results["plain_float_key"]["heldout_psnr"]),
            "order": "(tile << 32) | float bits of z, ties by gaussian index"},
        "depth_key_difference": key_diff,
        "kernel_check": check,
        "small_scene": small,
        "quality_curve": quality,
        "heldout_psnr": {"batched": rnd(r["heldout_psnr"]),
                         "plain": None if plain is None else rnd(plain["heldout_psnr"]),
                         "views": 8, "view_iterations": args.steps * args.batch + args.warmup * args.batch,
                         "note": "equal view-iterations: the plain run takes one Adam step per view, the batched "
                                 f"run one per {args.batch} views"},
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_step"])},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    for buf in target_bufs:
        buf.free()
    if args.calibrate_copy > 0:
        nbytes = args.calibrate_copy << 20
        src, dst = gs.device_alloc(nbytes), gs.device_alloc(nbytes)
        for _ in range(2):
            gs.device_copy(dst, src, nbytes)
        gs.device_free(src)
        gs.device_free(dst)
        log(f"copy calibration: k_copy16 read {args.calibrate_copy} MiB + wrote {args.calibrate_copy} MiB, twice")


if __name__ == "__main__":
    main()
