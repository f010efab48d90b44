"""The HIP path against the oracle at the benchmark configurations' own sizes.

* 1M gaussians, SH degree 3, 1280x720 (the full training step's scene: GT
  seed 1001 rendered at an orbit camera, training scene seed 2002):
  - a plain step with the float depth key ((tile << 32) | float bits of z):
    bit-exact projection and binning, forward, loss and backward within the
    bars of parity_check.py, the optimizer's gradient against the fp64
    projection adjoint; then densification at 1M (prune / clone / split
    counts and placement against the oracle's predicate on the GPU's stats);
  - an 8-view batched step (seeded jittered cameras, 16-bit depth key): every
    view checked as above, and the optimizer's gradient against the mean of
    the views' adjoints.
* The 16-scene batch (about 6M gaussians, 640x360 / 1280x720 / 1920x1080):
  one view of the smallest scene at each resolution checked as above inside
  a batch step of all 16 scenes.
The batched cases run with both backwards: the recompute walk and the
stored-state (bucket) sweep.
"""
import math

import numpy as np
import pytest

from helpers import gs, orc
from parity_check import check_adjoint, check_view, first_step_gradients, frozen

pytestmark = pytest.mark.gpu

N, DEG, W, H = 1_000_000, 3, 1280, 720


@pytest.fixture(scope="module", autouse=True)
def device():
    print(f"\n[gpu] {gs.init(0)}", flush=True)
    yield
    gs.synchronize()


def case(msg):
    print(f"[case] {msg}", flush=True)


def gt_targets(cams, n=N, deg=DEG, seed=1001):
    gt_p, gt_sh = gs.synthetic_scene(n, deg, seed)
    out = []
    with gs.Workspace(n, deg, cams[0].width, cams[0].height) as gt:
        gt.load(gt_p, gt_sh)
        for cam in cams:
            img = gt.render(gs.make_frame(cam, n, deg))
            img[..., 3] = 255
            out.append(img)
    return out


def test_full_step_1m_plain_float_key_then_densify():
    case("1M SH3 1280x720 plain step, float depth key, then densify")
    cam = gs.orbit_cameras(64, W, H)[5]
    target = gt_targets([cam])[0]
    params, sh = gs.synthetic_scene(N, DEG, 2002)
    cap = 1_300_000
    with gs.Workspace(cap, DEG, W, H) as ws:
        ws.load(params, sh)
        ws.set_target(target)
        frame = gs.make_frame(cam, N, DEG, tp=frozen())
        pairs, losses = ws.train_step(frame)
        g2 = check_view(ws, 0, pairs, params, sh, cam, target, losses, False, "1M plain")
        gp, gsh = first_step_gradients(ws, N)
        after = ws.store(N)[0]
        np.testing.assert_array_equal(after, params)  # rates 0
        check_adjoint(gp, gsh, params, sh, [cam], DEG, [g2], "1M plain")
        # Densify on the statistics this step left (thresholds that make every action occur).
        stats = ws.download(gs.BUF_STATS, np.float32, N * 4).reshape(N, 4)
        assert (stats[:, 1] > 0).sum() > 100_000
        grad = stats[:, 0] / np.maximum(stats[:, 1], 1)
        thr = float(np.quantile(grad[stats[:, 1] > 0], 0.9))
        tp = gs.TrainParams(grad_threshold=thr, clone_max_scale=0.02, prune_opacity=0.2)
        dframe = gs.make_frame(cam, N, DEG, step=100, tp=tp)
        slots, (freed, placed, pruned) = ws.densify(dframe)
        p_after = ws.store(min(slots, cap))[0]
    action, release, prune = orc.densify_mark(params, stats, thr, 0.02, 0.2)
    spawners = np.nonzero(action)[0]
    free_list = np.nonzero(release)[0]
    print(f"[info] densify at 1M: {len(spawners)} spawners ({(action == 1).sum()} clones), {pruned} pruned, "
          f"{placed} placed, slots {N} -> {slots}", flush=True)
    assert pruned == int(prune.sum()) and pruned > 1000
    assert freed == int(release.sum())
    assert (action == 1).sum() > 100 and (action == 2).sum() > 100
    assert placed == min(len(spawners), len(free_list) + cap - N)
    assert slots == N + max(0, placed - len(free_list))
    targets = np.array([free_list[k] if k < len(free_list) else N + k - len(free_list) for k in range(placed)])
    reused = np.zeros(cap, bool)
    reused[targets] = True
    gone = np.nonzero(prune.astype(bool) & ~reused[:N])[0]
    assert np.all(p_after[gone, 3] == -20.0) and np.all(p_after[gone, 7] == 0.0)
    src = spawners[:placed]
    np.testing.assert_array_equal(p_after[targets, 8:16], params[src, 8:16])
    clones = action[src] == 1
    np.testing.assert_array_equal(p_after[targets[clones], 0:8], params[src[clones], 0:8])
    splits = src[~clones]
    np.testing.assert_allclose(p_after[splits, 4:7], params[splits, 4:7] - np.log(np.float32(1.6)), atol=1e-6)


BACKWARDS = [0, 1]
BACKWARD_IDS = ["recompute", "stored-state"]


@pytest.mark.parametrize("stored", BACKWARDS, ids=BACKWARD_IDS)
def test_batched_8_views_1m_depth16(stored):
    case(f"1M SH3 1280x720, 8 jittered views in one batched step, 16-bit depth key, {BACKWARD_IDS[stored]} backward")
    cams = gs.jitter_cameras(8, W, H, 3003)
    targets = gt_targets(cams)
    bufs = [gs.DeviceBuffer(t) for t in targets]
    params, sh = gs.synthetic_scene(N, DEG, 2002)
    try:
        with gs.Workspace(N, DEG, W, H) as ws:
            ws.set_views(8)
            ws.set_tuning(stored_state=stored)
            ws.load(params, sh)
            frames = [gs.make_frame(c, N, DEG, tp=frozen(), flags=gs.FLAG_DEPTH16) for c in cams]
            pairs, losses = ws.train_views(frames, [b.ptr for b in bufs])
            g2s = []
            for v in range(8):
                g2s.append(check_view(ws, v, pairs[v], params, sh, cams[v], targets[v], losses[v], True,
                                      f"1M batched view {v}"))
            gp, gsh = first_step_gradients(ws, N)
            np.testing.assert_array_equal(ws.store(N)[0], params)
    finally:
        for b in bufs:
            b.free()
    check_adjoint(gp, gsh, params, sh, cams, DEG, g2s, "1M batched mean of 8 views")


@pytest.mark.parametrize("stored", BACKWARDS, ids=BACKWARD_IDS)
def test_multi_scene_batch_one_view_per_resolution(stored):
    case(f"16-scene batch (~6M gaussians): one view of the smallest scene at each resolution, "
         f"{BACKWARD_IDS[stored]} backward")
    rng_seed = 5000
    # The benchmark's workload draw (bench.py multi_scene_workload): counts and resolutions.
    for seed in range(rng_seed, rng_seed + 10000):
        rng = np.random.default_rng(seed)
        counts = np.floor(np.exp(rng.uniform(math.log(5e4), math.log(3e6), 16))).astype(int)
        res = [((640, 360), (1280, 720), (1920, 1080))[int(i)] for i in rng.integers(0, 3, 16)]
        if abs(counts.sum() - 6_000_000) <= 0.1 * 6_000_000:
            break
    picks = {}
    for s, (n, r) in enumerate(zip(counts, res)):
        if r not in picks or n < counts[picks[r]]:
            picks[r] = s
    scenes, cams, tgts, states = [], [], [], []
    try:
        for s, (n, (w, h)) in enumerate(zip(counts, res)):
            n = int(n)
            cs = gs.jitter_cameras(2, w, h, 3003 + s)
            gt_p, gt_sh = gs.synthetic_scene(n, DEG, 1001 + 100 * s)
            with gs.Workspace(n, DEG, w, h) as gt:
                gt.load(gt_p, gt_sh)
                ts = []
                for cam in cs:
                    img = gt.render(gs.make_frame(cam, n, DEG))
                    img[..., 3] = 255
                    ts.append(img)
            p, sh = gs.synthetic_scene(n, DEG, 2002 + 100 * s)
            ws = gs.Workspace(n, DEG, w, h)
            ws.set_views(2)
            ws.load(p, sh)
            scenes.append(ws)
            cams.append(cs)
            tgts.append((ts, [gs.DeviceBuffer(t) for t in ts]))
            states.append((p, sh) if s in picks.values() else None)
        with gs.Batch(scenes) as batch:
            batch.set_tuning(stored_state=stored)
            fps = [[gs.make_frame(c, int(n), DEG, tp=frozen(), flags=gs.FLAG_DEPTH16) for c in cs]
                   for n, cs in zip(counts, cams)]
            pairs, losses = batch.train(fps, [[b.ptr for b in t[1]] for t in tgts])
        for r, s in sorted(picks.items()):
            p, sh = states[s]
            first = 2 * s
            check_view(scenes[s], 0, pairs[first], p, sh, cams[s][0], tgts[s][0][0], losses[first], True,
                       f"batch scene {s} ({counts[s]} gaussians, {r[0]}x{r[1]}) view 0")
    finally:
        for ws in scenes:
            ws.close()
        for t in tgts:
            for b in t[1]:
                b.free()
