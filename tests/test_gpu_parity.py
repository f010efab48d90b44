"""HIP path against the CPU oracle, stage by stage, through the C ABI.

Bit-exact: projected records, tile counts, depth keys, the depth order, the
sorted pairs and tile ranges.  Within tolerance (stated per test): the
compositor (the kernels use the hardware exp), the loss gradient, the
backward's float-atomic sums and the projection adjoint.
"""
import math

import numpy as np
import pytest

from helpers import gs, make_camera, make_scene, noise_target, orc, smooth_target
from parity_check import check_backward, check_forward, check_loss

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    name = gs.init(0)
    print(f"\n[gpu] {name}", flush=True)
    yield name
    gs.synchronize()


def frozen(lam=0.2):
    """All learning rates zero: Adam's first moment then holds (1 - beta1) g."""
    return gs.TrainParams(lambda_dssim=lam, lr_position=0.0, lr_scale=0.0, lr_rotation=0.0, lr_opacity=0.0,
                          lr_sh_dc=0.0, lr_sh_rest=0.0)


def case(msg):
    print(f"[case] {msg}", flush=True)


def project_and_bin(ws, frame):
    ws.project(frame)
    return ws.bin(frame)


def gpu_state(ws, n, pairs):
    rec = ws.download(gs.BUF_RECORDS, np.float32, n * 12).reshape(n, 12)
    foot8 = ws.download(gs.BUF_FOOTPRINT, np.float32, n * 8).reshape(n, 8)
    foot = foot8[:, :4]
    rec_ = ws.download(gs.BUF_RECORDS, np.float32, n * 12).reshape(n, 12)
    np.testing.assert_array_equal(foot8[:, 4:6].view(np.uint32), rec_[:, 0:2].view(np.uint32))
    np.testing.assert_array_equal(foot8[:, 6].view(np.uint32), ws.download(gs.BUF_TILES, np.uint32, n))
    tiles = ws.download(gs.BUF_TILES, np.uint32, n)
    keys = ws.download(gs.BUF_DEPTH_KEYS, np.uint32, n)
    order = ws.download(gs.BUF_DEPTH_ORDER, np.uint32, n)
    pt = ws.download(gs.BUF_PAIR_TILES, np.uint32, pairs)
    pg = ws.download(gs.BUF_PAIR_GAUSSIANS, np.uint32, pairs)
    T = ws.tiles_x * ws.tiles_y
    ranges = ws.download(gs.BUF_RANGES, np.uint32, (T + 1) * 2).reshape(T + 1, 2)
    return rec, foot, tiles, keys, order, pt, pg, ranges


PROJECT_CASES = [
    ("small-sh1", 400, 1, 64, 48, 31, False),
    ("small-sh3-depth16", 2000, 3, 96, 72, 32, True),
    ("config2-200k-sh3", 200_000, 3, 1280, 720, 33, False),
]


@pytest.mark.parametrize("name,n,deg,w,h,seed,depth16", PROJECT_CASES, ids=[c[0] for c in PROJECT_CASES])
def test_project_and_bin_bit_exact(name, n, deg, w, h, seed, depth16):
    case(f"project+bin {name}: n={n} deg={deg} {w}x{h}")
    if n > 10000:
        params, sh = gs.synthetic_scene(n, deg, seed)
        cam = gs.look_at_camera((0.0, 0.5, 3.0), (0, 0, 0), w, h)
    else:
        params, sh = make_scene(n, deg, seed)
        cam = make_camera(w, h, 0.3)
    flags = gs.FLAG_DEPTH16 if depth16 else 0
    with gs.Workspace(n, deg, w, h) as ws:
        ws.load(params, sh)
        frame = gs.make_frame(cam, n, deg, flags=flags)
        pairs = project_and_bin(ws, frame)
        rec, foot, tiles, keys, order, pt, pg, ranges = gpu_state(ws, n, pairs)
    r_rec, r_foot, r_tiles, r_keys = orc.project32(params, sh, cam, deg, depth16=depth16)
    np.testing.assert_array_equal(tiles, r_tiles)
    # keys are sorted in place by the bin step: compare through the order; 16-bit keys carry the
    # slot's tile count in their upper half
    if depth16:
        np.testing.assert_array_equal(keys & 0xFFFF, r_keys[order] & 0xFFFF)
        np.testing.assert_array_equal(keys >> 16, tiles[order])
    else:
        np.testing.assert_array_equal(keys, r_keys[order])
    np.testing.assert_array_equal(order, np.argsort(r_keys, kind="stable").astype(np.uint32))
    np.testing.assert_array_equal(rec.view(np.uint32), r_rec.view(np.uint32))
    np.testing.assert_array_equal(foot.view(np.uint32), r_foot.view(np.uint32))
    o_pt, o_pg, o_ranges = orc.bin32(r_rec, r_foot, r_tiles, r_keys, w, h)
    assert pairs == len(o_pt) and pairs > 0
    np.testing.assert_array_equal(pt, o_pt)
    np.testing.assert_array_equal(pg, o_pg)
    np.testing.assert_array_equal(ranges, o_ranges)
    print(f"[ok] {name}: {int((tiles > 0).sum())} visible, {pairs} pairs", flush=True)


def test_bin_edge_cases():
    w, h = 64, 48
    cam = make_camera(w, h, 0.0)
    case("edge: nothing visible")
    params, sh = make_scene(50, 0, 1)
    params[:, 7] = 0.0  # all slots free
    with gs.Workspace(64, 0, w, h) as ws:
        ws.load(params, sh)
        frame = gs.make_frame(cam, 50, 0)
        assert project_and_bin(ws, frame) == 0
        ws.forward(frame)
        img = ws.download(gs.BUF_IMAGE, np.float32).reshape(-1, 4)
        assert np.all(img[:, 3] == 1.0) and np.all(img[:, :3] == 0.0)
    case("edge: one long tile list (> 4 LDS batches) and huge footprints")
    params, sh = make_scene(1500, 0, 2, spread=0.05)
    params[:10, 4:7] = math.log(2.0)  # a few covering the whole image
    with gs.Workspace(1500, 0, w, h) as ws:
        ws.load(params, sh)
        frame = gs.make_frame(cam, 1500, 0)
        pairs = project_and_bin(ws, frame)
        rec, foot, tiles, keys, order, pt, pg, ranges = gpu_state(ws, 1500, pairs)
        ws.forward(frame)
        img = ws.download(gs.BUF_IMAGE, np.float32).reshape(-1, 4)
        last = ws.download(gs.BUF_LAST, np.uint32, w * h)
    r_rec, r_foot, r_tiles, r_keys = orc.project32(params, sh, cam, 0)
    o_pt, o_pg, o_ranges = orc.bin32(r_rec, r_foot, r_tiles, r_keys, w, h)
    np.testing.assert_array_equal(pt, o_pt)
    np.testing.assert_array_equal(pg, o_pg)
    np.testing.assert_array_equal(ranges, o_ranges)
    assert (ranges[:, 1] - ranges[:, 0]).max() > 1024
    o_img, o_last = orc.forward32(r_rec, o_pg, o_ranges, w, h)
    assert np.mean(last == o_last) > 0.999
    assert np.abs(img - o_img).max() < 1e-4


def _pipeline(ws, frame, target, lam):
    pairs = project_and_bin(ws, frame)
    ws.forward(frame)
    ws.set_target(target)
    losses = ws.loss(frame)
    return pairs, losses


FORWARD_CASES = [("small", 300, 2, 96, 72, 41), ("config2-200k-sh3", 200_000, 3, 1280, 720, 42)]


@pytest.mark.parametrize("name,n,deg,w,h,seed", FORWARD_CASES, ids=[c[0] for c in FORWARD_CASES])
def test_forward_loss_backward_match_oracle(name, n, deg, w, h, seed):
    case(f"forward/loss/backward {name}")
    if n > 10000:
        params, sh = gs.synthetic_scene(n, deg, seed)
        cam = gs.look_at_camera((0.0, 0.5, 3.0), (0, 0, 0), w, h)
        target = smooth_target(w, h, seed)
    else:
        params, sh = make_scene(n, deg, seed)
        cam = make_camera(w, h, 1.1)
        target = noise_target(w, h, seed)
    bg = (0.1, 0.2, 0.3)
    lam = 0.2
    with gs.Workspace(n, deg, w, h) as ws:
        ws.load(params, sh)
        frame = gs.make_frame(cam, n, deg, background=bg, tp=frozen(lam))
        pairs, losses = _pipeline(ws, frame, target, lam)
        rec, foot, tiles, keys, order, pt, pg, ranges = gpu_state(ws, n, pairs)
        img = ws.download(gs.BUF_IMAGE, np.float32).reshape(-1, 4)
        last = ws.download(gs.BUF_LAST, np.uint32, w * h)
        dldc = ws.download(gs.BUF_DLDC, np.float32).reshape(-1, 4)
        ws.backward(frame)
        g2 = ws.download(gs.BUF_GRAD2D, np.float32, 9 * n).reshape(n, 9)
    # forward and loss against the oracle on the GPU's own records, pairs and image (parity_check.py)
    check_forward(rec, pg, ranges, img, last, w, h, name)
    L_gpu = check_loss(img, dldc, losses, target, bg, lam, w, h, name)
    # backward: float-atomic sums against the oracle's tile backward on the same state, each
    # gaussian within 1e-4 of the value's largest plus its share at ambiguous pixels (parity_check.py)
    check_backward(rec, pg, ranges, img, last, dldc, bg, w, h, g2, tiles, name)
    print(f"[ok] {name}: pairs {pairs}, loss {L_gpu:.6f}", flush=True)


def _first_step_gradients(ws, n):
    m = ws.download(gs.BUF_M_PARAMS, np.float32, n * 16).reshape(n, 16) / (1 - np.float32(0.9))
    coeffs = ws.sh_coeffs
    msh = ws.download(gs.BUF_M_SH, np.float32, coeffs * ws.capacity * 3).reshape(coeffs, ws.capacity, 3)
    msh = np.transpose(msh[:, :n, :], (1, 0, 2)) / (1 - np.float32(0.9))
    return m, msh


def test_adjoint_matches_oracle():
    case("adjoint: GPU grad2d through the fp64 projection adjoint")
    w, h, n, deg = 96, 72, 400, 3
    params, sh = make_scene(n, deg, 51)
    cam = make_camera(w, h, 2.2)
    target = smooth_target(w, h, 51)
    with gs.Workspace(n, deg, w, h) as ws:
        ws.load(params, sh)
        ws.set_target(target)
        frame = gs.make_frame(cam, n, deg, tp=frozen())
        ws.train_step(frame)
        g2 = ws.download(gs.BUF_GRAD2D, np.float32, 9 * n).reshape(n, 9).astype(np.float64)
        gp, gsh = _first_step_gradients(ws, n)
        after = ws.store(n)[0]
    np.testing.assert_array_equal(after, params)  # lr 0: parameters stay put
    r_gp, r_gsh = orc.adjoint64(params, sh, cam, deg, g2)
    groups = {"position": [0, 1, 2], "opacity": [3], "scale": [4, 5, 6], "rotation": [8, 9, 10, 11],
              "dc": [12, 13, 14]}
    for gname, cols in groups.items():
        scale = np.abs(r_gp[:, cols]).max()
        err = np.abs(gp[:, cols] - r_gp[:, cols]).max()
        assert scale > 0 and err <= 1e-4 * scale, f"{gname}: err {err:.3g} vs {scale:.3g}"
    scale = np.abs(r_gsh).max()
    assert np.abs(gsh - r_gsh).max() <= 1e-4 * scale


def test_gradients_against_fp64_reference_and_finite_differences():
    """The reference's check_gradients (train_selftest.cpp:356-486): a
    10-gaussian 64x48 scene; image within 2.5/255 and loss within 1e-4 of the
    fp64 render; every parameter gradient against the fp64 analytic backward
    (1e-4 of the group's largest) and against FP64 central differences
    (|g - fd| <= 0.05 |fd| + 2e-6, at most 5% off per group)."""
    case("gradients vs fp64 backward and central differences")
    w, h, n, deg = 64, 48, 10, 2
    params, sh = make_scene(n, deg, 61, spread=0.4)
    cam = make_camera(w, h, 0.3)
    target = smooth_target(w, h, 61)
    bg = (0.2, 0.1, 0.0)
    with gs.Workspace(n, deg, w, h) as ws:
        ws.load(params, sh)
        ws.set_target(target)
        frame = gs.make_frame(cam, n, deg, background=bg, tp=frozen(0.2))
        _, losses = ws.train_step(frame)
        img = ws.download(gs.BUF_IMAGE, np.float32).reshape(-1, 4)
        gp, gsh = _first_step_gradients(ws, n)
    p64, s64 = params.astype(np.float64), sh.astype(np.float64)
    L, _, r_gp, r_gsh = orc.backward64(p64, s64, cam, deg, target, 0.2, bg=bg)
    rgb, _ = orc.render64(p64, s64, cam, deg, bg)
    out = img[:, :3] + img[:, 3:4] * np.asarray(bg, np.float32)
    assert np.abs(out - rgb).max() < 2.5 / 255
    L_gpu = 0.8 * losses[0] / (3 * w * h) + 0.2 * (1 - losses[2] / (3 * w * h))
    assert abs(L_gpu - L) < 1e-4
    for cols in ([0, 1, 2], [3], [4, 5, 6], [8, 9, 10, 11], [12, 13, 14]):
        scale = np.abs(r_gp[:, cols]).max()
        assert np.abs(gp[:, cols] - r_gp[:, cols]).max() <= 1e-4 * scale, cols
    assert np.abs(gsh - r_gsh).max() <= 1e-4 * np.abs(r_gsh).max()
    rng = np.random.default_rng(4004)
    eps = 1e-4
    bad = checked = 0
    for _ in range(60):
        i, c = int(rng.integers(n)), int(rng.choice([0, 1, 2, 3, 4, 5, 6, 8, 9, 10, 11, 12, 13, 14]))
        pp, pm = p64.copy(), p64.copy()
        pp[i, c] += eps
        pm[i, c] -= eps
        fd = (orc.loss64(pp, s64, cam, deg, target, 0.2, bg=bg) - orc.loss64(pm, s64, cam, deg, target, 0.2,
                                                                             bg=bg)) / (2 * eps)
        checked += 1
        bad += abs(gp[i, c] - fd) > 0.05 * abs(fd) + 2e-6
    assert bad <= checked // 20, f"{bad}/{checked} off"


def test_adam_step_matches_oracle():
    case("adam: one step with the benchmark learning rates")
    w, h, n, deg = 96, 72, 300, 1
    params, sh = make_scene(n, deg, 71)
    cam = make_camera(w, h, 0.9)
    target = smooth_target(w, h, 71)
    with gs.Workspace(n, deg, w, h) as ws:
        ws.load(params, sh)
        ws.set_target(target)
        frame0 = gs.make_frame(cam, n, deg, tp=frozen())
        ws.train_step(frame0)  # step 0, lr 0: m, v hold the first gradient
        g, gsh = _first_step_gradients(ws, n)
        m0 = ws.download(gs.BUF_M_PARAMS, np.float32, n * 16).reshape(n, 16)
        v0 = ws.download(gs.BUF_V_PARAMS, np.float32, n * 16).reshape(n, 16)
        tp = gs.TrainParams()
        frame1 = gs.make_frame(cam, n, deg, step=1, tp=tp)
        ws.train_step(frame1)  # same view, same parameters: the same gradient again
        p1, sh1 = ws.store(n)
        m1 = ws.download(gs.BUF_M_PARAMS, np.float32, n * 16).reshape(n, 16)
    lr = np.zeros((n, 16), np.float32)
    lr[:, 0:3], lr[:, 3], lr[:, 4:7], lr[:, 8:12], lr[:, 12:15] = (tp.lr_position, tp.lr_opacity, tp.lr_scale,
                                                                   tp.lr_rotation, tp.lr_sh_dc)
    vis = np.any(g != 0, axis=1)
    want_p, want_m, _ = orc.adam(params, g, m0, v0, lr, bias1=frame1.bias1, bias2=frame1.bias2)
    np.testing.assert_allclose(m1[vis], want_m[vis], rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(p1, want_p, rtol=1e-5, atol=1e-6)


def _pcg_hash(v):
    v = np.uint64(v)
    state = (v * np.uint64(747796405) + np.uint64(2891336453)) & np.uint64(0xffffffff)
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(277803737)) & np.uint64(
        0xffffffff)
    return int((word >> np.uint64(22)) ^ word)


def test_densify_masks_and_slots():
    case("densify: prune / clone / split by compaction")
    w, h, n, deg, cap = 64, 48, 1000, 1, 1400
    rng = np.random.default_rng(81)
    params, sh = make_scene(n, deg, 81)
    params[:, 3] = rng.normal(-2.0, 3.0, size=n)
    params[rng.uniform(size=n) < 0.05, 7] = 0.0  # some free slots
    stats = np.zeros((n, 4), np.float32)
    stats[:, 1] = rng.integers(0, 5, size=n)
    stats[:, 0] = stats[:, 1] * rng.uniform(0, 4e-4, size=n)
    cam = make_camera(w, h, 0.0)
    tp = gs.TrainParams(clone_max_scale=0.1)
    frame = gs.make_frame(cam, n, deg, step=7, tp=tp)
    with gs.Workspace(cap, deg, w, h) as ws:
        ws.load(params, sh)
        ws.upload(gs.BUF_STATS, stats)
        slots, (freed, placed, pruned) = ws.densify(frame)
        p_after, sh_after = ws.store(min(slots, cap))
        st_after = ws.download(gs.BUF_STATS, np.float32, cap * 4)
    action, release, prune = orc.densify_mark(params, stats, tp.grad_threshold, tp.clone_max_scale,
                                              tp.prune_opacity)
    assert pruned == int(prune.sum()) and pruned > 0
    assert freed == int(release.sum())
    spawners = np.nonzero(action)[0]
    free_list = np.nonzero(release)[0]
    assert placed == min(len(spawners), len(free_list) + cap - n) and placed > 0
    assert slots == n + max(0, placed - len(free_list))
    assert np.all(st_after == 0)
    targets = [free_list[k] if k < len(free_list) else n + k - len(free_list) for k in range(placed)]
    reused = np.zeros(n, bool)
    reused[[t for t in targets if t < n]] = True
    assert np.all(p_after[:n][prune.astype(bool) & ~reused, 3] == -20.0)  # pruned and not refilled
    seed = frame.seed
    for k, i in enumerate(spawners[:placed]):
        target = free_list[k] if k < len(free_list) else n + k - len(free_list)
        np.testing.assert_array_equal(p_after[target, 8:16], params[i, 8:16])
        np.testing.assert_array_equal(sh_after[target], sh[i])
        if action[i] == 1:
            np.testing.assert_array_equal(p_after[target, 0:8], params[i, 0:8])
        else:
            np.testing.assert_allclose(p_after[target, 4:7], params[i, 4:7] - np.log(np.float32(1.6)), atol=1e-6)
            np.testing.assert_allclose(p_after[i, 4:7], params[i, 4:7] - np.log(np.float32(1.6)), atol=1e-6)
            # Box-Muller on the PCG stream (train_common.glsl:144-161)
            state = _pcg_hash(int(i) ^ _pcg_hash(seed))
            u = []
            for _ in range(8):
                state = _pcg_hash(state)
                u.append((state >> 8) / 16777216.0)
            def normal3(u1, u2, u3, u4):
                u1, u3 = max(u1, 1e-7), max(u3, 1e-7)
                r1, r2 = math.sqrt(-2 * math.log(u1)), math.sqrt(-2 * math.log(u3))
                return np.array([r1 * math.cos(2 * math.pi * u2), r1 * math.sin(2 * math.pi * u2),
                                 r2 * math.cos(2 * math.pi * u4)])
            na, nb = normal3(*u[0:4]), normal3(*u[4:8])
            q = params[i, 8:12].astype(np.float64)
            q /= np.linalg.norm(q)
            wq, xq, yq, zq = q
            R = np.array([[1 - 2 * (yq * yq + zq * zq), 2 * (xq * yq - wq * zq), 2 * (xq * zq + wq * yq)],
                          [2 * (xq * yq + wq * zq), 1 - 2 * (xq * xq + zq * zq), 2 * (yq * zq - wq * xq)],
                          [2 * (xq * zq - wq * yq), 2 * (yq * zq + wq * xq), 1 - 2 * (xq * xq + yq * yq)]])
            s = np.exp(params[i, 4:7].astype(np.float64))
            np.testing.assert_allclose(p_after[i, 0:3], params[i, 0:3] + R @ (na * s), atol=2e-5)
            np.testing.assert_allclose(p_after[target, 0:3], params[i, 0:3] + R @ (nb * s), atol=2e-5)


def test_train_views_matches_single_view_steps():
    """A batched step over four views, each on its own stream: every view's
    2D gradient equals that view's plain step from the same state, and the
    optimizer sees their mean (Adam's first moment over 1 - beta1)."""
    case("batched views vs single-view steps")
    w, h, n, deg, V = 96, 72, 500, 2, 4
    params, sh = make_scene(n, deg, 91)
    cams = [make_camera(w, h, 0.5 + 1.3 * v) for v in range(V)]
    targets = [smooth_target(w, h, 91 + v) for v in range(V)]
    s_g2, s_gp, s_gsh, s_loss, s_vis = [], [], [], [], []
    for v in range(V):
        with gs.Workspace(n, deg, w, h) as ws:
            ws.load(params, sh)
            ws.set_target(targets[v])
            _, losses = ws.train_step(gs.make_frame(cams[v], n, deg, tp=frozen()))
            s_g2.append(ws.download(gs.BUF_GRAD2D, np.float32, 9 * n).reshape(n, 9))
            gp, gsh = _first_step_gradients(ws, n)
            s_gp.append(gp)
            s_gsh.append(gsh)
            s_loss.append(losses)
            s_vis.append(ws.download(gs.BUF_TILES, np.uint32, n) > 0)
    bufs = [gs.DeviceBuffer(t) for t in targets]
    try:
        with gs.Workspace(n, deg, w, h) as ws:
            ws.set_views(V)
            ws.load(params, sh)
            frames = [gs.make_frame(c, n, deg, tp=frozen()) for c in cams]
            pairs, losses = ws.train_views(frames, [b.ptr for b in bufs])
            for v in range(V):
                g2 = ws.download(gs.BUF_GRAD2D, np.float32, 9 * n, view=v).reshape(n, 9)
                scale = np.abs(s_g2[v]).max(axis=0) + 1e-30
                assert (np.abs(g2 - s_g2[v]).max(axis=0) / scale).max() < 1e-5, v
                assert losses[v] == pytest.approx(s_loss[v], rel=1e-6)
            gp, gsh = _first_step_gradients(ws, n)
            stats = ws.download(gs.BUF_STATS, np.float32, n * 4).reshape(n, 4)
            after = ws.store(n)[0]
    finally:
        for b in bufs:
            b.free()
    np.testing.assert_array_equal(after, params)
    mean_gp, mean_gsh = np.mean(s_gp, axis=0), np.mean(s_gsh, axis=0)
    for cols in ([0, 1, 2], [3], [4, 5, 6], [8, 9, 10, 11], [12, 13, 14]):
        scale = np.abs(mean_gp[:, cols]).max()
        assert np.abs(gp[:, cols] - mean_gp[:, cols]).max() <= 1e-5 * scale, cols
    assert np.abs(gsh - mean_gsh).max() <= 1e-5 * np.abs(mean_gsh).max()
    np.testing.assert_array_equal(stats[:, 1], np.sum(s_vis, axis=0).astype(np.float32))


def test_gradients_at_config2_size_against_frozen_branch_differences():
    """FP64 central differences (h = 1e-4) on 1000 sampled parameters of the
    200k-gaussian SH3 1280x720 view, against the GPU's first-step gradient
    (train_selftest.cpp:356-486's check_gradients at this size).  The
    differences are of the fp64 loss with the compositor's branches frozen as
    in the unperturbed render (oracle.c orc_fd_frozen64): at 1280x720 a 1e-4
    move of a splat makes some pixel centre cross its 1/255 contour for most
    splats, a jump the analytic gradient (the reference's and this one) leaves
    out.  Samples: gaussians that drew (nonzero 2D gradient), every parameter
    column and the 45 SH rest values.  Rule per sample: |g - fd| <= 0.05 |fd|
    + 1e-4 max|fd| of its group + 2e-6 x (64 x 48) / (W x H) (the reference's
    absolute floor, scaled from its 64x48 test image to this mean over 3WH
    values); at most 5% of a group's samples off."""
    case("gradients vs frozen-branch central differences at 200k / 1280x720")
    n, deg, w, h, seed = 200_000, 3, 1280, 720, 42
    params, sh = gs.synthetic_scene(n, deg, seed)
    cam = gs.look_at_camera((0.0, 0.5, 3.0), (0, 0, 0), w, h)
    target = smooth_target(w, h, seed)
    bg, lam = (0.1, 0.2, 0.3), 0.2
    with gs.Workspace(n, deg, w, h) as ws:
        ws.load(params, sh)
        ws.set_target(target)
        ws.train_step(gs.make_frame(cam, n, deg, background=bg, tp=frozen(lam)), want_losses=False)
        g2 = ws.download(gs.BUF_GRAD2D, np.float32, 9 * n).reshape(n, 9)
        gp, gsh = _first_step_gradients(ws, n)
    drew = np.nonzero(np.any(g2 != 0, axis=1))[0]
    rng = np.random.default_rng(4004)
    cols = np.array([0, 1, 2, 3, 4, 5, 6, 8, 9, 10, 11, 12, 13, 14] + [16 + k for k in range(45)])
    gi = rng.choice(drew, 1000, replace=False).astype(np.uint32)
    ci = rng.choice(cols, 1000).astype(np.uint32)
    _, fd = orc.fd_frozen64(params.astype(np.float64), sh.astype(np.float64), cam, deg, target, lam, gi, ci, h=1e-4,
                            bg=bg)
    g = np.where(ci < 16, gp[gi, np.minimum(ci, 15)], gsh.reshape(n, -1)[gi, np.maximum(ci.astype(int) - 16, 0)])
    assert not np.isnan(fd).any()
    floor = 2e-6 * (64 * 48) / (w * h)
    groups = {"position": [0, 1, 2], "opacity": [3], "scale": [4, 5, 6], "rotation": [8, 9, 10, 11],
              "dc": [12, 13, 14], "sh rest": list(range(16, 61))}
    for gname, cs in groups.items():
        sel = np.isin(ci, cs)
        big = np.abs(fd[sel]).max()
        err = np.abs(g[sel] - fd[sel])
        off = err > 0.05 * np.abs(fd[sel]) + 1e-4 * big + floor
        rel = np.median(err / (np.abs(fd[sel]) + 1e-4 * big))
        print(f"[info] {gname}: {sel.sum()} samples, {off.sum()} off, median error {rel:.2g} of |fd|, largest |fd| "
              f"{big:.3g}", flush=True)
        assert sel.sum() >= 10 and big > 0
        assert off.sum() <= sel.sum() // 20, f"{gname}: {off.sum()} of {sel.sum()} off"


def test_needle_splats_forward_and_backward_match_oracle():
    """Needle splats (one axis 1-3, the others 1e-3 of the scene scale, seen
    at 1280x720: major-axis variance up to ~1e6 px^2) exercise the
    rasterizers' quadrant cull far out along the axis, where the conic's terms
    cancel: every pixel must still composite every splat the oracle (which
    culls nothing) composites -- each clear pixel within twice its evaluation
    spread -- and every backward sum must be within its share."""
    case("needle splats, 1280x720")
    from parity_check import check_forward
    n, deg, w, h = 3000, 1, 1280, 720
    rng = np.random.default_rng(121)
    params, sh = make_scene(n, deg, 121, spread=0.6)
    params[:, 4] = np.log(rng.uniform(1.0, 3.0, n))
    params[:, 5:7] = np.log(rng.uniform(5e-4, 2e-3, (n, 2)))
    params[:, 3] = rng.uniform(-2.0, 1.0, n)
    cam = gs.look_at_camera((0.0, 0.5, 3.0), (0, 0, 0), w, h)
    target = smooth_target(w, h, 121)
    bg = (0.0, 0.0, 0.0)
    with gs.Workspace(n, deg, w, h) as ws:
        ws.load(params, sh)
        frame = gs.make_frame(cam, n, deg, background=bg, tp=frozen(0.2))
        pairs, losses = _pipeline(ws, frame, target, 0.2)
        rec, foot, tiles, keys, order, pt, pg, ranges = gpu_state(ws, n, pairs)
        img = ws.download(gs.BUF_IMAGE, np.float32).reshape(-1, 4)
        last = ws.download(gs.BUF_LAST, np.uint32, w * h)
        dldc = ws.download(gs.BUF_DLDC, np.float32).reshape(-1, 4)
        ws.backward(frame)
        g2 = ws.download(gs.BUF_GRAD2D, np.float32, 9 * n).reshape(n, 9)
    cov = foot[:, :3]
    major = 0.5 * (cov[:, 0] + cov[:, 2]) + np.sqrt(0.25 * (cov[:, 0] - cov[:, 2]) ** 2 + cov[:, 1] ** 2)
    print(f"[info] needles: {int((tiles > 0).sum())} visible, {pairs} pairs, largest major-axis variance "
          f"{major.max():.3g} px^2", flush=True)
    assert major.max() > 2e5
    check_forward(rec, pg, ranges, img, last, w, h, "needles", max_ambiguous=0.75)
    # Far along a needle's axis fp32 evaluation is ill-conditioned nearly everywhere (terms of ~1e5
    # cancel to a power of ~-1): every sum must be within its share, however many need it.
    check_backward(rec, pg, ranges, img, last, dldc, bg, w, h, g2, tiles, "needles", cap=1.0)
