"""Pins the CPU oracle against what fixes it independently: libm, scipy's
spherical harmonics, closed forms, brute force, a second SSIM written with
scipy.ndimage, and central finite differences.  No GPU."""
import math

import numpy as np
import pytest
from scipy import ndimage, special

from helpers import make_camera, make_scene, noise_target, orc, smooth_target, gs


def _ulp_diff(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


def test_expf_within_two_ulp():
    xs = np.concatenate([np.linspace(-87.0, 88.0, 20001), np.linspace(-1.0, 1.0, 5001),
                         np.random.default_rng(1).uniform(-30, 30, 5000)]).astype(np.float32)
    got = np.array([orc.expf(float(x)) for x in xs], np.float32)
    want = np.exp(xs.astype(np.float64)).astype(np.float32)
    assert _ulp_diff(got, want).max() <= 2


def test_logf_within_two_ulp():
    rng = np.random.default_rng(2)
    xs = np.concatenate([np.exp(rng.uniform(-60, 60, 20000)), np.linspace(0.5, 2.0, 5001),
                         np.linspace(1.0, 255.0, 5001)]).astype(np.float32)
    got = np.array([orc.logf(float(x)) for x in xs], np.float32)
    want = np.log(xs.astype(np.float64)).astype(np.float32)
    # absolute error near log(1) = 0 is what matters there
    ok = (_ulp_diff(got, want) <= 2) | (np.abs(got - want) <= 2e-7)
    assert ok.all()


def _real_sh(l, m, polar, azimuth):
    if hasattr(special, "sph_harm_y"):
        Y = special.sph_harm_y(l, abs(m), polar, azimuth)
    else:
        Y = special.sph_harm(abs(m), l, azimuth, polar)
    if m > 0:
        return math.sqrt(2.0) * Y.real
    if m < 0:
        return math.sqrt(2.0) * Y.imag
    return Y.real


def test_sh_basis_matches_scipy_real_harmonics():
    """The 3DGS basis is the real SH built from scipy's Condon-Shortley
    Y_l^m: sqrt(2) Re Y_l^m (m > 0), Y_l^0, sqrt(2) Im Y_l^|m| (m < 0),
    ordered m = -l..l within each degree."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        polar, azimuth = math.acos(d[2]), math.atan2(d[1], d[0])
        b = orc.sh_basis64(d, 3)
        k = 0
        for l in range(4):
            for m in range(-l, l + 1):
                assert b[k] == pytest.approx(_real_sh(l, m, polar, azimuth), abs=1e-12), (l, m)
                k += 1


def _single_splat(opacity=0.95, scale=0.1, pos=(0.0, 0.0, 0.0), colour=(1.0, 0.0, 0.0)):
    p = np.zeros((1, 16), np.float64)
    p[0, 0:3] = pos
    p[0, 3] = math.log(opacity / (1 - opacity))
    p[0, 4:7] = math.log(scale)
    p[0, 7] = 1.0
    p[0, 8] = 1.0
    # DC such that colour = C0 dc + 0.5
    p[0, 12:15] = (np.asarray(colour) - 0.5) / 0.28209479177387814
    return p


def test_single_splat_falloff_known_answer():
    """engine/tests/test_splats.cpp:188-210: an isotropic splat of scale 0.1 at
    distance 3 falls off as o exp(-0.5 d^2 / (sigma_px^2 + 0.3))."""
    w = h = 64
    cam = gs.look_at_camera((0, 0, 3), (0, 0, 0), w, h)
    p = _single_splat()
    proj = orc.project64(p, None, cam, 0)[0]
    sigma = cam.fx * 0.1 / 3.0
    var = sigma * sigma + 0.3
    assert proj[0] == 1.0
    assert proj[1] == pytest.approx(32.0) and proj[2] == pytest.approx(32.0)
    assert proj[5] == pytest.approx(1 / var, rel=1e-6) and proj[7] == pytest.approx(1 / var, rel=1e-6)
    assert abs(proj[6]) < 1e-9
    rgb, T = orc.render64(p, None, cam, 0)
    for dx in (0, 2, 4):
        px, py = 32 + dx + 0.5 - 32.0, 0.5
        alpha = min(0.95 * math.exp(-0.5 * (px * px + py * py) / var), 0.99)
        assert rgb[32 * w + 32 + dx, 0] == pytest.approx(alpha, rel=1e-9)
        assert rgb[32 * w + 32 + dx, 1] == pytest.approx(0.0, abs=1e-12)
    assert rgb[2 * w + 2].max() < 1e-12


def test_depth_order_known_answer():
    """engine/tests/test_splats.cpp:212-229: the nearer of two splats on the
    axis dominates the centre pixel."""
    w = h = 64
    cam = gs.look_at_camera((0, 0, 3), (0, 0, 0), w, h)
    for red_in_front in (True, False):
        rz = 0.5 if red_in_front else -0.5
        p = np.concatenate([_single_splat(0.9, 0.1, (0, 0, rz), (1, 0, 0)),
                            _single_splat(0.9, 0.1, (0, 0, -rz), (0, 0, 1))])
        rgb, _ = orc.render64(p, None, cam, 0)
        c = rgb[32 * w + 32]
        if red_in_front:
            assert c[0] > 180 / 255 and c[2] < 60 / 255
        else:
            assert c[2] > 180 / 255 and c[0] < 60 / 255


def test_project32_agrees_with_project64():
    params, sh = make_scene(400, 3, 11)
    cam = make_camera(96, 72, 0.4)
    rec, foot, tiles, keys = orc.project32(params, sh, cam, 3)
    ref = orc.project64(params, sh, cam, 3)
    vis = tiles > 0
    assert vis.sum() > 100
    assert np.all(ref[vis, 0] == 1.0)
    np.testing.assert_allclose(rec[vis, 0:2], ref[vis, 1:3], rtol=1e-5, atol=1e-4)
    con32, con64 = rec[vis][:, [2, 3, 4]], ref[vis][:, 5:8]
    scale = np.abs(con64).max(axis=1, keepdims=True)
    assert np.max(np.abs(con32 - con64) / scale) < 1e-4
    np.testing.assert_allclose(rec[vis, 5], ref[vis, 4], rtol=1e-6)
    np.testing.assert_allclose(rec[vis, 6:9], ref[vis, 8:11], rtol=1e-5, atol=1e-5)
    z = keys[vis].view(np.float32)
    np.testing.assert_allclose(z, ref[vis, 3], rtol=1e-6)


def test_depth16_key_matches_reference_formula():
    params, sh = make_scene(300, 1, 12)
    cam = make_camera(64, 48, 1.0)
    _, _, tiles, keys = orc.project32(params, sh, cam, 1, depth16=True)
    ref = orc.project64(params, sh, cam, 1)
    vis = tiles > 0
    want = np.floor(np.clip(np.log2(ref[vis, 3] / 0.02) / np.log2(1000 / 0.02), 0, 1) * 65535)
    assert np.max(np.abs(keys[vis].astype(np.int64) - want.astype(np.int64))) <= 1
    assert np.all(keys[~vis] == 0xFFFFFFFF)


@pytest.mark.parametrize("side", [16, 32, 64])
def test_footprint_against_brute_force(side):
    """engine/tests/test_trainer.cpp:235-289: every tile holding a pixel with
    d^T conic d <= 2 ln(255 o) is emitted, once, and few others -- for the
    reference's 16 px tiles and the kernels' coarser 32 / 64 px pair bins."""
    w, h = 160, 112
    params, sh = make_scene(300, 1, 13)
    params[::2, 4] += 1.5  # half of them stretched
    cam = make_camera(w, h, 0.0)
    rec, foot, tiles, _ = orc.project32(params, sh, cam, 1)
    counts = orc.bin_counts32(rec, foot, tiles, w, h, side)
    if side == 16:
        np.testing.assert_array_equal(counts, tiles)
    tx = (w + side - 1) // side
    needed = emitted = 0
    ys, xs = np.mgrid[0:h, 0:w]
    for i in np.nonzero(tiles)[0]:
        q = rec[i].astype(np.float64)
        t = 2.0 * math.log(255.0 * q[5])
        dx, dy = q[0] - (xs + 0.5), q[1] - (ys + 0.5)
        inside = q[2] * dx * dx + 2 * q[3] * dx * dy + q[4] * dy * dy <= t
        want = set(((ys[inside] // side) * tx + xs[inside] // side).tolist())
        got = orc.footprint_tiles32(*[float(v) for v in (q[0], q[1], foot[i, 0], foot[i, 1], foot[i, 2],
                                                          foot[i, 3])], w, h, side)
        assert len(got) == counts[i]
        assert len(set(got.tolist())) == len(got), "duplicate tile"
        assert want <= set(got.tolist()), f"gaussian {i}: missing tiles"
        needed += len(want)
        emitted += len(got)
    assert needed > 100
    assert emitted <= needed + needed // 20


@pytest.mark.parametrize("shift", [1, 2])
def test_coarse_bins_hold_their_tiles(shift):
    """A bin of 16 << shift pixels lists a splat wherever one of its 16 px
    tiles does (so a raster tile walking its bin's list meets every splat of
    its own list, in the same depth order), and few more."""
    w, h, side = 200, 136, 16 << shift
    params, sh = make_scene(400, 2, 17)
    params[::3, 4:6] += 1.2
    cam = make_camera(w, h, 0.4)
    rec, foot, tiles, keys = orc.project32(params, sh, cam, 2, depth16=True)
    tx, bx = (w + 15) // 16, (w + side - 1) // side
    extra = total = 0
    for i in np.nonzero(tiles)[0]:
        args = [float(v) for v in (rec[i, 0], rec[i, 1], *foot[i])]
        fine = orc.footprint_tiles32(*args, w, h, 16)
        coarse = set(orc.footprint_tiles32(*args, w, h, side).tolist())
        want = {int((t // tx) >> shift) * bx + int((t % tx) >> shift) for t in fine.tolist()}
        assert want <= coarse, f"gaussian {i}"
        extra += len(coarse - want)
        total += len(coarse)
    assert total > 300 and extra <= total // 50
    counts = orc.bin_counts32(rec, foot, tiles, w, h, side)
    pt, pg, ranges = orc.bin32(rec, foot, counts, keys, w, h, side)
    fine_pt, fine_pg, fine_ranges = orc.bin32(rec, foot, tiles, keys, w, h)
    for t in range(tx * ((h + 15) // 16)):
        b = ((t // tx) >> shift) * bx + ((t % tx) >> shift)
        own = fine_pg[fine_ranges[t, 0]:fine_ranges[t, 1]].tolist()
        lst = pg[ranges[b, 0]:ranges[b, 1]].tolist()
        pos = {g: k for k, g in enumerate(lst)}
        assert all(g in pos for g in own) and [pos[g] for g in own] == sorted(pos[g] for g in own)


@pytest.mark.parametrize("side", [16, 32])
def test_bin_order_is_tile_depth_index(side):
    w, h = 64, 48
    params, sh = make_scene(400, 1, 37)
    cam = make_camera(w, h, 0.0)
    rec, foot, tiles, keys = orc.project32(params, sh, cam, 1, depth16=True)
    counts = orc.bin_counts32(rec, foot, tiles, w, h, side)
    pt, pg, ranges = orc.bin32(rec, foot, counts, keys, w, h, side)
    pairs = []
    for i in np.nonzero(tiles)[0]:
        for t in orc.footprint_tiles32(*[float(v) for v in (rec[i, 0], rec[i, 1], *foot[i])], w, h, side):
            pairs.append((int(t), int(keys[i]), int(i)))
    want = sorted(pairs)
    assert len(want) > (500 if side == 16 else 200) and len({p[1] for p in want}) > 100
    assert [p[0] for p in want] == pt.tolist()
    assert [p[2] for p in want] == pg.tolist()
    T = ((w + side - 1) // side) * ((h + side - 1) // side)
    for t in range(T):
        s, e = ranges[t]
        assert np.all(pt[s:e] == t)
        assert (e - s) == sum(1 for p in want if p[0] == t)


def test_forward32_matches_brute_force_render64():
    w, h = 96, 72
    params, sh = make_scene(300, 2, 21)
    cam = make_camera(w, h, 2.0)
    bg = (0.1, 0.2, 0.3)
    rec, foot, tiles, keys = orc.project32(params, sh, cam, 2)
    pt, pg, ranges = orc.bin32(rec, foot, tiles, keys, w, h)
    img, last = orc.forward32(rec, pg, ranges, w, h)
    out = img[:, :3] + img[:, 3:4] * np.asarray(bg, np.float32)
    ref, T = orc.render64(params, sh, cam, 2, bg)
    err = np.abs(out - ref).max(axis=1)
    assert np.quantile(err, 0.999) < 1e-4
    assert err.max() < 2.5 / 255


def test_threshold_flags_mark_only_decisions_at_the_edge():
    """The ambiguity flags of forward32/backward32 (used to bound comparisons
    with other fp32 evaluation orders) fire exactly where a splat's alpha
    sits on the 1/255 cut, and nowhere else."""
    w = h = 32
    d2 = 2 * 15.75 ** 2  # pixel (0, 0) seen from a splat at (16.25, 16.25)
    o = np.float32((1.0 / 255.0) / math.exp(-0.5 * 0.01 * d2))
    rec = np.zeros((2, 12), np.float32)
    rec[0, :9] = [16.25, 16.25, 0.01, 0.0, 0.01, o, 0.9, 0.5, 0.1]
    rec[1, :9] = [28.25, 28.25, 0.5, 0.0, 0.5, 0.8, 0.2, 0.3, 0.4]
    pg = np.array([0, 0, 0, 0, 1], np.uint32)  # tiles 0..2: splat 0; tile 3: splats 0, 1
    ranges = np.array([0, 1, 1, 2, 2, 3, 3, 5, 0, 0], np.uint32)
    img, last, amb = orc.forward32(rec, pg, ranges, w, h, flags=True)
    assert amb.sum() == 1 and amb[0]
    dldc = np.ones((w * h, 4), np.float32)
    g, tainted = orc.backward32(rec, pg, ranges, img, last, dldc, (0.0, 0.0, 0.0), w, h, flags=True)
    assert tainted.tolist() == [True, False]
    assert np.abs(g[1]).max() > 0


def _ssim_numpy(x, y, w, h):
    k = np.exp(-np.arange(-5, 6) ** 2 / (2 * 1.5 ** 2))
    k /= k.sum()
    blur = lambda a: ndimage.correlate1d(ndimage.correlate1d(a, k, axis=0, mode="constant"), k, axis=1,
                                         mode="constant")
    total = 0.0
    for c in range(3):
        a, b = x[:, c].reshape(h, w), y[:, c].reshape(h, w)
        mx, my = blur(a), blur(b)
        vx, vy, cxy = blur(a * a) - mx * mx, blur(b * b) - my * my, blur(a * b) - mx * my
        s = (2 * mx * my + 1e-4) * (2 * cxy + 9e-4) / ((mx * mx + my * my + 1e-4) * (vx + vy + 9e-4))
        total += s.sum()
    return total


def test_ssim64_against_scipy_and_identity():
    w, h = 40, 30
    rng = np.random.default_rng(5)
    x, y = rng.uniform(size=(w * h, 3)), rng.uniform(size=(w * h, 3))
    assert orc.ssim64(x, y, w, h) == pytest.approx(_ssim_numpy(x, y, w, h), rel=1e-12)
    s, g = orc.ssim64(x, x, w, h, grad=True)
    assert s == pytest.approx(3 * w * h, rel=1e-12)
    assert np.abs(g).max() < 1e-9


def test_loss_image_gradient_by_central_differences():
    w, h = 24, 20
    rng = np.random.default_rng(6)
    x, y = rng.uniform(0.2, 0.8, size=(w * h, 3)), rng.uniform(0.2, 0.8, size=(w * h, 3))
    L, g = orc.loss_image64(x, y, w, h, 0.2, grad=True)
    eps = 1e-6
    for idx in rng.choice(w * h * 3, 60, replace=False):
        if abs(x.flat[idx] - y.flat[idx]) < 1e-3:
            continue
        xp, xm = x.copy(), x.copy()
        xp.flat[idx] += eps
        xm.flat[idx] -= eps
        fd = (orc.loss_image64(xp, y, w, h, 0.2) - orc.loss_image64(xm, y, w, h, 0.2)) / (2 * eps)
        assert g.flat[idx] == pytest.approx(fd, rel=1e-5, abs=1e-10)


def test_backward64_by_central_differences():
    """FP64 central differences of the fp64 loss against the analytic backward
    (the reference's check_gradients, train_selftest.cpp:356-486, with its
    per-value rule |g - fd| <= 0.05 |fd| + 2e-6 and at most 5% off per group)."""
    w, h = 64, 48
    params, sh = make_scene(10, 2, 44, spread=0.4)
    cam = make_camera(w, h, 0.3)
    target = smooth_target(w, h, 7)
    p64, s64 = params.astype(np.float64), sh.astype(np.float64)
    L, g2, gp, gsh = orc.backward64(p64, s64, cam, 2, target, 0.2, bg=(0.2, 0.1, 0.0))
    eps = 1e-5
    groups = {"position": [0, 1, 2], "opacity": [3], "scale": [4, 5, 6], "rotation": [8, 9, 10, 11],
              "dc": [12, 13, 14]}
    for name, cols in groups.items():
        bad = checked = 0
        for i in range(params.shape[0]):
            for c in cols:
                pp, pm = p64.copy(), p64.copy()
                pp[i, c] += eps
                pm[i, c] -= eps
                fd = (orc.loss64(pp, s64, cam, 2, target, 0.2, bg=(0.2, 0.1, 0.0)) -
                      orc.loss64(pm, s64, cam, 2, target, 0.2, bg=(0.2, 0.1, 0.0))) / (2 * eps)
                checked += 1
                if abs(gp[i, c] - fd) > 0.05 * abs(fd) + 2e-6:
                    bad += 1
        assert bad <= checked // 20, f"{name}: {bad}/{checked} off"
    bad = checked = 0
    for i in range(params.shape[0]):
        for k in range(sh.shape[1]):
            for c in range(3):
                sp, sm = s64.copy(), s64.copy()
                sp[i, k, c] += eps
                sm[i, k, c] -= eps
                fd = (orc.loss64(p64, sp, cam, 2, target, 0.2, bg=(0.2, 0.1, 0.0)) -
                      orc.loss64(p64, sm, cam, 2, target, 0.2, bg=(0.2, 0.1, 0.0))) / (2 * eps)
                checked += 1
                if abs(gsh[i, k, c] - fd) > 0.05 * abs(fd) + 2e-6:
                    bad += 1
    assert bad <= checked // 20, f"sh: {bad}/{checked} off"
    assert np.abs(gp[:, [0, 3, 4, 12]]).max() > 1e-4  # the check is not vacuous


def test_adam_closed_form():
    rng = np.random.default_rng(8)
    g = rng.normal(size=100).astype(np.float32)
    p = rng.normal(size=100).astype(np.float32)
    zeros = np.zeros(100, np.float32)
    # lr 0: parameters stay, m / (1 - beta1) is the gradient (trainer.cpp:1802-1811)
    p1, m, v = orc.adam(p, g, zeros, zeros, zeros, bias1=1 / (1 - 0.9), bias2=1 / (1 - 0.999))
    np.testing.assert_array_equal(p1, p)
    np.testing.assert_allclose(m / (1 - 0.9), g, rtol=1e-6)
    np.testing.assert_allclose(v, g * g * (np.float32(1) - np.float32(0.999)), rtol=1e-6)
    # first step with bias correction moves each parameter by lr sign(g)
    lr = np.full(100, 1e-3, np.float32)
    p1, _, _ = orc.adam(p, g, zeros, zeros, lr, bias1=1 / (1 - 0.9), bias2=1 / (1 - 0.999))
    np.testing.assert_allclose(p1, p - 1e-3 * np.sign(g), rtol=1e-5, atol=1e-7)


def test_densify_mark_predicate():
    rng = np.random.default_rng(9)
    n = 2000
    params = np.zeros((n, 16), np.float32)
    params[:, 3] = rng.normal(-2, 3, size=n)
    params[:, 4:7] = rng.normal(-4, 1, size=(n, 3))
    params[:, 7] = (rng.uniform(size=n) > 0.1).astype(np.float32)
    stats = np.zeros((n, 4), np.float32)
    stats[:, 1] = rng.integers(0, 5, size=n)
    stats[:, 0] = stats[:, 1] * rng.uniform(0, 4e-4, size=n)
    action, release, prune = orc.densify_mark(params, stats, 2e-4, 0.01, 0.005)
    alive = params[:, 7] >= 0.5
    op = 1 / (1 + np.exp(-params[:, 3].astype(np.float64)))
    want_prune = alive & (op < 0.005)
    np.testing.assert_array_equal(prune.astype(bool), want_prune)
    np.testing.assert_array_equal(release.astype(bool), ~alive | want_prune)
    grad = np.divide(stats[:, 0], stats[:, 1], out=np.zeros(n, np.float32), where=stats[:, 1] > 0)
    flag = alive & ~want_prune & (stats[:, 1] > 0) & (grad >= 2e-4)
    big = np.exp(params[:, 4:7].max(axis=1).astype(np.float64)) > 0.01
    want = np.where(flag, np.where(big, 2, 1), 0)
    np.testing.assert_array_equal(action, want)


def test_backward32_tiles_matches_backward64():
    """The tile-list fp32 backward on the fp32 forward's state reproduces the
    brute-force fp64 analytic per-gaussian 2D gradient."""
    w, h = 64, 48
    params, sh = make_scene(60, 1, 45, spread=0.5)
    cam = make_camera(w, h, 0.7)
    bg = (0.2, 0.1, 0.0)
    target = smooth_target(w, h, 8)
    rec, foot, tiles, keys = orc.project32(params, sh, cam, 1)
    pt, pg, ranges = orc.bin32(rec, foot, tiles, keys, w, h)
    img, last = orc.forward32(rec, pg, ranges, w, h)
    x = img[:, :3].astype(np.float64) + img[:, 3:4].astype(np.float64) * np.asarray(bg)
    y = target.reshape(-1, 4)[:, :3] / 255.0
    _, dldc = orc.loss_image64(x, y, w, h, 0.2, grad=True)
    d4 = np.zeros((w * h, 4), np.float32)
    d4[:, :3] = dldc
    g32 = orc.backward32(rec, pg, ranges, img, last, d4, bg, w, h)
    _, g64, _, _ = orc.backward64(params.astype(np.float64), sh.astype(np.float64), cam, 1, target, 0.2, bg=bg)
    for c in range(9):
        scale = np.abs(g64[:, c]).max()
        assert scale > 0
        assert np.abs(g32[:, c] - g64[:, c]).max() <= 1e-3 * scale, c


def _rodrigues(w):
    w = np.asarray(w, np.float64)
    th = np.linalg.norm(w)
    if th == 0.0:
        return np.eye(3)
    k = w / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K


def moved_camera(cam, d=(0.0, 0.0, 0.0), w=(0.0, 0.0, 0.0)):
    """The trainer's pose correction applied to a camera: world -> camera
    [R | t] becomes [E R | E t + d], E the rotation by w (trainer.cpp:1294-1311)."""
    R = np.array(cam.rotation, np.float64).reshape(3, 3)
    t = np.array(cam.translation, np.float64)
    E = _rodrigues(w)
    R2, t2 = E @ R, E @ t + np.asarray(d, np.float64)
    out = gs.Camera()
    ctypes_copy = [("fx", cam.fx), ("fy", cam.fy), ("cx", cam.cx), ("cy", cam.cy), ("width", cam.width),
                   ("height", cam.height)]
    for name, v in ctypes_copy:
        setattr(out, name, v)
    out.rotation[:] = [float(v) for v in R2.reshape(-1)]
    out.translation[:] = [float(v) for v in t2]
    out.position[:] = [float(v) for v in -R2.T @ t2]
    return out


def test_pose_adjoint64_by_central_differences():
    """The summed pose gradient against central differences of the fp64 loss
    as the camera shifts (d) and turns (w) about its centre, SH degree 0 (the
    gradient leaves out colour's dependence on the view direction).  A camera
    move shifts every splat's alpha cutoff at once, a jump the analytic
    gradient leaves out (test_trainer.cpp:409-412), so the scene is a few
    wide, faint splats whose 1/255 contour lies outside the image and whose
    product of (1 - alpha) stays far above the 1e-4 stop: the loss is then
    smooth in the pose; the loss is D-SSIM alone (lambda 1: no kinks of
    |e| where a pixel's error changes sign).  The camera is stored in fp32,
    so the step is 1e-3: the moved camera's rounding (~1e-7) stays far below
    the difference, and the O(h^2) error measured ~1e-5 relative."""
    w, h = 64, 48
    rng = np.random.default_rng(48)
    n = 6
    params = np.zeros((n, 16), np.float64)
    params[:, 0:3] = 0.3 * rng.uniform(-1, 1, (n, 3))
    params[:, 4:7] = np.log(2.0 + 0.8 * rng.uniform(-1, 1, (n, 3)))
    q = rng.uniform(-1, 1, (n, 4))
    q[:, 0] += 1.5
    params[:, 8:12] = q
    params[:, 3] = np.log(0.4 / 0.6) + 0.3 * rng.uniform(-1, 1, n)
    params[:, 7] = 1.0
    params[:, 12:15] = 0.8 * rng.uniform(-1, 1, (n, 3))
    sh = np.zeros((n, 0, 3))
    cam = make_camera(w, h, 0.7)
    target = smooth_target(w, h, 9)
    bg = (0.2, 0.1, 0.0)
    # every splat above the cutoff over the whole image (the premise)
    rec = orc.project64(params, sh, cam, 0)
    yy, xx = np.mgrid[0:h, 0:w] + 0.5
    for r in rec:  # visible, mx, my, depth, opacity, conic abc, ...
        dx, dy = r[1] - xx, r[2] - yy
        power = -0.5 * (r[5] * dx * dx + r[7] * dy * dy) - r[6] * dx * dy
        assert r[0] == 1 and (r[4] * np.exp(power)).min() > 2.0 / 255.0
    _, g2, _, _ = orc.backward64(params, sh, cam, 0, target, 1.0, bg=bg)
    grad = orc.pose_adjoint64(params, sh, cam, 0, g2).sum(0)
    eps = 1e-3
    fd = np.zeros(6)
    for k in range(6):
        e = np.zeros(6)
        e[k] = eps
        lp = orc.loss64(params, sh, moved_camera(cam, e[:3], e[3:]), 0, target, 1.0, bg=bg)
        lm = orc.loss64(params, sh, moved_camera(cam, -e[:3], -e[3:]), 0, target, 1.0, bg=bg)
        fd[k] = (lp - lm) / (2 * eps)
    assert np.abs(fd).min() > 0.02 * np.abs(fd).max()  # every component is exercised
    np.testing.assert_allclose(grad, fd, rtol=5e-4)


def test_pose_adjoint64_identity_with_gaussian_gradients():
    """The reference's check of the pose gradient (test_trainer.cpp:404-464):
    moving the camera is moving every gaussian the other way, so dL/dd =
    W sum dL/dp and dL/dw = W sum((p - c) x dL/dp + tau), tau the
    orientation's share read off the quaternion gradient.  SH degree 0."""
    w, h = 96, 72
    params, sh = make_scene(150, 0, 47)
    cam = make_camera(w, h, 1.9)
    target = smooth_target(w, h, 13)
    p64, s64 = params.astype(np.float64), sh.astype(np.float64)
    _, g2, gp, _ = orc.backward64(p64, s64, cam, 0, target, 0.2, bg=(0.2, 0.3, 0.4))
    got = orc.pose_adjoint64(p64, s64, cam, 0, g2).sum(0)
    R = np.array(cam.rotation, np.float64).reshape(3, 3)
    c = np.array(cam.position, np.float64)
    sum_d = gp[:, 0:3].sum(0)
    rel = p64[:, 0:3] - c
    q = p64[:, 8:12]
    ln = np.linalg.norm(q, axis=1, keepdims=True)
    qw, u = q[:, :1] / ln, q[:, 1:] / ln
    gw, gu = gp[:, 8:9] * ln, gp[:, 9:12] * ln
    tau = 0.5 * (-gw * u + qw * gu + np.cross(u, gu))
    sum_w = (np.cross(rel, gp[:, 0:3]) + tau).sum(0)
    want = np.concatenate([R @ sum_d, R @ sum_w])
    assert np.abs(want).max() > 0
    # (the camera centre is stored in fp32)
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max(), (got, want)


def test_exposure_sums64_known_answers():
    """A render whose colour is the photo divided by k where both are opaque:
    the least-squares gain sum(C P) / sum(C^2) is k per channel; pixels with
    photo alpha <= 0.95 or transmittance >= 0.05 are left out, pixels with
    alpha > 0 are counted (train_forward.comp:142-151)."""
    rng = np.random.default_rng(3)
    n = 500
    target = rng.integers(1, 256, size=(n, 4), dtype=np.uint8)
    target[:, 3] = rng.choice([0, 100, 242, 243, 255], size=n)
    k = np.array([1.25, 0.8, 1.1])
    image = np.zeros((n, 4))
    image[:, :3] = target[:, :3] / 255.0 / k
    image[:, 3] = rng.choice([0.0, 0.049, 0.05, 0.3], size=n)
    s = orc.exposure_sums64(image, target)
    fit = (target[:, 3] >= 243) & (image[:, 3] < 0.05)
    assert s[3] == fit.sum() > 50
    assert s[7] == (target[:, 3] > 0).sum()
    np.testing.assert_allclose(s[4:7] / s[0:3], k, rtol=1e-12)
    y = orc.target_colour64(target, bg=(0.5, 0.25, 0.0), masked=True, exposure=1.0 / k)
    a = target[:, 3:4] / 255.0
    np.testing.assert_allclose(y, (0.5, 0.25, 0.0) * (1 - a) + image[:, :3] * a, rtol=1e-12)


def test_render64_tile_lists_equal_brute_force():
    """render64 walks per-tile lists of the splats whose 1/255 ellipse box
    meets the tile; the loop over every visible gaussian gives the same
    image bit for bit (same operations in the same order), including wide
    splats past the image edge and faint ones near the 1/255 cut."""
    for n, deg, w, h, seed, widen in ((300, 2, 96, 72, 5, 0.0), (150, 0, 70, 45, 6, 1.5), (60, 3, 33, 17, 7, 3.0)):
        params, sh = make_scene(n, deg, seed)
        params[: n // 3, 4:7] += widen
        params[n // 3: n // 2, 3] = -5.4  # opacity ~ 1/220: the ellipse is tiny or empty
        cam = make_camera(w, h, 0.7)
        rgb, _ = orc.render64(params, sh, cam, deg, bg=(0.1, 0.2, 0.3))
        brute = orc.render64_brute(params, sh, cam, deg, bg=(0.1, 0.2, 0.3))
        np.testing.assert_array_equal(rgb, brute)


def test_fd_frozen64_matches_the_analytic_backward():
    """Central differences with the compositor's branches frozen (the
    function the analytic backward differentiates) agree with the fp64
    analytic gradient to 1e-5 of each group's largest value plus 1e-4 of the
    difference itself, on every parameter of every gaussian of a 120-splat
    scene (SH degree 2, background, lambda 0.2); and plain central
    differences of the loss (jumps included) agree with them wherever no
    1/255 contour crosses a pixel centre (most samples at this size)."""
    w, h = 80, 60
    params, sh = make_scene(120, 2, 45, spread=0.5)
    cam = make_camera(w, h, 0.9)
    target = smooth_target(w, h, 8)
    bg = (0.2, 0.1, 0.0)
    p64, s64 = params.astype(np.float64), sh.astype(np.float64)
    L, _, gp, gsh = orc.backward64(p64, s64, cam, 2, target, 0.2, bg=bg)
    cols = [0, 1, 2, 3, 4, 5, 6, 8, 9, 10, 11, 12, 13, 14]
    gi = np.repeat(np.arange(120), len(cols) + 6)
    ci = np.tile(np.array(cols + [16, 17, 18, 16 + 3 * 7, 16 + 3 * 7 + 1, 16 + 3 * 7 + 2]), 120)
    L2, fd = orc.fd_frozen64(p64, s64, cam, 2, target, 0.2, gi, ci, h=1e-4, bg=bg)
    assert L2 == pytest.approx(L, rel=1e-12)
    assert not np.isnan(fd).any()
    g = np.where(ci < 16, gp[gi, np.minimum(ci, 15)], gsh.reshape(120, -1)[gi, np.maximum(ci - 16, 0)])
    groups = {"position": [0, 1, 2], "opacity": [3], "scale": [4, 5, 6], "rotation": [8, 9, 10, 11],
              "dc": [12, 13, 14], "sh": list(range(16, 16 + 24))}
    for name, cs in groups.items():
        sel = np.isin(ci, cs)
        scale = np.abs(g[sel]).max()
        assert scale > 0
        err = np.abs(g[sel] - fd[sel])
        assert np.all(err <= 1e-5 * scale + 1e-4 * np.abs(fd[sel])), f"{name}: worst {err.max():.3g} of {scale:.3g}"
    # plain differences of the loss with its jumps: mostly the same
    rng = np.random.default_rng(3)
    agree = 0
    picks = rng.choice(len(gi), 40, replace=False)
    for k in picks:
        i, c = int(gi[k]), int(ci[k])
        pp, pm = p64.copy(), p64.copy()
        sp, sm = s64.copy(), s64.copy()
        if c < 16:
            pp[i, c] += 1e-4
            pm[i, c] -= 1e-4
        else:
            sp[i].reshape(-1)[c - 16] += 1e-4
            sm[i].reshape(-1)[c - 16] -= 1e-4
        plain = (orc.loss64(pp, sp, cam, 2, target, 0.2, bg=bg) - orc.loss64(pm, sm, cam, 2, target, 0.2, bg=bg)) / 2e-4
        agree += abs(plain - fd[k]) <= 1e-3 * abs(fd[k]) + 1e-9
    assert agree >= 30, f"plain and frozen differences agree on only {agree} of 40 samples"


def _quat_rotation_by_axis_angle(q):
    """The rotation a quaternion (w, x, y, z) stands for, from first
    principles: angle 2 atan2(|v|, w) about v / |v|, through Rodrigues'
    formula R = I + sin(t) K + (1 - cos(t)) K^2 (K the cross-product matrix)."""
    q = np.asarray(q, np.float64) / np.linalg.norm(q)
    w, v = q[0], q[1:]
    s = np.linalg.norm(v)
    if s == 0.0:
        return np.eye(3)
    k = v / s
    t = 2.0 * math.atan2(s, w)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K


def _splat_2d(p, cam, transpose=False, clamp=True):
    """Screen mean and dilated 2D covariance of one gaussian, from first
    principles: Sigma = R S^2 R^T in the world, the camera's rotation V and
    the perspective Jacobian J at the camera-space mean (x/z and y/z clamped
    to 1.3x the half field of view, the rasterizers' guard band), Sigma_2D =
    J V Sigma V^T J^T + 0.3 I.  transpose / clamp=False: the two mistakes the
    test must catch."""
    V = np.asarray(cam.rotation, np.float64).reshape(3, 3)
    t = V @ p[0:3] + np.asarray(cam.translation, np.float64)
    R = _quat_rotation_by_axis_angle(p[8:12])
    if transpose:
        R = R.T
    S = np.diag(np.exp(p[4:7]))
    sigma = R @ S @ S @ R.T
    fx, fy, cx, cy = cam.fx, cam.fy, cam.cx, cam.cy
    u, v = t[0] / t[2], t[1] / t[2]
    if clamp:
        lx = 1.3 * max(cx, cam.width - cx) / fx
        ly = 1.3 * max(cy, cam.height - cy) / fy
        u, v = min(max(u, -lx), lx), min(max(v, -ly), ly)
    J = np.array([[fx / t[2], 0.0, -fx * u / t[2]], [0.0, fy / t[2], -fy * v / t[2]]])
    cov = J @ V @ sigma @ V.T @ J.T + 0.3 * np.eye(2)
    mean = np.array([fx * t[0] / t[2] + cx, fy * t[1] / t[2] + cy])
    return mean, cov


@pytest.mark.parametrize("where", ["head-on", "off-axis-clamped"])
def test_project64_rotated_anisotropic_closed_form(where):
    """orc_project64's mean, dilated covariance and conic for rotated,
    strongly anisotropic splats (scales 0.3 / 0.05 / 0.01, quaternions far from
    the identity), seen head-on near the image centre and far outside the
    field of view where the Jacobian's clamp is active, against the first-
    principles construction above (1e-10 relative).  The same construction
    with the rotation transposed, or without the clamp, misses the oracle by
    far more: the test tells those mistakes apart."""
    cam = gs.look_at_camera((0.4, 0.3, 3.0), (0, 0, 0), 160, 120)
    rng = np.random.default_rng(12)
    checked = 0
    for _ in range(20):
        p = np.zeros(16, np.float64)
        p[3], p[7] = 2.0, 1.0
        p[4:7] = np.log([0.3, 0.05, 0.01])
        p[8:12] = rng.normal(size=4) + np.array([0.2, 0.0, 0.0, 0.0])
        if where == "head-on":
            p[0:3] = rng.uniform(-0.2, 0.2, 3)
        else:
            p[0:3] = np.array([7.0, -2.5, -0.5]) + rng.uniform(-0.5, 0.5, 3)
        V = np.asarray(cam.rotation, np.float64).reshape(3, 3)
        t = V @ p[0:3] + np.asarray(cam.translation, np.float64)
        lx = 1.3 * max(cam.cx, cam.width - cam.cx) / cam.fx
        if where != "head-on":
            assert abs(t[0] / t[2]) > lx  # the clamp is active
        out = orc.project64(p[None, :], None, cam, 0)[0]
        assert out[0] == 1.0
        mean, cov = _splat_2d(p, cam)
        con = np.linalg.inv(cov)
        np.testing.assert_allclose(out[1:3], mean, rtol=1e-10)
        np.testing.assert_allclose(out[11:14], [cov[0, 0], cov[0, 1], cov[1, 1]], rtol=1e-10,
                                   atol=1e-10 * np.abs(cov).max())
        np.testing.assert_allclose(out[5:8], [con[0, 0], con[0, 1], con[1, 1]], rtol=1e-10,
                                   atol=1e-10 * np.abs(con).max())
        got = np.array([out[11], out[12], out[13]])
        for wrong in (dict(transpose=True), dict(clamp=False)):
            if where == "head-on" and "clamp" in wrong:
                continue  # inside the field of view the clamp changes nothing
            _, bad = _splat_2d(p, cam, **wrong)
            miss = np.abs(np.array([bad[0, 0], bad[0, 1], bad[1, 1]]) - got).max() / np.abs(got).max()
            assert miss > 1e-3, f"{wrong}: only {miss:.2g} off"
        checked += 1
    assert checked == 20
