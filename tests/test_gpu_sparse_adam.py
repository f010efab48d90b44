"""Sparse Adam (GS_HIP_FLAG_SPARSE_ADAM; the reference's kFlagSparseAdam,
engine/shaders/train_step.glsl:223-230): a gaussian that drew nothing in the
step's views (off screen, or behind an opaque wall past where every pixel it
covers had stopped) keeps its parameters, SH coefficients and every Adam
moment bit for bit, while every gaussian that drew is stepped as without
the flag (to the float atomics' reordering between two runs).  The moments start random, so a dense step would move the
gaussians that drew nothing (their momentum), which the test also checks.
"""
import numpy as np
import pytest

from helpers import gs, make_camera, make_scene, smooth_target

pytestmark = pytest.mark.gpu

W, H, DEG = 96, 72, 3


@pytest.fixture(scope="module", autouse=True)
def device():
    print(f"\n[gpu] {gs.init(0)}", flush=True)
    yield
    gs.synchronize()


def scene():
    """300 gaussians near the origin, four layers of a wide opaque wall at
    z = -1 (filling the view), 40 small gaussians behind the wall and 40 far
    off to the side; SH degree 3."""
    front, sh_front = make_scene(300, DEG, 91, spread=0.4)
    rng = np.random.default_rng(92)
    coeffs = (DEG + 1) ** 2 - 1
    wall = np.zeros((4, 16), np.float32)
    wall[:, 0:3] = [[0.0, 0.0, -1.0 - 0.02 * k] for k in range(4)]
    wall[:, 3] = 10.0  # alpha capped at 0.99
    wall[:, 4:7] = [np.log(10.0), np.log(10.0), np.log(0.01)]
    wall[:, 7] = 1.0
    wall[:, 8] = 1.0
    wall[:, 12:15] = 0.3
    hidden = np.zeros((40, 16), np.float32)
    hidden[:, 0:2] = rng.uniform(-0.8, 0.8, (40, 2))
    hidden[:, 2] = rng.uniform(-1.6, -1.3, 40)
    hidden[:, 3] = 1.0
    hidden[:, 4:7] = np.log(0.05)
    hidden[:, 7] = 1.0
    hidden[:, 8:12] = rng.normal(size=(40, 4))
    hidden[:, 12:15] = rng.uniform(0, 1, (40, 3))
    away = hidden.copy()
    away[:, 0] += 50.0
    away[:, 2] = 0.0
    params = np.concatenate([front, wall, hidden, away]).astype(np.float32)
    sh = np.concatenate([sh_front, rng.normal(0, 0.2, (84, coeffs, 3))]).astype(np.float32)
    return params, sh


def random_moments(n, coeffs, cap, seed):
    rng = np.random.default_rng(seed)
    m = rng.normal(0, 1e-3, (cap, 16)).astype(np.float32)
    v = rng.uniform(1e-8, 1e-6, (cap, 16)).astype(np.float32)
    msh = rng.normal(0, 1e-3, (coeffs, cap, 3)).astype(np.float32)
    vsh = rng.uniform(1e-8, 1e-6, (coeffs, cap, 3)).astype(np.float32)
    return m, v, msh, vsh


def run(flags, views):
    params, sh = scene()
    n = params.shape[0]
    coeffs = (DEG + 1) ** 2 - 1
    cams = [make_camera(W, H, a) for a in (0.0, 0.15)][:views]
    targets = [gs.DeviceBuffer(smooth_target(W, H, 93 + k)) for k in range(views)]
    with gs.Workspace(n, DEG, W, H) as ws:
        ws.set_views(views)
        ws.load(params, sh)
        m, v, msh, vsh = random_moments(n, coeffs, ws.capacity, 94)
        for which, arr in ((gs.BUF_M_PARAMS, m), (gs.BUF_V_PARAMS, v), (gs.BUF_M_SH, msh), (gs.BUF_V_SH, vsh)):
            ws.upload(which, arr)
        frames = [gs.make_frame(c, n, DEG, step=5, flags=flags) for c in cams]
        ws.train_views(frames, [t.ptr for t in targets])
        drew = np.zeros(n, bool)
        tiles = np.zeros(n, np.uint32)
        for k in range(views):
            g2 = ws.download(gs.BUF_GRAD2D, np.float32, 9 * n, view=k).reshape(n, 9)
            drew |= np.any(g2 != 0, axis=1)
            tiles = np.maximum(tiles, ws.download(gs.BUF_TILES, np.uint32, n, view=k))
        out = {
            "params": ws.store(n)[0], "sh": ws.store(n)[1],
            "m": ws.download(gs.BUF_M_PARAMS, np.float32, n * 16).reshape(n, 16),
            "v": ws.download(gs.BUF_V_PARAMS, np.float32, n * 16).reshape(n, 16),
            "msh": ws.download(gs.BUF_M_SH, np.float32).reshape(coeffs, ws.capacity, 3)[:, :n],
            "vsh": ws.download(gs.BUF_V_SH, np.float32).reshape(coeffs, ws.capacity, 3)[:, :n],
        }
    for t in targets:
        t.free()
    before = {"params": params, "sh": sh, "m": m[:n], "v": v[:n], "msh": msh[:, :n], "vsh": vsh[:, :n]}
    return out, before, drew, tiles


def rows(name, a, sel):
    return a[:, sel] if name in ("msh", "vsh") else a[sel]


@pytest.mark.parametrize("views", [1, 2])
def test_sparse_adam_keeps_gaussians_that_drew_nothing(views):
    print(f"[case] sparse adam, {views} view(s)", flush=True)
    sparse, before, drew, tiles = run(gs.FLAG_SPARSE_ADAM, views)
    dense, _, drew_d, _ = run(0, views)
    np.testing.assert_array_equal(drew, drew_d)
    n = drew.size
    hidden, away = np.arange(304, 344), np.arange(344, 384)
    assert not drew[away].any()
    assert (~drew[hidden]).sum() >= 30, f"only {(~drew[hidden]).sum()} of 40 gaussians behind the wall drew nothing"
    assert (tiles[hidden] > 0).sum() >= 30  # binned into tiles (seen), yet past every pixel's stop
    assert drew[:300].sum() >= 100 and drew[300]  # the wall's front layer (the ones behind may be past the stop)
    idle = ~drew
    print(f"[info] {idle.sum()} of {n} gaussians drew nothing ({(idle & (tiles > 0)).sum()} of them in some tile)",
          flush=True)
    for name in ("params", "sh", "m", "v", "msh", "vsh"):
        np.testing.assert_array_equal(rows(name, sparse[name], idle), rows(name, before[name], idle),
                                      err_msg=f"{name} of gaussians that drew nothing moved")
        # (two runs: the backward's float atomics may add in another order, an ulp or so)
        np.testing.assert_allclose(rows(name, sparse[name], drew), rows(name, dense[name], drew), rtol=1e-5,
                                   atol=1e-9, err_msg=f"{name} of gaussians that drew differ from a dense step")
    # Without the flag their momentum moves them.
    assert np.any(rows("params", dense["params"], idle) != rows("params", before["params"], idle))
    assert np.any(rows("sh", dense["sh"], idle) != rows("sh", before["sh"], idle))
