"""Shared test helpers: module loading, seeded scenes, cameras, targets."""
import importlib.util
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load(name, rel):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, rel))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


gs = _load("gs_hip", "on-divice-gsplate_amd/gs_hip.py")
orc = _load("gs_oracle", "oracle/oracle.py")


def make_scene(n, degree, seed, spread=0.7):
    """Scenes shaped like the reference's make_test_scene
    (engine/src/train_selftest.cpp:314-332): positions spread * U(-1, 1),
    scales log(0.08 + 0.06 U), quaternion U + 1.5 on w, logits 0.6 U - 0.2,
    DC 0.8 U, SH rest 0.25 U — drawn from numpy's PCG64, not mt19937."""
    rng = np.random.default_rng(seed)
    u = lambda *s: rng.uniform(-1.0, 1.0, size=s)
    params = np.zeros((n, 16), np.float32)
    params[:, 0:3] = spread * u(n, 3)
    params[:, 4:7] = np.log(0.08 + 0.06 * u(n, 3))
    params[:, 12:15] = 0.8 * u(n, 3)
    q = u(n, 4)
    q[:, 0] += 1.5
    params[:, 8:12] = q
    params[:, 3] = 0.6 * u(n) - 0.2
    params[:, 7] = 1.0
    coeffs = 0 if degree == 0 else (degree + 1) ** 2 - 1
    sh = (0.25 * u(n, coeffs, 3)).astype(np.float32)
    return params, sh


def make_camera(width, height, angle=0.0, focal_scale=0.9):
    """Orbit camera at radius 3 looking at the origin, fx = fy = 0.9 W
    (make_test_dataset, train_selftest.cpp:285-312)."""
    eye = (3.0 * math.sin(angle), 0.6 + 0.3 * math.sin(3.0 * angle), 3.0 * math.cos(angle))
    cam = gs.look_at_camera(eye, (0.0, 0.0, 0.0), width, height)
    cam.fx = cam.fy = focal_scale * width
    return cam


def noise_target(width, height, seed):
    rng = np.random.default_rng(seed)
    t = rng.integers(0, 256, size=(height, width, 4), dtype=np.uint8)
    t[..., 3] = 255
    return t


def smooth_target(width, height, seed):
    """A smooth RGBA8 target (low-frequency colour field), alpha 255."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    out = np.zeros((height, width, 4), np.uint8)
    for c in range(3):
        a, b, p = rng.uniform(0.02, 0.1), rng.uniform(0.02, 0.1), rng.uniform(0, 6.28)
        out[..., c] = np.clip(127.5 + 100 * np.sin(a * x + b * y + p), 0, 255).astype(np.uint8)
    out[..., 3] = 255
    return out
