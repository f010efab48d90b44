"""Writes tests/golden/c1_trajectory.json: the oracle's CPU restatement of
the training step run for the whole of the small-scene configuration, the
fixture the GPU trajectory test compares with.

Configuration: a seeded 20k-gaussian ground-truth scene (SH degree 0, the
benchmark generator: means in the unit ball, log-scales N(-4, 0.5),
normalised N(0,1) quaternions, logits N(0,1), DC U[0,1]; seed 1001) renders
24 targets (fp64 render rounded to RGBA8, alpha 255) at cameras evenly spaced
on a circle of radius 3 at height 0.5 looking at the origin, 60 degree
vertical field of view, 360 wide x 640 high; training starts from an
independently seeded 20k scene (seed 2002); loss 0.8 L1 + 0.2 (1 - SSIM);
Adam at 1.6e-4 for means and 5e-3 for every other parameter; 300 steps,
view k mod 24 at step k; no densification.

Recorded: a checksum of the targets (so a test regenerating them knows it
has the same ones), the mean PSNR over the 24 views (fp64 render) before and
after, and every 25th step's PSNR of the model then.  Run from the repository
root: python tests/golden/make_c1_trajectory.py
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from helpers import gs, orc  # noqa: E402

W, H, N, VIEWS, STEPS = 360, 640, 20_000, 24, 300
LR_MEANS, LR_OTHER, LAMBDA = 1.6e-4, 5e-3, 0.2


def setup():
    cams = gs.orbit_cameras(VIEWS, W, H)
    gt_p, gt_sh = gs.synthetic_scene(N, 0, 1001)
    targets = orc.render_targets64(gt_p, gt_sh, 0, cams)
    p0, s0 = gs.synthetic_scene(N, 0, 2002)
    lr = orc.adam_rates(N, LR_MEANS, LR_OTHER, LR_OTHER, LR_OTHER, LR_OTHER)
    return cams, targets, p0, s0, lr


def targets_digest(targets) -> str:
    h = hashlib.sha256()
    for t in targets:
        h.update(t.tobytes())
    return h.hexdigest()


def main():
    t0 = time.time()
    cams, targets, p, s, lr = setup()
    trace, state = {}, {}
    start = orc.mean_psnr(p, s, 0, cams, targets)
    for k in range(0, STEPS, 25):
        p, s = orc.train(p, s, 0, cams, targets, 25, lr, LR_OTHER, LAMBDA, order=[(k + j) % VIEWS for j in range(25)],
                         step0=k, state=state)
        trace[k + 25] = orc.mean_psnr(p, s, 0, cams, targets)
        print(f"step {k + 25}: {trace[k + 25]:.4f} dB ({time.time() - t0:.0f} s)", flush=True)
    out = {"config": __doc__.split("Recorded:")[0].strip(), "targets_sha256": targets_digest(targets),
           "psnr_start": start, "psnr_final": trace[STEPS], "psnr_every_25_steps": trace, "steps": STEPS}
    with open(os.path.join(HERE, "c1_trajectory.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"psnr {start:.4f} -> {trace[STEPS]:.4f} dB in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
