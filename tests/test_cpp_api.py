"""Runs the C++ API tests (tests/cpp) built by the package Makefile: scene
I/O, image files, JSON and the dataset loader without a GPU; the gs::Trainer
and gs::SplatRenderer on the GPU."""
import os
import subprocess

import pytest

from helpers import ROOT

BUILD = os.path.join(ROOT, "on-divice-gsplate_amd", "_build")


def _run(name, timeout):
    exe = os.path.join(BUILD, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built: run make -C on-divice-gsplate_amd tests")
    print(f"[case] {name}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_scene_io():
    _run("test_scene_io", 120)


def test_image_io():
    _run("test_image_io", 120)


def test_image_io_under_asan():
    """The image/JSON tests again, built with AddressSanitizer and UBSan."""
    _run("asan_image_io", 300)


def test_quadrant_cull():
    """The rasterizers' quadrant cull keeps every 8x8 box that holds a pixel a
    needle splat draws (csrc/cull.h, run on the host)."""
    _run("test_cull", 120)


def test_bucket_slots():
    """The stored-state backward's bucket slots: one per (bin, raster tile,
    bucket) within the view's reserve (csrc/gs_internal.h, on the host)."""
    _run("test_buckets", 120)


def test_dataset_loader():
    _run("test_dataset", 120)


@pytest.mark.gpu
def test_trainer_api():
    _run("test_trainer", 300)


@pytest.mark.gpu
def test_renderer_api():
    _run("test_renderer", 120)
