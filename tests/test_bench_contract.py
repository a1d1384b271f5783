"""bench.py's command-line contract on a host without a GPU: more than one
GPU is refused with a message (exit 2), no visible GPU is an error (exit 1),
and the feed taper it builds its workload on is the stated -12 dB."""

import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True,
                          text=True, timeout=300)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_no_gpu_is_an_error_not_a_number():
    import torch
    if torch.cuda.is_available():
        import pytest
        pytest.skip("a GPU is visible here")
    r = _run("--steps", "1", "--warmup", "0")
    assert r.returncode == 1
    assert "no GPU" in r.stderr and r.stdout == ""


def test_feed_waist_gives_minus_12_db_at_the_rim():
    """Far field of a Gaussian aperture of waist w0 at the rim angle
    theta_e = 2 atan(D / 4f), with the obliquity factor (1 + cos)/2 and the
    1 / distance space loss of a paraboloid (r = 2f / (1 + cos)), relative to
    boresight: 10^(-12/20)."""
    sys.path.insert(0, ROOT)
    import bench
    w0 = bench.feed_waist()
    te = 2 * np.arctan(bench.D_DISH / (4 * bench.F_DISH))
    pattern = np.exp(-(bench.K * w0 * np.sin(te) / 2) ** 2) * (1 + np.cos(te)) / 2
    space = (1 + np.cos(te)) / 2        # f / r at the rim over f / f at the vertex
    assert abs(20 * np.log10(pattern * space) - bench.EDGE_DB) < 1e-9
