"""The second pass's per-frame global transforms on the GPU
(vsMotionsToTransformBatch: vsMotionsToTransform, similarity model, every
frame of a clip in one launch) against the reference's own results
(tests/golden/transforms_golden.json, which the CPU oracle reproduces bit for
bit) and against the oracle on larger synthetic clips.

Tolerance.  Every double expression runs in the reference's order, but the
device's sin and cos may round differently from glibc's in the last bit.
The descent's gradient is a finite difference over h = 1e-6, so such a bit
moves a trial step by about 1e-10 of its size, and a comparison between two
nearly equal qualities late in a descent can then go the other way, moving a
parameter by at most the step size in force there.  The bar: x and y within
1e-6 px, alpha within 1e-9 rad, zoom within 1e-6 percent, extra equal, in
every frame.  On the reference's fixtures and on the three seeded synthetic
clips below (360 frames, outlier rounds included) every frame is
bit-identical -- the device's sincos rounds as glibc's on all of their
arguments -- and that is asserted; the tolerance is the bar for other
inputs."""
import json
import os
import sys

import numpy as np
import pytest

import pyoracle
from golden_util import golden_motions, hexbits, bits_to_double
from vidstab_amd import api

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = (1e-6, 1e-6, 1e-9, 1e-6)


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _close(a, b):
    return all(abs(a[i] - b[i]) <= TOL[i] for i in range(4)) and a[4] == b[4]


def _bits_equal(a, b):
    return all(np.float64(a[i]).view(np.uint64) == np.float64(b[i]).view(np.uint64) for i in range(4)) and \
        a[4] == b[4]


def test_gpu_second_pass_matches_reference_fixtures(golden):
    with open(os.path.join(HERE, "golden", "transforms_golden.json")) as f:
        tg = {c["name"]: c for c in json.load(f)["cases"]}
    total = exact = 0
    worst = np.zeros(4)
    for name, case in golden.items():
        frames = [golden_motions(fr) for fr in case["frames"]]
        _say(f"[gpu] second pass: case {name}, {len(frames)} frames")
        got = api.motions_to_transforms(case["width"], case["height"], frames, smooth_zoom=1)
        for t, (g, w) in enumerate(zip(got, tg[name]["frames"])):
            want = tuple(bits_to_double(hexbits(w[i])) for i in range(4)) + (w[4],)
            assert _close(g, want), f"{name} frame {t}: {g} vs {want}"
            worst = np.maximum(worst, [abs(g[i] - want[i]) for i in range(4)])
            exact += _bits_equal(g, want)
            total += 1
    _say(f"[gpu] second pass vs reference: {exact}/{total} frames bit-identical, largest differences "
         f"x {worst[0]:.2e} y {worst[1]:.2e} alpha {worst[2]:.2e} zoom {worst[3]:.2e}")
    assert total > 100
    # on these inputs the device's sincos rounds as glibc's does: every frame is exact
    assert exact == total


def _synthetic_clip(seed, nframes, width, height):
    """LocalMotions of a drifting, rotating camera with a moving object (a
    cluster of outlier fields), uneven field counts, some frames empty"""
    rng = np.random.default_rng(seed)
    frames = []
    for f in range(nframes):
        n = 0 if f % 37 == 5 else int(rng.integers(6, 1500))
        a = rng.normal(0.0, 0.004)
        tx, ty = rng.normal(0.0, 6.0, 2)
        fx = rng.integers(16, width - 16, n)
        fy = rng.integers(16, height - 16, n)
        rx, ry = fx - width // 2, fy - height // 2
        vx = np.rint(np.cos(a) * rx + np.sin(a) * ry - rx + tx + rng.normal(0, 0.7, n)).astype(int)
        vy = np.rint(-np.sin(a) * rx + np.cos(a) * ry - ry + ty + rng.normal(0, 0.7, n)).astype(int)
        obj = rng.random(n) < 0.15
        vx[obj] += int(rng.integers(-30, 30))
        vy[obj] += int(rng.integers(-30, 30))
        match = np.abs(rng.normal(3.0, 1.0, n))
        match[rng.random(n) < 0.05] += 20.0
        frames.append([(int(vx[i]), int(vy[i]), int(fx[i]), int(fy[i]), 32, 0,
                        int(np.float64(match[i]).view(np.uint64))) for i in range(n)])
    return frames


@pytest.mark.parametrize("seed,width,height,smooth", [(1, 1920, 1080, 0), (2, 3840, 2160, 1), (3, 640, 360, 1)])
def test_gpu_second_pass_matches_oracle_on_synthetic_clips(seed, width, height, smooth):
    frames = _synthetic_clip(seed, 120, width, height)
    _say(f"[gpu] second pass synthetic seed {seed}: {len(frames)} frames {width}x{height}")
    got = api.motions_to_transforms(width, height, frames, smooth_zoom=smooth)
    exact = 0
    rounds = set()
    for t, (g, motions) in enumerate(zip(got, frames)):
        tr, res, r = pyoracle.motions_to_transform(width, height, motions, smooth_zoom=smooth)
        want = (tr.x, tr.y, tr.alpha, tr.zoom, tr.extra)
        assert _close(g, want), f"seed {seed} frame {t} ({len(motions)} motions): {g} vs {want}"
        exact += _bits_equal(g, want)
        rounds.add(r)
    _say(f"[gpu] synthetic seed {seed}: {exact}/{len(frames)} bit-identical, descent rounds seen {sorted(rounds)}")
    assert 0 in rounds and max(rounds) >= 2      # empty frames and the outlier rounds both occur
    assert exact == len(frames)


def test_gpu_second_pass_edge_cases():
    """no frames; frames without motions; a single motion"""
    assert api.motions_to_transforms(640, 360, []) == []
    one = [(2, -1, 100, 80, 16, 0, int(np.float64(1.5).view(np.uint64)))]
    got = api.motions_to_transforms(640, 360, [[], one, []], smooth_zoom=1)
    tr, _, _ = pyoracle.motions_to_transform(640, 360, one, smooth_zoom=1)
    assert got[0] == (0.0, 0.0, 0.0, 0.0, 1) and got[2] == (0.0, 0.0, 0.0, 0.0, 1)
    assert _close(got[1], (tr.x, tr.y, tr.alpha, tr.zoom, tr.extra))
