"""The second pass (vsMotionsToTransform, similarity model) on the CPU oracle.

tests/golden/transforms_golden.json holds the reference's own
vsMotionsToTransform (src/localmotion2transform.c:213-283, built from its
sources with its release flags) of every frame of the planar detect fixtures,
run on the fixtures' LocalMotions with the default transform configuration
(fov 0, no lens) and smoothZoom 1.  The oracle must reproduce every double
bit for bit.  Beside the fixtures, closed forms pin what they leave open: an
empty frame, a pure translation (the gradient descent stays on the exact
integer mean), and the zoom switch."""
import json
import os

import numpy as np
import pytest

import pyoracle
from golden_util import golden_motions, hexbits

HERE = os.path.dirname(os.path.abspath(__file__))
TGOLDEN = os.path.join(HERE, "golden", "transforms_golden.json")


@pytest.fixture(scope="module")
def tgolden():
    with open(TGOLDEN) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


def _bits(d):
    return int(np.array([d], dtype=np.float64).view(np.uint64)[0])


def test_fixtures_cover_every_planar_case(golden, tgolden):
    for name, case in golden.items():
        assert name in tgolden
        assert len(tgolden[name]["frames"]) == len(case["frames"])


@pytest.mark.parametrize("name", ["c1_full", "c1_stride", "c1_tripod", "c1_step5", "c1_step3", "c1_step8acc6",
                                  "c1_step1", "c1_acc4", "odd_502x298", "c2_1080p", "c3_4k", "c4_clip0",
                                  "c4_clip1", "c4_clip2", "c5_4kblur"])
def test_oracle_second_pass_matches_reference(golden, tgolden, name):
    case, tcase = golden[name], tgolden[name]
    rounds_seen = set()
    for t, (fr, want) in enumerate(zip(case["frames"], tcase["frames"])):
        lms = golden_motions(fr)
        tr, res, rounds = pyoracle.motions_to_transform(case["width"], case["height"], lms, smooth_zoom=1)
        got = [_bits(tr.x), _bits(tr.y), _bits(tr.alpha), _bits(tr.zoom), tr.extra]
        exp = [hexbits(want[0]), hexbits(want[1]), hexbits(want[2]), hexbits(want[3]), want[4]]
        assert got == exp, f"{name} frame {t}: {got} != {exp}"
        rounds_seen.add(rounds)
    if name == "c1_full":
        # the fixture set exercises the outlier rounds, not only the first descent
        assert max(rounds_seen) >= 2


def test_empty_frame_is_flagged():
    tr, res, rounds = pyoracle.motions_to_transform(640, 360, [])
    assert (tr.x, tr.y, tr.alpha, tr.zoom, tr.extra) == (0.0, 0.0, 0.0, 0.0, 1)


def test_pure_translation_stays_at_the_integer_mean():
    """every field moved by (3, -2) with equal matches: the mean motion is
    exact, the quality is 0 there (below the 0.01 threshold), so no descent
    step is taken and the result is the translation itself"""
    lms = [(3, -2, 40 + 37 * (i % 12), 30 + 41 * (i // 12), 32, 0, _bits(1.0)) for i in range(60)]
    tr, res, rounds = pyoracle.motions_to_transform(640, 480, lms, smooth_zoom=1)
    assert (tr.x, tr.y, tr.alpha, tr.zoom, tr.extra) == (3.0, -2.0, 0.0, 0.0, 0)
    assert res == 0.0 and rounds == 1


def test_rotation_is_recovered_and_zoom_switch():
    """fields displaced by a small rotation about the frame centre: the fitted
    alpha has the rotation's sign and size (the descent's regulariser and
    integer displacements allow a loose bound); smoothZoom 0 zeroes the zoom"""
    w, h, a = 1280, 720, 0.01
    lms = []
    for i in range(20):
        for j in range(11):
            fx, fy = 60 + 60 * i, 60 + 60 * j
            rx, ry = fx - w // 2, fy - h // 2
            vx = round(np.cos(a) * rx + np.sin(a) * ry - rx)
            vy = round(-np.sin(a) * rx + np.cos(a) * ry - ry)
            lms.append((vx, vy, fx, fy, 32, 0, _bits(1.0)))
    tr1, _, _ = pyoracle.motions_to_transform(w, h, lms, smooth_zoom=1)
    tr0, _, _ = pyoracle.motions_to_transform(w, h, lms, smooth_zoom=0)
    assert 0.5 * a < tr1.alpha < 1.5 * a
    assert tr0.zoom == 0.0 and (tr0.x, tr0.y, tr0.alpha) == (tr1.x, tr1.y, tr1.alpha)
