"""The CPU oracle's show = 1 / 2 overlays against the reference's fixtures
(tests/golden/show_golden.json, written by tests/golden/make_golden.c from the
reference's vsMotionDetection with conf.show set): the frame the caller
handed in, after the call, bit for bit (planar luma, packed RGBA / BGR24), and
every LocalMotion -- the reference's line drawing writes the end point into
the field position of leftward or upward axis-parallel motions, so show
changes the returned motions."""
import json
import os

import numpy as np
import pytest

import pyoracle
from golden_util import case_params, golden_motions

HERE = os.path.dirname(os.path.abspath(__file__))
FAST = ["s_c1_show1", "s_c1_show2", "s_c1_step3_show2", "s_odd_show2", "s_rgba_show2", "s_bgr24_show1"]
SLOW = ["s_1080p_show2"]


@pytest.fixture(scope="module")
def show_golden():
    with open(os.path.join(HERE, "golden", "show_golden.json")) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


def show_frames(case):
    """the fixture's frames (planar luma or packed plane 0), regenerated"""
    from vidstab_amd.synth import Clip, pack
    clip = Clip(case["config"], case["clip"], width=case["width"], height=case["height"],
                nframes=len(case["frames"]))
    for t in range(len(case["frames"])):
        if case["bpp"] == 1:
            yield clip.render(t, stride=case["stride"])
        else:
            yield pack(clip.render(t), case["bpp"], stride=case["stride"], width=case["width"])


def _run(case):
    from vidstab_amd.synth import hash_plane
    wb = case["width"] * case["bpp"]
    det = pyoracle.Detector(case["width"], case["height"], bpp=case["bpp"], show=case["show"],
                            **case_params(case))
    moved = 0
    for t, (fr, frame) in enumerate(zip(case["frames"], show_frames(case))):
        assert "%016x" % hash_plane(frame, wb) == fr["in"], f"generator drift frame {t}"
        got = det.detect(frame)
        assert "%016x" % hash_plane(frame, wb) == fr["drawn"], f"{case['name']} overlay frame {t}"
        assert got == golden_motions(fr), f"{case['name']} frame {t}"
        moved += sum(1 for m in got if (m[0] == 0) != (m[1] == 0) and min(m[0], m[1]) < 0)
    return moved


@pytest.mark.parametrize("name", FAST)
def test_show_oracle_matches_reference(show_golden, name):
    _run(show_golden[name])


@pytest.mark.slow
@pytest.mark.parametrize("name", SLOW)
def test_show_oracle_matches_reference_large(show_golden, name):
    _run(show_golden[name])


def test_show_fixtures_hold_moved_field_positions(show_golden):
    """the fixtures contain motions whose field position the line drawing
    rewrote, so that quirk is pinned by the reference and not assumed"""
    moved = sum(_run(show_golden[n]) for n in FAST)
    assert moved > 0
