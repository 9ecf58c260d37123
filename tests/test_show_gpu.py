"""GPU detect with show = 1 / 2 through the vid.stab C API against the
reference's show fixtures (tests/golden/show_golden.json): the caller's frame
after each call (planar luma, packed RGBA / BGR24) and every LocalMotion, bit
for bit, per frame and through the batch entry point."""
import json
import os
import sys

import pytest

from golden_util import case_params, golden_motions
from vidstab_amd import api
from vidstab_amd.synth import hash_plane

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["s_c1_show1", "s_c1_show2", "s_c1_step3_show2", "s_odd_show2", "s_rgba_show2", "s_bgr24_show1",
         "s_1080p_show2"]


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def show_golden():
    with open(os.path.join(HERE, "golden", "show_golden.json")) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


def _frames(case):
    from vidstab_amd.synth import Clip, pack
    clip = Clip(case["config"], case["clip"], width=case["width"], height=case["height"],
                nframes=len(case["frames"]))
    out = []
    for t in range(len(case["frames"])):
        if case["bpp"] == 1:
            out.append(clip.render(t, stride=case["stride"]))
        else:
            out.append(pack(clip.render(t), case["bpp"], stride=case["stride"], width=case["width"]))
    return out


def _detector(case):
    pformat = case["pformat"] or api.PF_YUV420P
    return api.Detector(case["width"], case["height"], pformat=pformat, show=case["show"], **case_params(case))


@pytest.mark.parametrize("name", CASES)
def test_show_per_frame_matches_reference_fixture(show_golden, name):
    case = show_golden[name]
    wb = case["width"] * case["bpp"]
    _say(f"[gpu] show fixture {name} show={case['show']} bpp {case['bpp']}")
    det = _detector(case)
    for t, (fr, frame) in enumerate(zip(case["frames"], _frames(case))):
        got = det.detect(frame)
        assert "%016x" % hash_plane(frame, wb) == fr["drawn"], f"{name} overlay frame {t}"
        assert got == golden_motions(fr), f"{name} frame {t}"
    det.close()


@pytest.mark.parametrize("name", ["s_c1_show2", "s_rgba_show2"])
def test_show_batch_matches_reference_fixture(show_golden, name):
    case = show_golden[name]
    wb = case["width"] * case["bpp"]
    _say(f"[gpu] show batch {name}")
    det = _detector(case)
    frames = _frames(case)
    got = det.detect_batch(frames[:2]) + det.detect_batch(frames[2:])
    for t, fr in enumerate(case["frames"]):
        assert "%016x" % hash_plane(frames[t], wb) == fr["drawn"], f"{name} overlay frame {t}"
        assert got[t] == golden_motions(fr), f"{name} frame {t}"
    det.close()
