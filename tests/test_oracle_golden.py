"""The CPU oracle against the reference's golden fixtures.

The fixtures (tests/golden/detect_golden.json) were written by
tests/golden/make_golden.c from the reference's own vsMotionDetection built
from its sources with its release flags.  Every case regenerates the seeded
frames, so the input hash also pins the generator.
"""
import pytest

import pyoracle
from golden_util import case_clip, case_params, golden_motions
from vidstab_amd.synth import hash_plane

FAST = ["c1_stride", "c1_tripod", "c1_step5", "c1_step3", "c1_step8acc6", "c1_step1",
        "c1_acc4", "odd_502x298"]
SLOW = ["c1_full", "c2_1080p", "c4_clip2", "c4_clip0", "c4_clip1", "c3_4k", "c5_4kblur"]


def _run(case):
    clip, stride = case_clip(case)
    det = pyoracle.Detector(case["width"], case["height"], **case_params(case))
    st = det.state
    assert [st.coarse.fieldSize, st.coarse.maxShift, st.coarse.fieldNum, st.coarse.maxFields] \
        == case["coarse"]
    assert [st.fine.fieldSize, st.fine.maxShift, st.fine.fieldNum, st.fine.maxFields] \
        == case["fine"]
    for t, fr in enumerate(case["frames"]):
        luma = clip.render(t, stride=stride)
        assert "%016x" % hash_plane(luma, case["width"]) == fr["in"], f"generator drift frame {t}"
        got = det.detect(luma)
        assert "%016x" % hash_plane(det.blurred()) == fr["blur"], f"blur frame {t}"
        want = golden_motions(fr)
        assert got == want, f"{case['name']} frame {t}: {len(got)} vs {len(want)} motions"


@pytest.mark.parametrize("name", FAST)
def test_oracle_matches_reference(golden, name):
    _run(golden[name])


@pytest.mark.slow
@pytest.mark.parametrize("name", SLOW)
def test_oracle_matches_reference_large(golden, name):
    _run(golden[name])


def test_oracle_matches_whole_clip_hashes():
    """the reference's whole-clip hashes (tests/golden/clip_hashes.json, written
    by make_golden's hashes mode): every 640x360 clip of the multi-clip
    workload in full, and the first frames of every other clip (1080p, 4K step
    4, 720p, 1440p) -- the same hash of LocalMotions the GPU tests use"""
    from golden_util import hash_motions, load_clip_hashes
    from vidstab_amd.synth import Clip
    hashes = load_clip_hashes()
    assert {"w2_full", "w3_full"} <= set(hashes) and sum(k.startswith("w4_clip") for k in hashes) == 64
    for name, case in hashes.items():
        nfr = len(case["frames"]) if case["width"] == 640 else 3
        clip = Clip(case["config"], case["clip"])
        det = pyoracle.Detector(case["width"], case["height"], **case_params(case))
        for t in range(nfr):
            hin, cnt, hl, hb = case["frames"][t]
            luma = clip.render(t)
            assert "%016x" % hash_plane(luma) == hin, f"{name}: generator drift frame {t}"
            got = det.detect(luma)
            assert len(got) == cnt and "%016x" % hash_motions(got) == hl, f"{name} frame {t}"
            assert "%016x" % hash_plane(det.blurred()) == hb, f"{name} blur frame {t}"
