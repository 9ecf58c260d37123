"""GPU parity of the packed RGB24 / BGR24 / RGBA detect pass (the unblurred
plane 0, all bytes of a pixel in the SAD, the mean channel contrast, the
raster grid with the unshifted position first, one 1-px refinement level)
through the vid.stab C API, against the reference's packed fixtures
(tests/golden/packed_golden.json) and the CPU oracle on fresh seeds.

Bar: bit-identical LocalMotions (integer shifts and the IEEE bits of
contrast and match) and a bit-identical working plane.
"""
import json
import os
import sys

import numpy as np
import pytest

import pyoracle
from golden_util import case_params, golden_motions
from vidstab_amd import api
from vidstab_amd.synth import Clip, hash_plane, pack

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["p_rgb24_c1", "p_bgr24_pad", "p_rgba_c1", "p_rgb24_tripod", "p_rgb24_step4", "p_rgba_step8acc6",
         "p_rgb24_step1", "p_rgba_odd", "p_rgb24_1080p"]


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def packed_golden():
    with open(os.path.join(HERE, "golden", "packed_golden.json")) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


def _frames(case):
    clip = Clip(case["config"], case["clip"], width=case["width"], height=case["height"],
                nframes=len(case["frames"]))
    return [pack(clip.render(t), case["bpp"], stride=case["stride"], width=case["width"])
            for t in range(len(case["frames"]))]


def _detector(case):
    return api.Detector(case["width"], case["height"], pformat=case["pformat"], **case_params(case))


@pytest.mark.parametrize("name", CASES)
def test_packed_per_frame_matches_reference_fixture(packed_golden, name):
    case = packed_golden[name]
    _say(f"[gpu] packed fixture {name} {case['width']}x{case['height']} bpp {case['bpp']} stride {case['stride']}")
    det = _detector(case)
    for t, (fr, frame) in enumerate(zip(case["frames"], _frames(case))):
        got = det.detect(frame)
        assert "%016x" % hash_plane(det.blurred()) == fr["blur"], f"{name} working plane frame {t}"
        assert got == golden_motions(fr), f"{name} frame {t}"
    det.close()


@pytest.mark.parametrize("name", ["p_rgb24_c1", "p_rgb24_tripod", "p_rgba_odd"])
def test_packed_batch_matches_reference_fixture(packed_golden, name):
    case = packed_golden[name]
    _say(f"[gpu] packed batch {name}")
    det = _detector(case)
    frames = _frames(case)
    got = det.detect_batch(frames[:3]) + det.detect_batch(frames[3:])
    for t, fr in enumerate(case["frames"]):
        assert got[t] == golden_motions(fr), f"{name} frame {t}"
    det.close()


@pytest.mark.parametrize("lds", [0, 1])
@pytest.mark.parametrize("K", [4, 6, 7, 8])
@pytest.mark.parametrize("name", ["p_rgba_step8acc6", "p_rgb24_1080p"])
def test_every_packed_search_plan_is_bit_identical(packed_golden, monkeypatch, name, K, lds):
    monkeypatch.setenv("VIDSTAB_AMD_SEARCH_K", str(K))
    monkeypatch.setenv("VIDSTAB_AMD_SEARCH_LDS", str(lds))
    case = packed_golden[name]
    _say(f"[gpu] packed plan K={K} lds={lds} fixture {name}")
    det = _detector(case)
    got = det.detect_batch(_frames(case))
    for t, fr in enumerate(case["frames"]):
        assert got[t] == golden_motions(fr), f"{name} K={K} lds={lds} frame {t}"
    det.close()


def test_packed_device_resident_entry_point(packed_golden):
    import torch
    case = packed_golden["p_rgb24_1080p"]
    _say("[gpu] packed device-resident 1080p")
    host = np.stack(_frames(case))
    dev = torch.from_numpy(host).to("cuda:0")
    torch.cuda.synchronize()
    det = _detector(case)
    fb = host.shape[1] * host.shape[2]
    got = det.detect_device([dev.data_ptr() + i * fb for i in range(host.shape[0])],
                            [case["stride"]] * host.shape[0], keep=True)
    for t, fr in enumerate(case["frames"]):
        assert got[t] == golden_motions(fr), f"frame {t}"
    det.close()


@pytest.mark.parametrize("pformat,clip_id,step", [(api.PF_RGB24, 51, 6), (api.PF_RGBA, 52, 3),
                                                   (api.PF_BGR24, 53, 7)])
def test_packed_matches_oracle_on_fresh_seeds(pformat, clip_id, step):
    clip = Clip(1, clip_id, nframes=5)
    s = clip.spec
    bpp = 4 if pformat == api.PF_RGBA else 3
    _say(f"[gpu] packed oracle seed clip {clip_id} bpp {bpp} step {step}")
    kw = dict(shakiness=s.shakiness, accuracy=8, stepsize=step, mincontrast=s.mincontrast)
    gpu = api.Detector(s.width, s.height, pformat=pformat, **kw)
    ref = pyoracle.Detector(s.width, s.height, bpp=bpp, **kw)
    for t in range(5):
        frame = pack(clip.render(t), bpp)
        assert gpu.detect(frame) == ref.detect(frame), f"frame {t}"
        assert np.array_equal(gpu.blurred(), ref.blurred()), f"working plane frame {t}"
    gpu.close()


def test_packed_shift_is_recovered():
    """blocky random colour tiles, the next frame shifted so that
    cur(x, y) = prev(x + 8, y - 6): the motions find d = (8, -6) and equal
    the oracle's"""
    rng = np.random.default_rng(42)
    w, h, bpp = 320, 240, 3
    blocks = rng.integers(0, 256, size=(h // 8 + 1, w // 8 + 1, bpp)).astype(np.uint8)
    f1 = np.repeat(np.repeat(blocks, 8, axis=0), 8, axis=1)[:h, :w]
    f2 = np.zeros_like(f1)
    f2[6:, :w - 8] = f1[:h - 6, 8:]
    det = api.Detector(w, h, pformat=api.PF_RGB24)
    det.detect(f1.reshape(h, w * bpp))
    got = det.detect(f2.reshape(h, w * bpp))
    ref = pyoracle.Detector(w, h, bpp=bpp)
    ref.detect(f1.reshape(h, w * bpp))
    assert got == ref.detect(f2.reshape(h, w * bpp))
    assert len(got) > 4
    vx = np.median([m[0] for m in got])
    vy = np.median([m[1] for m in got])
    assert (vx, vy) == (8, -6)
    det.close()
