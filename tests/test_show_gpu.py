"""GPU detect with show = 1 / 2 through the vid.stab C API against the
reference's show fixtures (tests/golden/show_golden.json): the caller's frame
after each call (planar luma, packed RGBA / BGR24) and every LocalMotion, bit
for bit, per frame, through the batch entry point, with the frames in device
memory (vsMotionDetectionDevice) and through the multi-clip form."""
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


def _to_device(frames):
    """the frames as one device block, each frame's rows at its own stride"""
    import numpy as np
    import torch
    host = np.stack(frames)
    dev = torch.from_numpy(host).to("cuda:0")
    torch.cuda.synchronize()
    step = host.shape[1] * host.shape[2]
    return dev, [dev.data_ptr() + i * step for i in range(len(frames))], host.shape[2]


@pytest.mark.parametrize("name", CASES)
def test_show_device_frames_match_reference_fixture(show_golden, name):
    """the overlays drawn into device-resident frames, two device batches"""
    import torch
    case = show_golden[name]
    wb = case["width"] * case["bpp"]
    _say(f"[gpu] show device frames {name} show={case['show']} bpp {case['bpp']}")
    frames = _frames(case)
    dev, ptrs, stride = _to_device(frames)
    det = _detector(case)
    k = max(1, len(frames) // 2)
    got = det.detect_device(ptrs[:k], [stride] * k, keep=True) + \
        det.detect_device(ptrs[k:], [stride] * (len(frames) - k), keep=True)
    det.close()
    torch.cuda.synchronize()
    drawn = dev.cpu().numpy()
    for t, fr in enumerate(case["frames"]):
        assert "%016x" % hash_plane(drawn[t], wb) == fr["drawn"], f"{name} overlay frame {t}"
        assert got[t] == golden_motions(fr), f"{name} frame {t}"


def test_show_multi_clip_matches_reference_fixture(show_golden):
    """the planar show cases as clips of one vsMotionDetectMulti call"""
    import torch
    names = ["s_c1_show1", "s_c1_show2", "s_c1_step3_show2", "s_odd_show2", "s_1080p_show2"]
    _say(f"[gpu] show multi-clip {names}")
    clips, devs = [], []
    for name in names:
        case = show_golden[name]
        dev, ptrs, stride = _to_device(_frames(case))
        devs.append(dev)
        clips.append(dict(width=case["width"], height=case["height"], ptrs=ptrs, linesize=stride,
                          show=case["show"], **case_params(case)))
    got = api.detect_multi(clips)
    torch.cuda.synchronize()
    for name, dev, motions in zip(names, devs, got):
        case = show_golden[name]
        drawn = dev.cpu().numpy()
        for t, fr in enumerate(case["frames"]):
            assert "%016x" % hash_plane(drawn[t], case["width"]) == fr["drawn"], f"{name} overlay frame {t}"
            assert motions[t] == golden_motions(fr), f"{name} frame {t}"
