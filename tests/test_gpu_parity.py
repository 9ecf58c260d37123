"""GPU parity: the MI355X library through the vid.stab C API against the
reference's golden fixtures and against the CPU oracle on fresh seeds.

Bar: bit-identical LocalMotions (positions, integer shifts, and the IEEE bits
of contrast and match) and bit-identical blurred planes.
"""
import sys

import numpy as np
import pytest

import pyoracle
from golden_util import case_clip, case_params, golden_motions
from vidstab_amd import api
from vidstab_amd.synth import Clip, hash_plane, pack

pytestmark = pytest.mark.gpu

CASES = ["c1_full", "c1_stride", "c1_tripod", "c1_step5", "c1_step3", "c1_step8acc6",
         "c1_step1", "c1_acc4", "odd_502x298", "c2_1080p", "c4_clip2", "c4_clip0",
         "c4_clip1", "c3_4k", "c5_4kblur"]


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.mark.parametrize("name", CASES)
def test_c_api_matches_reference_fixture(golden, name):
    case = golden[name]
    _say(f"[gpu] fixture {name} {case['width']}x{case['height']} stride {case['stride']}")
    clip, stride = case_clip(case)
    det = api.Detector(case["width"], case["height"], **case_params(case))
    for t, fr in enumerate(case["frames"]):
        luma = clip.render(t, stride=stride)
        got = det.detect(luma)
        assert "%016x" % hash_plane(det.blurred()) == fr["blur"], f"{name} blur frame {t}"
        want = golden_motions(fr)
        assert got == want, f"{name} frame {t}: {len(got)} vs {len(want)} motions"
    det.close()


@pytest.mark.parametrize("chunk", [1, 7, 16, 40])
def test_batch_entry_point_equals_per_frame(golden, chunk):
    case = golden["c1_full"]
    _say(f"[gpu] batch chunk {chunk}")
    clip, stride = case_clip(case)
    det = api.Detector(case["width"], case["height"], **case_params(case))
    frames = [clip.render(t, stride=stride) for t in range(len(case["frames"]))]
    got = []
    for i in range(0, len(frames), chunk):
        got += det.detect_batch(frames[i:i + chunk])
    for t, fr in enumerate(case["frames"]):
        assert got[t] == golden_motions(fr), f"frame {t}"
    det.close()


def test_batch_with_tripod_crossing_batches(golden):
    case = golden["c1_tripod"]
    clip, stride = case_clip(case)
    det = api.Detector(case["width"], case["height"], **case_params(case))
    frames = [clip.render(t, stride=stride) for t in range(len(case["frames"]))]
    got = det.detect_batch(frames[:2]) + det.detect_batch(frames[2:5]) + det.detect_batch(frames[5:])
    for t, fr in enumerate(case["frames"]):
        assert got[t] == golden_motions(fr), f"frame {t}"
    det.close()


def test_device_resident_entry_point(golden):
    import torch
    case = golden["c2_1080p"]
    _say("[gpu] device-resident 1080p")
    clip, stride = case_clip(case)
    host = np.stack([clip.render(t, stride=stride) for t in range(len(case["frames"]))])
    dev = torch.from_numpy(host).to("cuda:0")
    torch.cuda.synchronize()
    det = api.Detector(case["width"], case["height"], **case_params(case))
    fb = host.shape[1] * host.shape[2]
    got = det.detect_device([dev.data_ptr() + i * fb for i in range(host.shape[0])],
                            [stride] * host.shape[0], keep=True)
    for t, fr in enumerate(case["frames"]):
        assert got[t] == golden_motions(fr), f"frame {t}"
    det.close()


@pytest.mark.parametrize("config,clip_id,nframes,step", [(1, 31, 8, 6), (1, 32, 6, 4), (1, 33, 6, 2),
                                                        (4, 7, 3, 6), (4, 9, 3, 6)])
def test_matches_oracle_on_fresh_seeds(config, clip_id, nframes, step):
    clip = Clip(config, clip_id)
    s = clip.spec
    _say(f"[gpu] oracle seed config {config} clip {clip_id} {s.width}x{s.height} step {step}")
    kw = dict(shakiness=s.shakiness, accuracy=s.accuracy, stepsize=step, mincontrast=s.mincontrast)
    gpu = api.Detector(s.width, s.height, **kw)
    ref = pyoracle.Detector(s.width, s.height, **kw)
    for t in range(nframes):
        luma = clip.render(t)
        assert gpu.detect(luma) == ref.detect(luma), f"frame {t}"
        assert np.array_equal(gpu.blurred(), ref.blurred()), f"blur frame {t}"
    gpu.close()


@pytest.mark.parametrize("width,height,step,pformat", [(72, 72, 6, api.PF_GRAY8), (75, 75, 7, api.PF_GRAY8),
                                                       (66, 66, 1, api.PF_GRAY8), (74, 90, 6, api.PF_YUV420P),
                                                       (75, 75, 6, api.PF_RGB24), (66, 66, 1, api.PF_RGB24)])
def test_smallest_accepted_frames_match_oracle(width, height, step, pformat):
    """frames at the edge of what vsMotionDetectInit accepts: a 3 x 3 grid of
    fields with little or no spacing, search windows that touch the frame's
    edges (tests/test_small_frames_cpu.py covers the refusal)"""
    clip = Clip(1, 61, width=width, height=height, nframes=6)
    bpp = 3 if pformat == api.PF_RGB24 else 1
    _say(f"[gpu] smallest frames {width}x{height} step {step} bpp {bpp}")
    kw = dict(shakiness=5, accuracy=6, stepsize=step, mincontrast=0.1)
    gpu = api.Detector(width, height, pformat=pformat, **kw)
    ref = pyoracle.Detector(width, height, bpp=bpp, **kw)
    found = 0
    for t in range(6):
        frame = clip.render(t) if bpp == 1 else pack(clip.render(t), bpp)
        want = ref.detect(frame)
        assert gpu.detect(frame) == want, f"frame {t}"
        assert np.array_equal(gpu.blurred(), ref.blurred()), f"working plane frame {t}"
        found += len(want)
    gpu.close()
    ref.close()
    assert found > 0


def test_flat_frames_give_empty_lists():
    """no contrast anywhere: every frame yields an empty LocalMotions list"""
    det = api.Detector(640, 360)
    flat = np.full((360, 640), 77, dtype=np.uint8)
    for _ in range(3):
        assert det.detect(flat) == []
    det.close()


@pytest.mark.parametrize("lds", [0, 1])
@pytest.mark.parametrize("K", [4, 6, 7, 8])
@pytest.mark.parametrize("name", ["c1_step5", "c2_1080p", "c3_4k"])
def test_every_search_plan_is_bit_identical(golden, monkeypatch, name, K, lds):
    """the autotuner may pick any plan (K rows per lane -- 7 is the
    three-column tile where it applies, the four-column K = 4 elsewhere --,
    field by scalar loads or staged in LDS for fields up to 128 px): each
    must reproduce the fixture"""
    monkeypatch.setenv("VIDSTAB_AMD_SEARCH_K", str(K))
    monkeypatch.setenv("VIDSTAB_AMD_SEARCH_LDS", str(lds))
    case = golden[name]
    _say(f"[gpu] plan K={K} lds={lds} fixture {name}")
    clip, stride = case_clip(case)
    det = api.Detector(case["width"], case["height"], **case_params(case))
    frames = [clip.render(t, stride=stride) for t in range(len(case["frames"]))]
    got = det.detect_batch(frames)
    for t, fr in enumerate(case["frames"]):
        assert got[t] == golden_motions(fr), f"{name} K={K} lds={lds} frame {t}"
    det.close()
