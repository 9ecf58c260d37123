"""GPU parity of the multi-batch pipeline and of whole workload clips.

One detect call on more frames than a device batch runs the pipelined path:
three slot sets in flight (on one shared compute stream, or on one stream
each with VIDSTAB_AMD_STREAMS=3), batch k+1's coarse results
turned into fine jobs before the host waits for batch k, cross-set event
ordering, and slot sets reused after three batches.  These tests drive that
path with small batches on the fixtures, and at the default batch on the
full-length workload clips against per-frame hashes written by the reference
(tests/golden/clip_hashes.json: input luma, LocalMotions count and hash,
blurred plane).  Bar: bit-identical LocalMotions and blurred planes.
"""
import sys

import numpy as np
import pytest

import pyoracle
from golden_util import case_clip, case_params, golden_motions, hash_motions, load_clip_hashes
from vidstab_amd import api
from vidstab_amd.synth import Clip, hash_plane

pytestmark = pytest.mark.gpu

KEPT_BATCHES = 3   # the detector keeps the planes of its latest three device batches


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def clip_hashes():
    return load_clip_hashes()


def _to_device(frames):
    import torch
    host = np.stack(frames)
    dev = torch.from_numpy(host).to("cuda:0")
    torch.cuda.synchronize()
    fb = host.shape[1] * host.shape[2]
    return dev, [dev.data_ptr() + i * fb for i in range(host.shape[0])]


def _kept_frames(n, batch):
    """frame indices whose blurred planes are still on the device after one
    call over n frames in batches of `batch`"""
    starts = list(range(0, n, batch))[-KEPT_BATCHES:]
    return list(range(starts[0], n)) if starts else []


@pytest.mark.parametrize("entry", ["host", "device"])
@pytest.mark.parametrize("name", ["c1_full", "c1_tripod"])
@pytest.mark.parametrize("batch", [4, 7])
def test_multi_batch_call_matches_reference(golden, monkeypatch, batch, name, entry):
    """all frames of a fixture in ONE call at a small device batch"""
    monkeypatch.setenv("VIDSTAB_AMD_MAX_BATCH", str(batch))
    case = golden[name]
    n = len(case["frames"])
    _say(f"[gpu] pipeline {name}: {n} frames, batch {batch}, {entry} frames")
    clip, stride = case_clip(case)
    frames = [clip.render(t, stride=stride) for t in range(n)]
    det = api.Detector(case["width"], case["height"], **case_params(case))
    if entry == "host":
        got = det.detect_batch(frames)
    else:
        dev, ptrs = _to_device(frames)
        got = det.detect_device(ptrs, [stride] * n, keep=True)
    for t, fr in enumerate(case["frames"]):
        assert got[t] == golden_motions(fr), f"{name} batch {batch} frame {t}"
    for t in _kept_frames(n, batch):
        assert "%016x" % hash_plane(det.blurred(n - 1 - t)) == case["frames"][t]["blur"], f"blur frame {t}"
    det.close()


def test_wide_box_blur_across_batches():
    """stepsize 80 at accuracy 6 gives an 81-px box, blurred by the two-pass
    kernels through per-slot intermediate planes; three batches of two frames
    blur concurrently on the sets' streams"""
    import os
    old = os.environ.get("VIDSTAB_AMD_MAX_BATCH")
    os.environ["VIDSTAB_AMD_MAX_BATCH"] = "2"
    try:
        clip = Clip(1, 40)
        s = clip.spec
        kw = dict(shakiness=5, accuracy=6, stepsize=80, mincontrast=s.mincontrast)
        frames = [clip.render(t) for t in range(8)]
        _say("[gpu] wide box 81 px, batch 2, 8 frames")
        gpu = api.Detector(s.width, s.height, **kw)
        got = gpu.detect_batch(frames)
        ref = pyoracle.Detector(s.width, s.height, **kw)
        blurred = []
        for t, luma in enumerate(frames):
            assert got[t] == ref.detect(luma), f"frame {t}"
            blurred.append(ref.blurred())
        for t in _kept_frames(8, 2):
            assert np.array_equal(gpu.blurred(7 - t), blurred[t]), f"blur frame {t}"
        gpu.close()
    finally:
        if old is None:
            del os.environ["VIDSTAB_AMD_MAX_BATCH"]
        else:
            os.environ["VIDSTAB_AMD_MAX_BATCH"] = old


def _check_hashes(name, case, got, first=0):
    for k, m in enumerate(got):
        t = first + k
        hin, cnt, hl, _ = case["frames"][t]
        assert len(m) == cnt, f"{name} frame {t}: {len(m)} vs {cnt} motions"
        assert "%016x" % hash_motions(m) == hl, f"{name} frame {t}: LocalMotions differ"


def _whole_clip(case):
    clip = Clip(case["config"], case["clip"])
    n = len(case["frames"])
    frames = [clip.render(t) for t in range(n)]
    for t in (0, n - 1):
        assert "%016x" % hash_plane(frames[t]) == case["frames"][t][0], f"input frame {t}"
    return frames


@pytest.mark.parametrize("name", ["w2_full", "w3_full"])
def test_whole_workload_clip_device_frames(clip_hashes, name):
    """the 300-frame 1080p and 240-frame 4K step-4 clips through
    vsMotionDetectionDevice at the default batch of 64 (what the benchmark
    times): every frame's LocalMotions and the kept blurred planes"""
    case = clip_hashes[name]
    _say(f"[gpu] whole clip {name} {case['width']}x{case['height']} x {len(case['frames'])}, device frames")
    frames = _whole_clip(case)
    n = len(frames)
    dev, ptrs = _to_device(frames)
    det = api.Detector(case["width"], case["height"], **case_params(case))
    got = det.detect_device(ptrs, [case["stride"]] * n, keep=True)
    _check_hashes(name, case, got)
    for t in _kept_frames(n, 64):
        assert "%016x" % hash_plane(det.blurred(n - 1 - t)) == case["frames"][t][3], f"{name} blur frame {t}"
    det.close()


def test_whole_1080p_clip_host_frames_and_per_frame_calls(clip_hashes):
    """the same 1080p clip as pageable host frames: one vsMotionDetectionBatch
    call, then one vsMotionDetection call per frame (ffmpeg's pattern) with
    the blurred plane of every frame"""
    case = clip_hashes["w2_full"]
    frames = _whole_clip(case)
    _say("[gpu] whole clip w2_full, host frames, batch call")
    det = api.Detector(case["width"], case["height"], **case_params(case))
    _check_hashes("w2_full", case, det.detect_batch(frames))
    det.close()
    _say("[gpu] whole clip w2_full, one call per frame")
    det = api.Detector(case["width"], case["height"], **case_params(case))
    for t, luma in enumerate(frames):
        _check_hashes("w2_full", case, [det.detect(luma)], first=t)
        assert "%016x" % hash_plane(det.blurred()) == case["frames"][t][3], f"blur frame {t}"
    det.close()


def test_multi_clip_workload_matches_reference(clip_hashes):
    """the 64 clips of the batched multi-clip workload (360p to 2160p, 30 to
    120 frames) in one vsMotionDetectMulti call: every frame of every clip
    against the reference's hashes"""
    import torch
    names = sorted((k for k in clip_hashes if k.startswith("w4_clip")), key=lambda k: int(k[7:]))
    assert len(names) == 64
    clips, keep = [], []
    for name in names:
        case = clip_hashes[name]
        clip = Clip(case["config"], case["clip"])
        n = len(case["frames"])
        host = np.empty((n, case["height"], case["width"]), dtype=np.uint8)
        for t in range(n):
            clip.render(t, out=host[t])
        assert "%016x" % hash_plane(host[n - 1]) == case["frames"][n - 1][0], f"{name} input"
        dev = torch.from_numpy(host).to("cuda:0")
        keep.append(dev)
        fb = case["width"] * case["height"]
        clips.append(dict(width=case["width"], height=case["height"], shakiness=case["shakiness"],
                          accuracy=case["accuracy"], stepsize=case["stepsize"],
                          mincontrast=float(np.array([int(case["mincontrast"], 16)], dtype=np.uint64)
                                            .view(np.float64)[0]),
                          linesize=case["width"], ptrs=[dev.data_ptr() + t * fb for t in range(n)]))
    torch.cuda.synchronize()
    _say(f"[gpu] multi-clip workload: 64 clips, {sum(len(c['ptrs']) for c in clips)} frames")
    got = api.detect_multi(clips, keep=True)
    for name, motions in zip(names, got):
        _check_hashes(name, clip_hashes[name], motions)
