"""The batched multi-clip form (vsMotionDetectMulti) against per-clip plain
runs of the same frames: every clip's LocalMotions must be bit-identical to
what one vsMotionDetection stream per clip returns (which the fixtures of
test_gpu_parity.py pin to the reference).  Clips of uneven size and length
go through one call; small batch sizes leave idle slots and shorter clips
finishing early, which the fixed-shape graphs must skip."""
import sys

import numpy as np
import pytest

from vidstab_amd import api
from vidstab_amd.synth import Clip

pytestmark = pytest.mark.gpu

# (config, clip, width, height, frames, stepsize); a mix of geometries, two
# clips sharing a geometry (one search launch), lengths 1 .. 11
CLIPS = [(4, 40, 640, 360, 11, 6), (4, 41, 1280, 720, 7, 6), (4, 42, 640, 360, 5, 6),
         (1, 43, 502, 298, 9, 6), (4, 44, 1920, 1080, 4, 6), (1, 45, 640, 360, 1, 6),
         (1, 46, 640, 360, 6, 4)]


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _device_clips(clips):
    import torch
    out, keep = [], []
    for config, cid, w, h, n, step in clips:
        clip = Clip(config, cid, width=w, height=h, nframes=n)
        s = clip.spec
        host = np.stack([clip.render(t) for t in range(n)])
        dev = torch.from_numpy(host).to("cuda:0")
        keep.append(dev)
        fb = w * h
        out.append(dict(width=w, height=h, shakiness=s.shakiness, accuracy=s.accuracy, stepsize=step,
                        mincontrast=s.mincontrast, linesize=w,
                        ptrs=[dev.data_ptr() + t * fb for t in range(n)]))
    torch.cuda.synchronize()
    return out, keep


@pytest.mark.parametrize("batch", [8, 3, 1])
def test_multi_clip_equals_per_clip(monkeypatch, batch):
    monkeypatch.setenv("VIDSTAB_AMD_MULTI_BATCH", str(batch))
    _say(f"[gpu] multi-clip, {len(CLIPS)} clips, batch {batch}")
    clips, keep = _device_clips(CLIPS)
    got = api.detect_multi(clips)
    assert len(got) == len(clips)
    for c, cl in enumerate(clips):
        det = api.Detector(cl["width"], cl["height"], shakiness=cl["shakiness"], accuracy=cl["accuracy"],
                           stepsize=cl["stepsize"], mincontrast=cl["mincontrast"])
        want = det.detect_device(cl["ptrs"], [cl["linesize"]] * len(cl["ptrs"]), keep=True)
        det.close()
        assert len(got[c]) == len(want)
        for t in range(len(want)):
            assert got[c][t] == want[t], f"clip {c} frame {t}"
        assert got[c][0] == []
        if len(want) > 1:
            assert len(got[c][1]) > 0
    del keep


def test_kept_state_between_calls(monkeypatch):
    """a call over clips of the same shapes as the previous call reuses its
    device state: the frames change (here: each clip played backwards), the
    results follow them -- equal to per-clip runs of the new frames -- and
    going back to the first frames gives the first results again"""
    monkeypatch.setenv("VIDSTAB_AMD_MULTI_BATCH", "3")
    clips, keep = _device_clips(CLIPS)
    first = api.detect_multi(clips)
    back = [dict(cl, ptrs=cl["ptrs"][::-1]) for cl in clips]
    second = api.detect_multi(back)
    for c, cl in enumerate(back):
        det = api.Detector(cl["width"], cl["height"], shakiness=cl["shakiness"], accuracy=cl["accuracy"],
                           stepsize=cl["stepsize"], mincontrast=cl["mincontrast"])
        want = det.detect_device(cl["ptrs"], [cl["linesize"]] * len(cl["ptrs"]), keep=True)
        det.close()
        assert second[c] == want, f"clip {c}"
    assert api.detect_multi(clips) == first
    del keep


@pytest.mark.parametrize("batch", [16, 3])
def test_multi_clip_tripod_equals_per_clip(monkeypatch, batch):
    """virtualTripod in the multi-clip form: per clip, the lead-in frames
    emit the empty vector, the reference frame's plane is frozen in the batch
    that holds it, and every later frame is measured against it -- equal to a
    per-clip tripod stream.  Reference frames at the first frame, inside a
    batch, on a batch boundary (batch 3: frame 2, frame 5), past the clip's
    end, and a clip without tripod in the same call"""
    monkeypatch.setenv("VIDSTAB_AMD_MULTI_BATCH", str(batch))
    tripods = [3, 1, 6, 0, 2, 9, 4]
    _say(f"[gpu] multi-clip tripod {tripods}, batch {batch}")
    clips, keep = _device_clips(CLIPS)
    for cl, tp in zip(clips, tripods):
        cl["tripod"] = tp
    got = api.detect_multi(clips)
    for c, cl in enumerate(clips):
        det = api.Detector(cl["width"], cl["height"], shakiness=cl["shakiness"], accuracy=cl["accuracy"],
                           stepsize=cl["stepsize"], mincontrast=cl["mincontrast"], tripod=cl["tripod"])
        want = det.detect_device(cl["ptrs"], [cl["linesize"]] * len(cl["ptrs"]), keep=True)
        det.close()
        assert len(got[c]) == len(want)
        for t in range(len(want)):
            assert got[c][t] == want[t], f"clip {c} (tripod {cl['tripod']}) frame {t}"
            if t < max(cl["tripod"], 1):
                assert got[c][t] == [], f"clip {c} lead-in frame {t}"
    # clip 0 (11 frames, tripod 3) measures frames 3.. against frame 2
    assert any(len(m) > 0 for m in got[0][3:])
    # the same clips without tripod next (kept device state): the chains again
    for cl in clips:
        cl["tripod"] = 0
    plain = api.detect_multi(clips)
    for c, cl in enumerate(clips):
        det = api.Detector(cl["width"], cl["height"], shakiness=cl["shakiness"], accuracy=cl["accuracy"],
                           stepsize=cl["stepsize"], mincontrast=cl["mincontrast"])
        assert plain[c] == det.detect_device(cl["ptrs"], [cl["linesize"]] * len(cl["ptrs"]), keep=True)
        det.close()
    del keep
