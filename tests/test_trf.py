"""The .trf writer (vsPrepareFile / vsWriteToFile, reference
src/serialize.c:144-190, 252-280, 355-409) against the reference's own
output: the golden fixtures hold FNV-1a hashes and lengths of the binary and
ASCII .trf streams that the reference wrote for its LocalMotions.

On the CPU the fixture's LocalMotions go through this library's writer; on
the GPU the whole pipeline does (GPU detect -> writer), which is the
end-to-end byte identity a vidstabtransform consumer sees."""
import ctypes
import os

import numpy as np
import pytest

from golden_util import case_clip, case_params, hexbits, bits_to_double
from vidstab_amd import api

ASCII, BINARY = 1, 2
_libc = ctypes.CDLL("libc.so.6")
_libc.fopen.restype = ctypes.c_void_p
_libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
_libc.fclose.argtypes = [ctypes.c_void_p]


def fnv1a(data):
    h = 0xCBF29CE484222325
    for b in data:
        h ^= b
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def _vector(lms):
    """LocalMotions from fixture tuples; returns (vector, keep-alive)"""
    n = len(lms)
    arr = (api.LocalMotion * max(n, 1))()
    ptrs = (ctypes.c_void_p * max(n, 1))()
    for i, m in enumerate(lms):
        lm = arr[i]
        lm.v.x, lm.v.y, lm.f.x, lm.f.y, lm.f.size = m[0], m[1], m[2], m[3], m[4]
        lm.contrast = bits_to_double(m[5])
        lm.match = bits_to_double(m[6])
        ptrs[i] = ctypes.cast(ctypes.byref(lm), ctypes.c_void_p)
    vec = api.VSVector()
    vec.data = ctypes.cast(ptrs, ctypes.POINTER(ctypes.c_void_p))
    vec.buffersize = max(n, 1)
    vec.nelems = n
    return vec, (arr, ptrs)


def _md_for(case, mode):
    md = api.VSMotionDetect()
    shak, acc, step = case["md_conf"]
    md.conf.shakiness, md.conf.accuracy, md.conf.stepSize = shak, acc, step
    md.conf.contrastThreshold = bits_to_double(hexbits(case["mincontrast"]))
    md.serializationMode = mode
    return md


def _write(path, case, mode, frames):
    L = api.lib()
    md = _md_for(case, mode)
    f = _libc.fopen(path.encode(), b"wb")
    assert f
    assert L.vsPrepareFile(ctypes.byref(md), ctypes.c_void_p(f)) == 0
    for t, lms in enumerate(frames):
        md.frameNum = t + 1
        vec, keep = _vector(lms)
        assert L.vsWriteToFile(ctypes.byref(md), ctypes.c_void_p(f), ctypes.byref(vec)) == 0
    _libc.fclose(ctypes.c_void_p(f))
    with open(path, "rb") as fh:
        return fh.read()


def _fixture_frames(case):
    return [[(m[0], m[1], m[2], m[3], m[4], hexbits(m[5]), hexbits(m[6])) for m in fr["lm"]]
            for fr in case["frames"]]


@pytest.mark.parametrize("name", ["c1_full", "c1_tripod", "c1_step5", "odd_502x298", "c2_1080p", "c3_4k"])
def test_writer_reproduces_reference_trf(golden, tmp_path, name):
    case = golden[name]
    frames = _fixture_frames(case)
    for mode, key in ((BINARY, "trf_bin"), (ASCII, "trf_ascii")):
        data = _write(str(tmp_path / f"{name}.{mode}.trf"), case, mode, frames)
        assert len(data) == case[key + "_len"], f"{name} {key} length"
        assert "%016x" % fnv1a(data) == case[key], f"{name} {key} bytes"


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1_full", "c2_1080p"])
def test_gpu_detect_then_writer_is_byte_identical(golden, tmp_path, name):
    case = golden[name]
    clip, stride = case_clip(case)
    det = api.Detector(case["width"], case["height"], **case_params(case))
    frames = det.detect_batch([clip.render(t, stride=stride) for t in range(len(case["frames"]))])
    det.close()
    for mode, key in ((BINARY, "trf_bin"), (ASCII, "trf_ascii")):
        data = _write(str(tmp_path / f"{name}.{mode}.trf"), case, mode, frames)
        assert len(data) == case[key + "_len"] and "%016x" % fnv1a(data) == case[key], f"{name} {key}"


_DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "vid.stab_amd", "lib", "vidstabdetect_like")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1_full", "c1_stride", "c1_tripod", "odd_502x298", "c2_1080p", "c3_4k"])
def test_ffmpeg_call_sequence_in_c_is_byte_identical(golden, tmp_path, name):
    """tests/c/vidstabdetect_like.c drives the library exactly as ffmpeg's
    vidstabdetect filter does (vsFrameInfoInit, vsMotionDetectInit,
    vsPrepareFile, then vsMotionDetection / vsWriteToFile / vs_vector_del per
    frame, vsMotionDetectionCleanup): its .trf, binary and ASCII, must be the
    reference's byte for byte"""
    import subprocess
    assert os.path.exists(_DRIVER), "the C driver is not built (make -C vid.stab_amd)"
    case = golden[name]
    clip, stride = case_clip(case)
    n = len(case["frames"])
    raw = tmp_path / "frames.raw"
    with open(raw, "wb") as fh:
        for t in range(n):
            fh.write(np.ascontiguousarray(clip.render(t, stride=stride), dtype=np.uint8).tobytes())
    p = case_params(case)
    for mode, key in ((BINARY, "trf_bin"), (ASCII, "trf_ascii")):
        out = tmp_path / f"{name}.{mode}.trf"
        args = [_DRIVER, str(raw), str(case["width"]), str(case["height"]), str(n), str(stride),
                str(p["shakiness"]), str(p["accuracy"]), str(p["stepsize"]), case["mincontrast"],
                str(p["tripod"]), str(mode), str(out)]
        r = subprocess.run(args, capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stderr
        data = out.read_bytes()
        assert len(data) == case[key + "_len"] and "%016x" % fnv1a(data) == case[key], f"{name} {key}"
