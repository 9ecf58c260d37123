"""The allocation-free second-pass entry point (vsg_motions_to_transform_static)
against the allocating one (vsg_motions_to_transform), through the C ABI.

Both run the same kernel; the static variant stages inputs and outputs
through the library's fixed device array instead of per-call allocations.
Bar: every transform, residual and round count bit-identical, on two calls
in a row, the second with a smaller input (the array keeps the first call's
contents, which must not leak into the second).
"""
import ctypes
import sys

import numpy as np
import pytest

from golden_util import golden_motions
from vidstab_amd import api

c_int, c_void_p = ctypes.c_int, ctypes.c_void_p


def _bind():
    L = api.lib()
    for name in ("vsg_motions_to_transform", "vsg_motions_to_transform_static"):
        f = getattr(L, name)
        f.argtypes = [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p,
                      c_void_p]
        f.restype = c_int
    return L


def _arrays(frames):
    """fixture frames -> (lm int16 [n][4], match float64 [n], count int32 [frames])"""
    lm, match, count = [], [], []
    for fr in frames:
        ms = golden_motions(fr)
        count.append(len(ms))
        for (vx, vy, fx, fy, _size, _cbits, mbits) in ms:
            lm.append((fx, fy, vx, vy))
            match.append(mbits)
    lm = np.array(lm, dtype=np.int16).reshape(-1, 4)
    match = np.array(match, dtype=np.uint64).view(np.float64)
    return np.ascontiguousarray(lm), np.ascontiguousarray(match), np.array(count, dtype=np.int32)


def _run(L, name, w, h, arrays, smooth_zoom=1):
    lm, match, count = arrays
    n = len(count)
    out = (api.VSTransform * n)()
    res = np.zeros(n, dtype=np.float64)
    rounds = np.zeros(n, dtype=np.int32)
    rc = getattr(L, name)(0, w, h, lm.ctypes.data, match.ctypes.data, count.ctypes.data, n, smooth_zoom,
                          ctypes.addressof(out), res.ctypes.data, rounds.ctypes.data)
    assert rc == 0, L.vsg_last_error().decode()
    t = np.frombuffer(bytes(out), dtype=np.uint8).reshape(n, -1)[:, :48]   # the six doubles
    extra = [o.extra for o in out]
    return t.copy(), extra, res.view(np.uint64).copy(), rounds.copy()


@pytest.mark.gpu
def test_static_entry_point_matches_allocating_one(golden):
    L = _bind()
    first, second = golden["c1_full"], golden["c1_stride"]
    a1, a2 = _arrays(first["frames"]), _arrays(second["frames"])
    assert len(a2[2]) < len(a1[2]) and len(a2[0]) < len(a1[0])
    for label, case, arr in (("first call, 60 frames", first, a1), ("second call, 8 frames", second, a2)):
        print(f"[gpu] no-alloc second pass: {label}", file=sys.stderr, flush=True)
        want = _run(L, "vsg_motions_to_transform", case["width"], case["height"], arr)
        got = _run(L, "vsg_motions_to_transform_static", case["width"], case["height"], arr)
        assert np.array_equal(got[0], want[0]), f"{label}: transforms differ"
        assert got[1] == want[1] and np.array_equal(got[2], want[2]) and np.array_equal(got[3], want[3]), label


def test_static_entry_point_rejects_oversized_inputs_before_touching_the_gpu():
    """the limits are checked on the host (no device call is made)"""
    L = _bind()
    count = np.ones(8193, dtype=np.int32)
    lm = np.zeros((8193, 4), dtype=np.int16)
    match = np.zeros(8193, dtype=np.float64)
    out = (api.VSTransform * 8193)()
    assert L.vsg_motions_to_transform_static(0, 640, 360, lm.ctypes.data, match.ctypes.data, count.ctypes.data,
                                             8193, 0, ctypes.addressof(out), None, None) != 0
    assert b"8192 frames" in L.vsg_last_error()
    count = np.full(2, 1 << 17, dtype=np.int32)
    count[1] += 1
    assert L.vsg_motions_to_transform_static(0, 640, 360, lm.ctypes.data, match.ctypes.data, count.ctypes.data,
                                             2, 0, ctypes.addressof(out), None, None) != 0
    assert b"262144 motions" in L.vsg_last_error()
