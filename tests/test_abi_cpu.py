"""CPU-only checks of the drop-in boundary: exported symbols, struct layouts,
host helpers, and error paths that need no GPU."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from conftest import REPO
import pyoracle
from vidstab_amd import api

HEADERS = sorted(glob.glob(os.path.join(REPO, "include", "*.h")))


def _declared_symbols():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = "\n".join(l for l in txt.splitlines() if not l.lstrip().startswith("#"))
        for m in re.finditer(r"VS_API\s+[^;(]*?\b(\w+)\s*\(", txt):
            names.add(m.group(1))
        for m in re.finditer(r"extern\s+VS_API\s+[\w\s\*]+?\b(\w+)\s*;", txt):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = api.lib()
    names = _declared_symbols()
    assert len(names) > 60
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_library_exports_no_cpp_symbols():
    """the dynamic symbol table holds C names only: the C++ template
    instantiations of the host code stay local to the library"""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", api.lib_path("libvidstab_amd.so")], capture_output=True, text=True,
                         check=True).stdout
    names = [l.split()[-1] for l in out.splitlines() if l.strip()]
    assert len(names) > 100
    cxx = [n for n in names if n.startswith("_Z")]
    assert not cxx, cxx[:5]


def test_abi_struct_sizes_match_reference_layout():
    # sizes and offsets of the reference ABI on x86-64
    assert ctypes.sizeof(api.VSFrameInfo) == 28
    assert ctypes.sizeof(api.VSFrame) == 48
    assert ctypes.sizeof(api.VSMotionDetectConfig) == 48
    assert ctypes.sizeof(api.VSMotionDetectFields) == 104
    assert ctypes.sizeof(api.VSMotionDetect) == 496
    assert ctypes.sizeof(api.LocalMotion) == 32
    assert ctypes.sizeof(api.VSTransform) == 56
    assert api.VSMotionDetect.conf.offset == 32
    assert api.VSMotionDetect.fieldscoarse.offset == 80
    assert api.VSMotionDetect.fieldsfine.offset == 184
    assert api.VSMotionDetect.curr.offset == 288
    assert api.VSMotionDetect.prev.offset == 432
    assert api.VSMotionDetect.hasSeenOneFrame.offset == 480
    assert api.VSMotionDetect.frameNum.offset == 492
    assert api.LocalMotion.contrast.offset == 16 and api.LocalMotion.match.offset == 24


def test_second_pass_struct_layouts_match_reference():
    """the transform pass's structs that vsLocalmotions2Transforms takes, and
    the lens estimate's, as the reference lays them out on x86-64 (measured
    with offsetof on src/transform.h and src/lensdistortion.h in a release
    build): in include/ and in the Python mirrors"""
    import subprocess
    import tempfile
    want = {"VSTransformData": 720, "VSTransformData.conf": 216, "VSTransformData.lensMode": 328,
            "VSTransformData.lensK": 336, "VSTransformData.lensMaps": 352, "VSTransformData.initialized": 712,
            "VSTransformConfig": 112, "VSTransformConfig.modName": 56, "VSTransformConfig.estimateLensDistortion": 84,
            "VSTransformConfig.lensK": 96, "VSTransformConfig.fov": 104, "VSTransformations": 24,
            "VSTransformations.len": 12, "VSLensPlaneMap": 120, "VSLensEstimateConfig": 72,
            "VSLensEstimateConfig.f": 64, "VSLensEstimate": 48, "VSLensEstimate.determined": 44,
            "VSLensDistortion": 32}
    probe = "#include <stddef.h>\n#include <stdio.h>\n#include \"transform.h\"\n#include \"lensdistortion.h\"\n"
    probe += "int main(void) {\n"
    for key in want:
        if "." in key:
            t, m = key.split(".")
            probe += f'  printf("{key} %zu\\n", offsetof({t}, {m}));\n'
        else:
            probe += f'  printf("{key} %zu\\n", sizeof({key}));\n'
    probe += "  return 0;\n}\n"
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "probe.c")
        open(src, "w").write(probe)
        exe = os.path.join(tmp, "probe")
        subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), "-o", exe, src], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = {l.split()[0]: int(l.split()[1]) for l in out.splitlines()}
    assert got == want
    for key, v in want.items():
        if "." in key:
            t, m = key.split(".")
            assert getattr(getattr(api, t), m).offset == v, key
        elif hasattr(api, key):
            assert ctypes.sizeof(getattr(api, key)) == v, key


@pytest.mark.parametrize("n", [0, 1, 2, 8, 25, 77, 154])
def test_spiral_rank_closed_form_matches_walk(n):
    """the kernels' closed-form spiral rank (and its inverse) against the
    reference's step-by-step spiral, simulated by the oracle"""
    L = api.lib()
    g = 2 * n + 1
    a = (ctypes.c_int * (g * g))()
    b = (ctypes.c_int * (g * g))()
    pyoracle.lib().or_spiral_walk(n, a, b)
    seen = set()
    for r in range(g * g):
        assert L.vsg_spiral_rank(a[r], b[r]) == r
        ua, ub = ctypes.c_int(), ctypes.c_int()
        L.vsg_spiral_unrank(r, ctypes.byref(ua), ctypes.byref(ub))
        assert (ua.value, ub.value) == (a[r], b[r])
        seen.add((a[r], b[r]))
    assert len(seen) == g * g and all(abs(x) <= n and abs(y) <= n for x, y in seen)


def test_magic_division_is_exact():
    """floor(acc/size) == mulhi(acc, ceil(2^32/size)) for every blur accumulator"""
    for size in list(range(2, 300)) + [511, 1023, 2047, 4095]:
        mul = ((1 << 32) + size - 1) // size
        acc = np.arange(0, 255 * size + 1, dtype=np.uint64)
        q = (acc * np.uint64(mul)) >> np.uint64(32)
        assert np.array_equal(q, acc // np.uint64(size)), size


def _hip_device_count():
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return n.value if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 else 0
    except OSError:
        return 0


def test_init_rejects_unsupported_configs_without_touching_the_gpu():
    L = api.lib()
    for pf, step in ((api.PF_RGB24, 6), (api.PF_PACKED, 6), (api.PF_YUV420P, 0)):
        fi = api.VSFrameInfo()
        L.vsFrameInfoInit(ctypes.byref(fi), 640, 360, pf if pf != api.PF_PACKED else api.PF_YUV420P)
        fi.pFormat = pf
        md = api.VSMotionDetect()
        conf = L.vsMotionDetectGetDefaultConfig(b"t")
        conf.stepSize = step
        assert L.vsMotionDetectInit(ctypes.byref(md), ctypes.byref(conf), ctypes.byref(fi)) == -1


@pytest.mark.skipif(_hip_device_count() > 0, reason="a GPU is present")
def test_init_without_gpu_reports_vs_error():
    d = None
    with pytest.raises(RuntimeError):
        d = api.Detector(640, 360)
    assert d is None


def test_frameinfo_init_matches_reference_conventions():
    L = api.lib()
    fi = api.VSFrameInfo()
    assert L.vsFrameInfoInit(ctypes.byref(fi), 640, 360, api.PF_YUV420P) == 1
    assert (fi.planes, fi.log2ChromaW, fi.log2ChromaH, fi.bytesPerPixel) == (3, 1, 1, 1)
    assert L.vsFrameInfoInit(ctypes.byref(fi), 641, 360, api.PF_YUV420P) == 0
    assert L.vsFrameInfoInit(ctypes.byref(fi), 641, 361, api.PF_GRAY8) == 1
    assert L.vsFrameInfoInit(ctypes.byref(fi), 64, 64, api.PF_RGBA) == 1 and fi.bytesPerPixel == 4
