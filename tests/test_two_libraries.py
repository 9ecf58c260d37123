"""Both vid.stab libraries in one process, linked as one ffmpeg binary with the
vidstabdetect and vidstabtransform filters would be.

tests/c/two_libraries.c (built by oracle/Makefile against the reference's
headers) replaces the allocation / log hooks and remaps VS_OK / VS_ERROR as
ffmpeg's vidstabutils does, runs the detect filter's call sequence, writes
the .trf stream, then reads it back, fits the transforms and smooths them
through the transform-side calls.  It is linked twice: with this
library ahead of the reference library (detection binds here, the transform
side -- vsLocalmotions2Transforms, the lens estimate and the per-frame fits --
here as well, the rest of the transform side (reading the .trf, smoothing)
to the reference, the shared helpers and hook variables to one definition),
and with the reference library alone.  Both runs must print the
same .trf bytes and the same transforms bit for bit, with every hook
allocation freed.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(REPO, "oracle", "_ref")


def _run(exe, out, frames):
    print(f"[gpu] {os.path.basename(exe)} {frames} frames", file=sys.stderr, flush=True)
    p = subprocess.run([exe, out, str(frames)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, f"{exe}: rc {p.returncode}\n{p.stdout}\n{p.stderr}"
    return p.stdout.splitlines()


def test_mixed_process_matches_the_reference_alone(tmp_path):
    amd, ref = os.path.join(REF, "two_libs_amd"), os.path.join(REF, "two_libs_ref")
    if not (os.path.exists(amd) and os.path.exists(ref)):
        pytest.skip("two-library programs not built (make -C oracle ref needs the reference sources)")
    frames = 20
    mixed = _run(amd, str(tmp_path / "amd.trf"), frames)
    alone = _run(ref, str(tmp_path / "ref.trf"), frames)
    assert mixed[0] == "vsMotionDetection from libvidstab_amd.so"
    assert alone[0] == "vsMotionDetection from libvidstab.so"
    # the second pass's entry binds here too: the lens estimate and the
    # per-frame fits run on the GPU
    assert mixed[1] == "vsLocalmotions2Transforms from libvidstab_amd.so"
    assert alone[1] == "vsLocalmotions2Transforms from libvidstab.so"
    assert mixed[-1] == alone[-1] == "hook allocations all freed"
    assert f"frames read {frames}, transforms {frames}" in alone
    assert mixed[2:] == alone[2:]
