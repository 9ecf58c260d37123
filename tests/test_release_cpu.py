"""The bulk release of LocalMotions arrays (vid.stab_amd/csrc/host/
vs_release.c): several caller threads release arrays of random lengths and
reuse them at once, built plain and under ThreadSanitizer (host code only,
gcc).  Every block the vector code allocated through the vs_malloc hooks must
come back through vs_free exactly once by the time vsLocalMotionsReleaseWait
returns, and every released vector must be empty on return
(tests/c/release_stress.c)."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(REPO, "vid.stab_amd", "csrc", "host")
SRC = [os.path.join(REPO, "tests", "c", "release_stress.c"), os.path.join(HOST, "vs_release.c"),
       os.path.join(HOST, "vs_vector.c"), os.path.join(HOST, "vs_glue.c")]


def _build(tmp_path, name, flags):
    exe = str(tmp_path / name)
    subprocess.run(["gcc", *flags, "-std=gnu11", "-I", os.path.join(REPO, "include"), "-o", exe, *SRC,
                    "-lpthread"], check=True)
    return exe


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
@pytest.mark.parametrize("threads,rounds,seed", [(1, 3000, 1), (4, 2000, 2), (8, 1000, 3)])
def test_release_returns_every_block_once(tmp_path, threads, rounds, seed):
    exe = _build(tmp_path, "release_o2", ["-O2"])
    p = subprocess.run([exe, str(threads), str(rounds), str(seed)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_release_under_thread_sanitizer(tmp_path):
    exe = _build(tmp_path, "release_tsan", ["-O1", "-g", "-fsanitize=thread"])
    p = subprocess.run([exe, "4", "500", "5"], capture_output=True, text=True, timeout=180)
    if "FATAL: ThreadSanitizer: unexpected memory mapping" in p.stderr:
        pytest.skip("ThreadSanitizer cannot map its shadow memory in this environment")
    assert p.returncode == 0 and "WARNING: ThreadSanitizer" not in p.stderr, p.stderr[-3000:]
