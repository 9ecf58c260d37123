"""The detect pass's host thread pool (vid.stab_amd/csrc/host/vs_pool.c): many
back-to-back runs of random sizes on two pools, runs smaller than the pool
(workers sitting a run out), pauses past the spin window (sleeping workers),
built once plain and once under ThreadSanitizer (host code only, gcc).  Each
run must execute every index exactly once and publish its writes to the
caller (tests/c/pool_stress.c)."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(REPO, "tests", "c", "pool_stress.c"), os.path.join(REPO, "vid.stab_amd", "csrc", "host", "vs_pool.c")]
INC = os.path.join(REPO, "vid.stab_amd", "csrc", "host")


def _build(tmp_path, name, flags):
    exe = str(tmp_path / name)
    subprocess.run(["gcc", *flags, "-I", INC, "-o", exe, *SRC, "-lpthread"], check=True)
    return exe


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
@pytest.mark.parametrize("threads,runs,seed", [(12, 6000, 1), (4, 6000, 2), (2, 3000, 3)])
def test_pool_runs_every_item_once(tmp_path, threads, runs, seed):
    exe = _build(tmp_path, "pool_o2", ["-O2"])
    p = subprocess.run([exe, str(threads), str(runs), str(seed)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_pool_under_thread_sanitizer(tmp_path):
    exe = _build(tmp_path, "pool_tsan", ["-O1", "-g", "-fsanitize=thread"])
    p = subprocess.run([exe, "12", "2000", "5"], capture_output=True, text=True, timeout=180)
    if "FATAL: ThreadSanitizer: unexpected memory mapping" in p.stderr:
        pytest.skip("ThreadSanitizer cannot map its shadow memory in this environment")
    assert p.returncode == 0 and "WARNING: ThreadSanitizer" not in p.stderr, p.stderr[-3000:]
