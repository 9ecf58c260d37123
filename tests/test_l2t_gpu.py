"""The second pass's LocalMotions -> transforms step on the GPU
(vsLocalmotions2Transforms, vsEstimateLensDistortion) against the reference's
own results on seeded synthetic clips (tests/golden/l2t_golden.json, written
by tests/golden/make_golden_l2t.c from the reference build): camera moves
under the similarity and the rotational model (fov 65 and 90 degrees), seen
through barrel and pincushion lenses or none, with outlier fields and an
empty frame.

Bars.
- vsMotionsToTransform with the rotational model (no lens): every double
  expression runs in the reference's compiled order, so the bar is the one of
  the similarity model (tests/test_transform_gpu.py): x, y within 1e-6 px,
  alpha within 1e-9 rad, zoom within 1e-6 %, extra equal -- and every frame of
  these fixtures bit-identical, asserted.
- The lens estimate: each frame's fit sums its points in another order than
  the reference (a workgroup reduction against a sequential loop), so the
  objective differs in its last bits and Brent's search, whose steps compare
  objective values, may take another path to the same minimum (a few
  evaluations more or fewer): k within 1e-7 (measured: within 2e-8), the
  determined flag, the rejected and used counts equal, residual within 1e-9
  relative, uncertainty within 1e-6 relative.
- Transforms through the estimated lens (vsLensMotionsToTransform at that k):
  x, y within 1e-8 px, alpha within 1e-11 rad, zoom within 1e-8 % (measured:
  3e-11 px, 4e-14 rad, 7e-12 %), extra equal; the lens strength left in the
  transform data within 1e-7.  Frames fitted without a lens, and the simple
  transform, bit-identical.
"""
import json
import os
import sys

import numpy as np
import pytest

from golden_util import bits_to_double, hexbits
from vidstab_amd import api

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL_FIT = (1e-6, 1e-6, 1e-9, 1e-6)
TOL_LENS = (1e-8, 1e-8, 1e-11, 1e-8)


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def l2t():
    with open(os.path.join(HERE, "golden", "l2t_golden.json")) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


def _frames(case):
    """motion tuples (v.x, v.y, f.x, f.y, size, contrast bits, match bits) per frame"""
    out = []
    for fr in case["frames"]:
        ms = []
        for i in range(0, len(fr), 5):
            fx, fy, vx, vy, match = fr[i:i + 5]
            mbits = int(np.array([match], dtype=np.float64).view(np.uint64)[0])
            cbits = int(np.array([0.5], dtype=np.float64).view(np.uint64)[0])
            ms.append((vx, vy, fx, fy, 32, cbits, mbits))
        out.append(ms)
    return out


def _want(row):
    return [bits_to_double(hexbits(v)) for v in row[:-1]] + [row[-1]]


def _bits(x):
    return int(np.array([x], dtype=np.float64).view(np.uint64)[0])


def test_rotational_model_fit_matches_reference(l2t):
    """vsMotionsToTransform with the case's fov (and fov 0), lens estimate off"""
    total = exact = 0
    for name, case in l2t.items():
        if case["simple"]:
            continue
        frames = _frames(case)
        _say(f"[gpu] l2t fit: case {name}, fov {case['fov']}, {len(frames)} frames")
        got, lens_k = api.localmotions_to_transforms(case["width"], case["height"], frames, fov=case["fov"],
                                                     estimate=0, smooth_zoom=case["smoothZoom"])
        assert lens_k == 0.0
        for t, (g, w) in enumerate(zip(got, case["motions_to_transform"])):
            want = _want(w)
            g4 = (g[0], g[1], g[2], g[3], g[5])
            assert all(abs(g4[i] - want[i]) <= TOL_FIT[i] for i in range(4)) and g4[4] == want[4], \
                f"{name} frame {t}: {g4} vs {want}"
            exact += all(_bits(g4[i]) == _bits(want[i]) for i in range(4))
            total += 1
    _say(f"[gpu] l2t fit: {exact}/{total} frames bit-identical")
    assert total > 100
    assert exact == total


def test_lens_estimate_matches_reference(l2t):
    for name, case in l2t.items():
        frames = _frames(case)
        focal = api.lib().focal_from_fov(case["fov"], case["width"])
        got = api.estimate_lens(case["width"], case["height"], frames, focal=focal)
        w = case["estimate_result"]
        want = {k: (bits_to_double(hexbits(v)) if isinstance(v, str) else v) for k, v in w.items()}
        _say(f"[gpu] l2t estimate {name}: k {got['k']:.8f} vs {want['k']:.8f}, unc {got['uncertainty']:.6f} vs "
             f"{want['uncertainty']:.6f}, residual {got['residual']:.8f} vs {want['residual']:.8f}, rejected "
             f"{got['rejected']} vs {want['rejected']}, evals {got['iterations']} vs {want['iterations']}")
        assert abs(got["k"] - want["k"]) <= 1e-7, name
        assert got["determined"] == want["determined"], name
        assert got["rejected"] == want["rejected"] and got["used"] == want["used"], name
        assert abs(got["residual"] - want["residual"]) <= 1e-9 * max(1.0, abs(want["residual"])), name
        assert abs(got["uncertainty"] - want["uncertainty"]) <= 1e-6 * abs(want["uncertainty"]), name


def test_localmotions_to_transforms_matches_reference(l2t):
    """the whole entry point with each case's transform configuration"""
    for name, case in l2t.items():
        frames = _frames(case)
        _say(f"[gpu] l2t entry point: case {name}")
        got, lens_k = api.localmotions_to_transforms(case["width"], case["height"], frames, fov=case["fov"],
                                                     estimate=case["estimate"], simple=case["simple"],
                                                     smooth_zoom=case["smoothZoom"])
        want_k = bits_to_double(hexbits(case["lensK"]))
        assert abs(lens_k - want_k) <= 1e-7, f"{name}: lensK {lens_k} vs {want_k}"
        assert (lens_k == 0.0) == (want_k == 0.0), name
        tol = TOL_LENS if want_k != 0.0 else TOL_FIT
        worst = np.zeros(4)
        for t, (g, w) in enumerate(zip(got, case["transforms"])):
            want = _want(w)
            for i in range(4):
                worst[i] = max(worst[i], abs(g[i] - want[i]))
            assert all(abs(g[i] - want[i]) <= tol[i] for i in range(4)), f"{name} frame {t}: {g} vs {want}"
            assert g[5] == want[5], f"{name} frame {t}: extra {g[5]} vs {want[5]}"
            assert abs(g[4] - want[4]) <= 1e-7, f"{name} frame {t}: barrel"
        _say(f"[gpu] l2t entry point {name}: lensK {lens_k:.8f} vs {want_k:.8f}, largest differences "
             f"x {worst[0]:.2e} y {worst[1]:.2e} alpha {worst[2]:.2e} zoom {worst[3]:.2e}")
        if case["simple"] or want_k == 0.0:
            # the host's simple transform, or the GPU fit without a lens: exact
            for t, (g, w) in enumerate(zip(got, case["transforms"])):
                want = _want(w)
                assert all(_bits(g[i]) == _bits(want[i]) for i in range(4)), f"{name} frame {t} bits"


def test_store_transforms_writes_global_motions(l2t, tmp_path, monkeypatch):
    """storeTransforms: one record per frame in global_motions.trf (the
    reference's text layout, a comment line after each)"""
    case = l2t["fov65"]
    monkeypatch.chdir(tmp_path)
    got, _ = api.localmotions_to_transforms(case["width"], case["height"], _frames(case), fov=case["fov"],
                                            estimate=0, smooth_zoom=1, store=True)
    lines = open(tmp_path / "global_motions.trf").read().splitlines()
    assert len(lines) == 2 * len(got)
    for t, g in enumerate(got):
        rec = lines[2 * t].split()
        assert rec[0] == "0" and int(rec[5]) == g[5]
        assert abs(float(rec[1]) - g[0]) < 1e-5 and abs(float(rec[2]) - g[1]) < 1e-5
        assert lines[2 * t + 1].startswith("#")
    assert lines[2 * 3 + 1] == "# no fields"
