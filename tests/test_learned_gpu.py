"""The learned-proposal form on a held-out 4K motion-blur clip (config 5,
clip 0; the trainer uses clips 100 and up).

The exact full search stays the reference: the learned form's LocalMotions (its
coarse and fine fields both come from window searches around the network's
displacements) are compared field by field with the plain form's (which the fixtures pin to the
reference), and the simple global transform of every frame is compared in
pixels at the frame corners.  The ablation runs the same method with an
untrained network (all-zero weights: every proposal is "no motion") and
shows what the trained network gains: far fewer exact-search fallbacks, at
the same agreement."""
import os
import sys

import numpy as np
import pytest

from vidstab_amd import api, learned
from vidstab_amd.synth import Clip

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "..", "vid.stab_amd", "learned", "proposal_c5.bin")
NFRAMES = 72   # several batches in flight at once


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def clip4k():
    import torch
    clip = Clip(5, 0)
    s = clip.spec
    host = np.stack([clip.render(t) for t in range(NFRAMES)])
    dev = torch.from_numpy(host).to("cuda:0")
    torch.cuda.synchronize()
    fb = s.width * s.height
    ptrs = [dev.data_ptr() + t * fb for t in range(NFRAMES)]
    kw = dict(shakiness=s.shakiness, accuracy=s.accuracy, stepsize=s.stepsize, mincontrast=s.mincontrast)
    det = api.Detector(s.width, s.height, **kw)
    plain = det.detect_device(ptrs, [s.width] * NFRAMES, keep=True)
    S = det.md.fieldscoarse.fieldSize
    det.close()
    yield dict(spec=s, kw=kw, ptrs=ptrs, plain=plain, S=S, dev=dev)


def _run_learned(c, path):
    s = c["spec"]
    det = api.Detector(s.width, s.height, **c["kw"])
    det.set_proposal(path)
    got = det.detect_device(c["ptrs"], [s.width] * NFRAMES, keep=True)
    stats = det.proposal_stats(0)
    fine = det.proposal_stats(1)
    det.close()
    return got, stats, fine


def test_learned_form_agrees_with_full_search(clip4k):
    assert os.path.exists(WEIGHTS), "trained weights missing"
    got, (fields, fallbacks), (ffields, ffallbacks) = _run_learned(clip4k, WEIGHTS)
    cn, ce, an, ae = api.compare_forms(clip4k["plain"], got, clip4k["S"])
    s = clip4k["spec"]
    errs = [api.transform_distance(s.width, s.height, api.simple_transform(s.width, s.height, p),
                                   api.simple_transform(s.width, s.height, q))
            for p, q in zip(clip4k["plain"][1:], got[1:])]
    _say(f"[gpu] learned: {fields} coarse fields, {fallbacks} fallbacks; {ffields} fine fields, {ffallbacks} "
         f"fallbacks; coarse equal {ce}/{cn}, all equal {ae}/{an}, mean transform error {np.mean(errs):.4f} px")
    assert fields == cn
    assert fallbacks <= 0.15 * fields
    # every fine field searched went through the window search (kept or not
    # by the cleanup afterwards), and most windows were accepted
    assert ffields >= an - cn > 0
    assert ffallbacks <= 0.25 * ffields
    assert ae >= 0.99 * an
    assert (ae - ce) >= 0.99 * (an - cn)
    assert np.mean(errs) <= 0.05


def test_ablation_untrained_network(clip4k, tmp_path):
    """all-zero weights: proposals are zero shifts, so nearly every field
    falls back; the trained network must cut the fallbacks sharply while
    agreeing with the full search at least as well"""
    layers, thr = learned.load_weights(WEIGHTS)
    zero = [(np.zeros_like(w), np.zeros_like(b), d) for (w, b, d) in layers]
    zpath = tmp_path / "zero.bin"
    learned.save_weights(str(zpath), zero, thr)
    got0, (f0, b0), _ = _run_learned(clip4k, str(zpath))
    got1, (f1, b1), _ = _run_learned(clip4k, WEIGHTS)
    c0 = api.compare_forms(clip4k["plain"], got0, clip4k["S"])
    c1 = api.compare_forms(clip4k["plain"], got1, clip4k["S"])
    _say(f"[gpu] ablation: untrained {b0}/{f0} fallbacks, equal {c0[3]}/{c0[2]}; "
         f"trained {b1}/{f1} fallbacks, equal {c1[3]}/{c1[2]}")
    assert f0 == f1 > 0
    assert b1 <= 0.5 * b0
    assert c1[3] >= c0[3] - 0.01 * c0[2]


def test_fallback_list_reproduces_plain(clip4k, tmp_path):
    """a residual threshold of ~0 rejects every window minimum above SAD 0,
    so (nearly) every coarse and fine field is listed for the exact search
    of the fallback kernel: the LocalMotions must be the plain form's, bit
    for bit, field by field"""
    layers, _ = learned.load_weights(WEIGHTS)
    path = tmp_path / "thr0.bin"
    learned.save_weights(str(path), layers, 1e-9)
    got, (f, b), (ff, fb) = _run_learned(clip4k, str(path))
    cn, ce, an, ae = api.compare_forms(clip4k["plain"], got, clip4k["S"])
    _say(f"[gpu] threshold 0: {b}/{f} coarse and {fb}/{ff} fine fields to the fallback list; equal {ae}/{an}")
    assert f == cn and b >= 0.99 * f
    assert ff >= an - cn and fb >= 0.99 * ff
    assert [len(x) for x in got] == [len(x) for x in clip4k["plain"]]
    assert ae == an


def test_fused_first_layers_do_not_change_the_result(clip4k, monkeypatch):
    """the network's first two layers fused in one pass (the default) or run
    as two launches (VIDSTAB_AMD_CONV_FUSE=0, read when the proposal is
    attached): the same proposals, so the same LocalMotions, bit for bit"""
    got_fused, st_fused, _ = _run_learned(clip4k, WEIGHTS)
    monkeypatch.setenv("VIDSTAB_AMD_CONV_FUSE", "0")
    got_split, st_split, _ = _run_learned(clip4k, WEIGHTS)
    assert st_fused == st_split
    assert got_fused == got_split
