"""Helpers shared by the tests that replay the golden fixtures."""
import struct

import numpy as np


def hexbits(h):
    return int(h, 16)


def golden_motions(frame):
    """fixture LocalMotions -> tuples (vx, vy, fx, fy, size, contrast_bits, match_bits)"""
    return [(m[0], m[1], m[2], m[3], m[4], hexbits(m[5]), hexbits(m[6])) for m in frame["lm"]]


def bits_to_double(b):
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def case_clip(case):
    """the synthetic clip of a fixture case and its luma stride"""
    from vidstab_amd.synth import Clip
    clip = Clip(case["config"], case["clip"], width=case["width"], height=case["height"],
                nframes=len(case["frames"]))
    return clip, case["stride"]


def case_params(case):
    return dict(shakiness=case["shakiness"], accuracy=case["accuracy"],
                stepsize=case["stepsize"],
                mincontrast=bits_to_double(hexbits(case["mincontrast"])),
                tripod=case["tripod"])


CLIP_HASHES = "clip_hashes.json"
FNV_BASIS = 0xCBF29CE484222325


def hash_motions(motions):
    """FNV-1a over a frame's LocalMotions as the whole-clip fixtures hash them:
    per motion little-endian int16 v.x, v.y, f.x, f.y, f.size, then the IEEE
    bits of contrast and match (26 bytes)"""
    if not motions:
        return FNV_BASIS
    from vidstab_amd.synth import hash_plane
    rec = np.zeros(len(motions), dtype=[("v", "<i2", 5), ("c", "<u8"), ("m", "<u8")])
    rec["v"] = [m[:5] for m in motions]
    rec["c"] = [m[5] for m in motions]
    rec["m"] = [m[6] for m in motions]
    return hash_plane(np.frombuffer(rec.tobytes(), dtype=np.uint8).reshape(1, -1))


def load_clip_hashes():
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", CLIP_HASHES)) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}
