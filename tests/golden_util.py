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
