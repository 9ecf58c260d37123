"""Locations of the in-tree native libraries."""
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_ROOT, "lib")


def lib_path(name):
    p = os.path.join(LIB_DIR, name)
    if not os.path.exists(p):
        raise RuntimeError(f"native library missing: {p} -- run __graft_entry__.build()")
    return p
