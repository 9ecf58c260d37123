"""Seeded synthetic clips of the detect workloads (wraps csrc/synth/vsynth.c).

The generator is integer / IEEE-basic-op only, so a (config, clip) pair gives
byte-identical frames on every x86-64 host.
"""
import ctypes
import os

import numpy as np

from ._paths import lib_path


class VSynMotion(ctypes.Structure):
    _fields_ = [("tx", ctypes.c_double), ("ty", ctypes.c_double), ("alpha", ctypes.c_double)]


class VSynRect(ctypes.Structure):
    _fields_ = [("w", ctypes.c_int), ("h", ctypes.c_int), ("x0", ctypes.c_double),
                ("y0", ctypes.c_double), ("vx", ctypes.c_double), ("vy", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("tex", ctypes.c_void_p)]


class VSynClip(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("margin", ctypes.c_int),
                ("seed", ctypes.c_uint64), ("noise_sigma_q8", ctypes.c_int),
                ("nrects", ctypes.c_int), ("rects", VSynRect * 8),
                ("blur_len_max", ctypes.c_int), ("master", ctypes.c_void_p)]


class VSynSpec(ctypes.Structure):
    _fields_ = [("config", ctypes.c_int), ("clip", ctypes.c_int), ("width", ctypes.c_int),
                ("height", ctypes.c_int), ("nframes", ctypes.c_int), ("shakiness", ctypes.c_int),
                ("accuracy", ctypes.c_int), ("stepsize", ctypes.c_int),
                ("mincontrast", ctypes.c_double), ("seed", ctypes.c_uint64)]


_lib = None


def _L():
    global _lib
    if _lib is None:
        p = lib_path("libvsynth.so")
        L = ctypes.CDLL(p)
        L.vsyn_spec.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(VSynSpec)]
        L.vsyn_open.argtypes = [ctypes.POINTER(VSynSpec), ctypes.POINTER(VSynClip),
                                ctypes.POINTER(VSynMotion)]
        L.vsyn_clip_free.argtypes = [ctypes.POINTER(VSynClip)]
        L.vsyn_render.argtypes = [ctypes.POINTER(VSynClip), ctypes.POINTER(VSynMotion),
                                  ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        L.vsyn_hash_plane.restype = ctypes.c_uint64
        L.vsyn_hash_plane.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.vsyn_value_noise.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_uint64]
        L.vsyn_sin.restype = ctypes.c_double
        L.vsyn_sin.argtypes = [ctypes.c_double]
        L.vsyn_cos.restype = ctypes.c_double
        L.vsyn_cos.argtypes = [ctypes.c_double]
        _lib = L
    return _lib


def spec(config, clip=0):
    s = VSynSpec()
    if _L().vsyn_spec(config, clip, ctypes.byref(s)) != 0:
        raise ValueError(f"unknown workload config {config}")
    return s


class Clip:
    """Frames of one synthetic clip.  `overrides` may change width/height/nframes
    (the motion path and texture keep their seed)."""

    def __init__(self, config, clip=0, width=None, height=None, nframes=None):
        self.spec = spec(config, clip)
        if width:
            self.spec.width, self.spec.height = width, height
        if nframes and nframes > self.spec.nframes:
            self.spec.nframes = nframes
        self.width, self.height = self.spec.width, self.spec.height
        self.nframes = self.spec.nframes
        self._clip = VSynClip()
        self._path = (VSynMotion * self.nframes)()
        if _L().vsyn_open(ctypes.byref(self.spec), ctypes.byref(self._clip), self._path) != 0:
            raise RuntimeError("vsyn_open failed")

    def render(self, t, stride=None, out=None):
        stride = stride or self.width
        if out is None:
            out = np.full((self.height, stride), 0x5A, dtype=np.uint8)
        _L().vsyn_render(ctypes.byref(self._clip), self._path, t, out.ctypes.data, stride)
        return out

    def motion(self, t):
        m = self._path[t]
        return (m.tx, m.ty, m.alpha)

    def close(self):
        if self._clip.master:
            _L().vsyn_clip_free(ctypes.byref(self._clip))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pack(luma, bpp, stride=None, width=None):
    """packed frame (H, stride) of bpp bytes a pixel from a luma plane: bytes
    Y, (7*Y) mod 256, 255 - Y and (bpp 4) Y xor 0x5A, as vsyn_pack writes"""
    w = width or luma.shape[1]
    y = np.asarray(luma, dtype=np.uint8)[:, :w]
    h = y.shape[0]
    stride = stride or w * bpp
    out = np.full((h, stride), 0x5A, dtype=np.uint8)
    px = out[:, :w * bpp].reshape(h, w, bpp)
    px[..., 0] = y
    px[..., 1] = (y.astype(np.uint32) * 7).astype(np.uint8)
    px[..., 2] = 255 - y
    if bpp > 3:
        px[..., 3] = y ^ 0x5A
    return out


def hash_plane(a, width=None):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    w = width or a.shape[1]
    return int(_L().vsyn_hash_plane(a.ctypes.data, w, a.shape[0], a.shape[1]))
