"""ctypes wrapper of the CPU oracle (oracle/vsoracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker of the GPU library.  The product
never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libvsref.so")


class Field(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int16), ("y", ctypes.c_int16), ("size", ctypes.c_int16)]


class Vec(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int16), ("y", ctypes.c_int16)]


class LocalMotion(ctypes.Structure):
    _fields_ = [("v", Vec), ("f", Field), ("contrast", ctypes.c_double),
                ("match", ctypes.c_double)]


class VSTransform(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double), ("alpha", ctypes.c_double),
                ("zoom", ctypes.c_double), ("barrel", ctypes.c_double),
                ("rshutter", ctypes.c_double), ("extra", ctypes.c_int)]


class OrFields(ctypes.Structure):
    _fields_ = [("fieldSize", ctypes.c_int), ("maxShift", ctypes.c_int),
                ("stepSize", ctypes.c_int), ("fieldNum", ctypes.c_int),
                ("maxFields", ctypes.c_int), ("fieldRows", ctypes.c_int),
                ("contrastThreshold", ctypes.c_double), ("fields", ctypes.POINTER(Field))]


class OrState(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int),
                ("shakiness", ctypes.c_int), ("accuracy", ctypes.c_int),
                ("stepSize", ctypes.c_int), ("virtualTripod", ctypes.c_int),
                ("contrastThreshold", ctypes.c_double),
                ("coarse", OrFields), ("fine", OrFields),
                ("blurSize", ctypes.c_int),
                ("curr", ctypes.c_void_p), ("prev", ctypes.c_void_p), ("tmp", ctypes.c_void_p),
                ("hasSeenOneFrame", ctypes.c_int), ("frameNum", ctypes.c_int),
                ("nCoarse", ctypes.c_int), ("nFineAll", ctypes.c_int), ("nFineKept", ctypes.c_int),
                ("lastT", VSTransform), ("lastMeanMatch", ctypes.c_double), ("bpp", ctypes.c_int),
                ("show", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.or_new.restype = ctypes.POINTER(OrState)
        L.or_new.argtypes = [ctypes.c_int] * 5 + [ctypes.c_double, ctypes.c_int]
        L.or_new_packed.restype = ctypes.POINTER(OrState)
        L.or_new_packed.argtypes = [ctypes.c_int] * 6 + [ctypes.c_double, ctypes.c_int]
        L.or_contrast_packed.restype = ctypes.c_double
        L.or_contrast_packed.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5
        L.or_sad_packed.restype = ctypes.c_uint32
        L.or_sad_packed.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 7
        L.or_field_search_packed.restype = ctypes.c_int
        L.or_field_search_packed.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(OrFields), ctypes.POINTER(Field), ctypes.c_int,
                                             ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint32)]
        L.or_delete.argtypes = [ctypes.POINTER(OrState)]
        L.or_detect.restype = ctypes.c_int
        L.or_detect.argtypes = [ctypes.POINTER(OrState), ctypes.c_void_p, ctypes.c_int,
                                ctypes.POINTER(LocalMotion), ctypes.c_int]
        L.or_blurred.restype = ctypes.c_void_p
        L.or_blurred.argtypes = [ctypes.POINTER(OrState)]
        L.or_boxblur.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 4
        L.or_blur_size.restype = ctypes.c_int
        L.or_blur_size.argtypes = [ctypes.c_int] * 3
        L.or_sad.restype = ctypes.c_uint32
        L.or_sad.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 6
        L.or_spiral_rank.restype = ctypes.c_int
        L.or_spiral_rank.argtypes = [ctypes.c_int, ctypes.c_int]
        L.or_spiral_walk.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int)]
        L.or_cleanmean.restype = ctypes.c_double
        L.or_cleanmean.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.or_contrast.restype = ctypes.c_double
        L.or_contrast.argtypes = [ctypes.c_int, ctypes.c_int]
        L.or_select.restype = ctypes.c_int
        L.or_select.argtypes = [ctypes.POINTER(OrFields), ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
        L.or_simple_transform.restype = VSTransform
        L.or_simple_transform.argtypes = [ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(LocalMotion), ctypes.c_int]
        L.or_transform_point.argtypes = [ctypes.POINTER(VSTransform)] + [ctypes.c_int] * 4 + \
            [ctypes.POINTER(ctypes.c_int)] * 2
        L.or_motions_to_transform.restype = VSTransform
        L.or_motions_to_transform.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(LocalMotion), ctypes.c_int,
                                              ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def motions_to_list(arr, n):
    """LocalMotions as comparable tuples (doubles as their IEEE bit patterns)."""
    out = []
    for i in range(n):
        m = arr[i]
        c = np.array([m.contrast], dtype=np.float64).view(np.uint64)[0]
        t = np.array([m.match], dtype=np.float64).view(np.uint64)[0]
        out.append((m.v.x, m.v.y, m.f.x, m.f.y, m.f.size, int(c), int(t)))
    return out


def local_motions(motions):
    """motion tuples (vx, vy, fx, fy, size, contrast_bits, match_bits) -> LocalMotion array"""
    arr = (LocalMotion * max(1, len(motions)))()
    for i, m in enumerate(motions):
        arr[i].v.x, arr[i].v.y = m[0], m[1]
        arr[i].f.x, arr[i].f.y, arr[i].f.size = m[2], m[3], m[4]
        arr[i].contrast = float(np.array([m[5]], dtype=np.uint64).view(np.float64)[0])
        arr[i].match = float(np.array([m[6]], dtype=np.uint64).view(np.float64)[0])
    return arr


def motions_to_transform(width, height, motions, smooth_zoom=0):
    """the second pass's vsMotionsToTransform of one frame (similarity model):
    (VSTransform, residual, gradient-descent rounds)"""
    L = lib()
    arr = local_motions(motions)
    res, rounds = ctypes.c_double(0.0), ctypes.c_int(0)
    t = L.or_motions_to_transform(width, height, arr, len(motions), smooth_zoom, ctypes.byref(res),
                                  ctypes.byref(rounds))
    return t, res.value, rounds.value


class Detector:
    """One vsMotionDetection stream on the CPU oracle.  bpp 1: planar luma;
    3 or 4: packed RGB24 / BGR24 / RGBA frames (plane 0, width*bpp bytes a row)."""

    def __init__(self, width, height, shakiness=5, accuracy=15, stepsize=6,
                 mincontrast=0.25, tripod=0, bpp=1, show=0):
        self._L = lib()
        self.width, self.height, self.bpp = width, height, bpp
        if bpp == 1:
            self.s = self._L.or_new(width, height, shakiness, accuracy, stepsize,
                                    float(mincontrast), tripod)
        else:
            self.s = self._L.or_new_packed(width, height, bpp, shakiness, accuracy, stepsize,
                                           float(mincontrast), tripod)
        if not self.s:
            raise RuntimeError("or_new failed")
        st = self.s.contents
        st.show = show
        self.maxout = st.coarse.fieldNum + st.fine.fieldNum + 1
        self.buf = (LocalMotion * self.maxout)()

    def detect(self, luma):
        """luma (or packed plane 0): uint8 array (H, stride) -> list of motion tuples.
        With show > 0 the overlays are drawn into the array (it must be
        C-contiguous uint8, as the reference draws into the caller's frame)."""
        if self.s.contents.show:
            assert luma.dtype == np.uint8 and luma.flags.c_contiguous
        luma = np.ascontiguousarray(luma, dtype=np.uint8)
        n = self._L.or_detect(self.s, luma.ctypes.data, luma.shape[1], self.buf, self.maxout)
        return motions_to_list(self.buf, n)

    def blurred(self):
        """the working plane (packed frames: the unblurred copy, width*bpp bytes a row)"""
        p = self._L.or_blurred(self.s)
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint8)),
                                     shape=(self.height, self.width * self.bpp)).copy()

    @property
    def state(self):
        return self.s.contents

    def close(self):
        if self.s:
            self._L.or_delete(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RefDetector:
    """The reference's own vsMotionDetection (oracle/_ref/libvsref.so: its C
    sources built with its release flags, SIMD dispatch and OpenMP), driven
    through the same C API.  Used as the CPU baseline and as a cross-check."""

    def __init__(self, width, height, shakiness=5, accuracy=15, stepsize=6, mincontrast=0.25,
                 tripod=0, threads=0, pformat=None):
        from vidstab_amd import api  # struct layouts only (identical ABI)
        if not os.path.exists(REF_PATH):
            raise RuntimeError(f"reference build missing: {REF_PATH}")
        self.api = api
        L = ctypes.CDLL(REF_PATH)
        L.vsFrameInfoInit.argtypes = [ctypes.POINTER(api.VSFrameInfo)] + [ctypes.c_int] * 3
        L.vsMotionDetectGetDefaultConfig.restype = api.VSMotionDetectConfig
        L.vsMotionDetectGetDefaultConfig.argtypes = [ctypes.c_char_p]
        L.vsMotionDetectInit.argtypes = [ctypes.POINTER(api.VSMotionDetect),
                                         ctypes.POINTER(api.VSMotionDetectConfig),
                                         ctypes.POINTER(api.VSFrameInfo)]
        L.vsMotionDetection.argtypes = [ctypes.POINTER(api.VSMotionDetect),
                                        ctypes.POINTER(api.VSVector), ctypes.POINTER(api.VSFrame)]
        L.vsMotionDetectionCleanup.argtypes = [ctypes.POINTER(api.VSMotionDetect)]
        L.vs_vector_del.argtypes = [ctypes.POINTER(api.VSVector)]
        ctypes.c_int.in_dll(L, "vs_log_level").value = 0
        self.L = L
        self.fi = api.VSFrameInfo()
        self.packed = pformat is not None and pformat > api.PF_PACKED
        L.vsFrameInfoInit(ctypes.byref(self.fi), width, height, pformat if self.packed else api.PF_YUV420P)
        self.md = api.VSMotionDetect()
        conf = L.vsMotionDetectGetDefaultConfig(b"ref")
        conf.shakiness, conf.accuracy, conf.stepSize = shakiness, accuracy, stepsize
        conf.contrastThreshold = float(mincontrast)
        conf.virtualTripod = tripod
        conf.numThreads = threads
        if L.vsMotionDetectInit(ctypes.byref(self.md), ctypes.byref(conf), ctypes.byref(self.fi)) != 0:
            raise RuntimeError("reference vsMotionDetectInit failed")
        self.width, self.height = width, height
        self._chroma = np.full((height // 2) * (width // 2), 128, dtype=np.uint8)

    def detect(self, luma, keep=True):
        api = self.api
        fr = api.VSFrame()
        fr.data[0] = luma.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        fr.linesize[0] = luma.shape[1]
        if not self.packed:
            fr.data[1] = fr.data[2] = self._chroma.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
            fr.linesize[1] = fr.linesize[2] = self.width // 2
        vec = api.VSVector()
        self.L.vsMotionDetection(ctypes.byref(self.md), ctypes.byref(vec), ctypes.byref(fr))
        out = api.motions_of(vec) if keep else vec.nelems
        self.L.vs_vector_del(ctypes.byref(vec))
        return out

    def close(self):
        if self.md.initialized:
            self.L.vsMotionDetectionCleanup(ctypes.byref(self.md))


def ref_localmotions_to_transforms(width, height, frames, fov=0.0, estimate=1):
    """The reference's own vsLocalmotions2Transforms (oracle/_ref/libvsref.so)
    on a clip's LocalMotions with the reference's default transform
    configuration apart from fov and estimateLensDistortion -> (list of (x, y,
    alpha, zoom, barrel, extra), the lens strength left in the transform data).
    The CPU baseline of the second pass."""
    from vidstab_amd import api  # struct layouts only (identical ABI)
    if not os.path.exists(REF_PATH):
        raise RuntimeError(f"reference build missing: {REF_PATH}")
    L = ctypes.CDLL(REF_PATH)
    ctypes.c_int.in_dll(L, "vs_log_level").value = 0
    L.vsFrameInfoInit.argtypes = [ctypes.POINTER(api.VSFrameInfo)] + [ctypes.c_int] * 3
    L.vsTransformGetDefaultConfig.restype = api.VSTransformConfig
    L.vsTransformGetDefaultConfig.argtypes = [ctypes.c_char_p]
    L.vsLocalmotions2Transforms.argtypes = [ctypes.POINTER(api.VSTransformData), ctypes.POINTER(api.VSVector),
                                            ctypes.POINTER(api.VSTransformations)]
    fi = api.VSFrameInfo()
    L.vsFrameInfoInit(ctypes.byref(fi), width, height, api.PF_YUV420P)
    clip = frames if isinstance(frames, api.ClipMotions) else api.ClipMotions(frames)
    many = api._many(clip)
    td = api.VSTransformData()
    td.fiSrc = fi
    td.fiDest = fi
    td.conf = L.vsTransformGetDefaultConfig(b"ref")
    td.conf.fov = fov
    td.conf.estimateLensDistortion = estimate
    trans = api.VSTransformations()
    if L.vsLocalmotions2Transforms(ctypes.byref(td), ctypes.byref(many), ctypes.byref(trans)) != 0:
        raise RuntimeError("reference vsLocalmotions2Transforms failed")
    out = [(t.x, t.y, t.alpha, t.zoom, t.barrel, t.extra) for t in trans.ts[:trans.len]]
    free = ctypes.CFUNCTYPE(None, ctypes.c_void_p).in_dll(L, "vs_free")
    free(ctypes.cast(trans.ts, ctypes.c_void_p))
    return out, td.lensK
