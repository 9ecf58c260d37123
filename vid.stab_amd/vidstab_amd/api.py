"""ctypes binding of lib/libvidstab_amd.so: the vid.stab detect C API.

Struct layouts mirror include/*.h (which mirror the reference's ABI), so the
Python side exercises the same entry points a C caller such as ffmpeg's
vidstabdetect filter uses:  vsFrameInfoInit -> vsMotionDetectInit ->
{vsMotionDetection; vsWriteToFile; vs_vector_del} -> vsMotionDetectionCleanup.
"""
import ctypes
import os

import numpy as np

from ._paths import lib_path

c_int, c_short, c_double, c_void_p = ctypes.c_int, ctypes.c_short, ctypes.c_double, ctypes.c_void_p

PF_GRAY8, PF_YUV420P, PF_YUV422P, PF_YUV444P = 0, 1, 2, 3
PF_PACKED, PF_RGB24, PF_BGR24, PF_RGBA = 8, 9, 10, 11
ASCII_SERIALIZATION_MODE, BINARY_SERIALIZATION_MODE = 1, 2


class VSFrameInfo(ctypes.Structure):
    _fields_ = [("width", c_int), ("height", c_int), ("planes", c_int),
                ("log2ChromaW", c_int), ("log2ChromaH", c_int), ("pFormat", c_int),
                ("bytesPerPixel", c_int)]


class VSFrame(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(ctypes.c_uint8) * 4), ("linesize", c_int * 4)]


class VSMotionDetectConfig(ctypes.Structure):
    _fields_ = [("shakiness", c_int), ("accuracy", c_int), ("stepSize", c_int), ("algo", c_int),
                ("virtualTripod", c_int), ("show", c_int), ("contrastThreshold", c_double),
                ("modName", ctypes.c_char_p), ("numThreads", c_int)]


class Field(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int16), ("y", ctypes.c_int16), ("size", ctypes.c_int16)]


class Vec(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int16), ("y", ctypes.c_int16)]


class LocalMotion(ctypes.Structure):
    _fields_ = [("v", Vec), ("f", Field), ("contrast", c_double), ("match", c_double)]


class VSTransform(ctypes.Structure):
    _fields_ = [("x", c_double), ("y", c_double), ("alpha", c_double), ("zoom", c_double),
                ("barrel", c_double), ("rshutter", c_double), ("extra", c_int)]


class VSMotionDetectFields(ctypes.Structure):
    _fields_ = [("maxShift", c_int), ("stepSize", c_int), ("fieldNum", c_int),
                ("maxFields", c_int), ("contrastThreshold", c_double), ("fieldSize", c_int),
                ("fieldRows", c_int), ("fields", ctypes.POINTER(Field)), ("useOffset", c_short),
                ("offset", VSTransform)]


class VSMotionDetect(ctypes.Structure):
    _fields_ = [("fi", VSFrameInfo), ("conf", VSMotionDetectConfig),
                ("fieldscoarse", VSMotionDetectFields), ("fieldsfine", VSMotionDetectFields),
                ("curr", VSFrame), ("currorig", VSFrame), ("currtmp", VSFrame), ("prev", VSFrame),
                ("hasSeenOneFrame", c_short), ("initialized", c_int),
                ("serializationMode", c_int), ("frameNum", c_int)]


class VSVector(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(c_void_p)), ("buffersize", c_int), ("nelems", c_int)]


class VSTransformations(ctypes.Structure):
    _fields_ = [("ts", ctypes.POINTER(VSTransform)), ("current", c_int), ("len", c_int),
                ("warned_end", c_short)]


class VSTransformConfig(ctypes.Structure):
    _fields_ = [("relative", c_int), ("smoothing", c_int), ("crop", c_int), ("invert", c_int),
                ("zoom", c_double), ("optZoom", c_int), ("zoomSpeed", c_double), ("interpolType", c_int),
                ("maxShift", c_int), ("maxAngle", c_double), ("modName", ctypes.c_char_p), ("verbose", c_int),
                ("simpleMotionCalculation", c_int), ("storeTransforms", c_int), ("smoothZoom", c_int),
                ("camPathAlgo", c_int), ("estimateLensDistortion", c_int), ("lensCorrection", c_int),
                ("lensK", c_double), ("fov", c_double)]


class VSLensPlaneMap(ctypes.Structure):
    _fields_ = [("active", c_int), ("mode", c_int), ("k", c_double), ("cdx", c_double), ("cdy", c_double),
                ("csx", c_double), ("csy", c_double), ("tMaxU", c_double), ("tMaxD", c_double),
                ("tDomD", c_double), ("invRho2", c_double), ("f", c_double), ("sxShift", c_int),
                ("syShift", c_int), ("idxScaleU", ctypes.c_int32), ("idxScaleD", ctypes.c_int32),
                ("gU", c_void_p), ("gD", c_void_p)]


class VSTransformData(ctypes.Structure):
    _fields_ = [("fiSrc", VSFrameInfo), ("fiDest", VSFrameInfo), ("src", VSFrame), ("destbuf", VSFrame),
                ("dest", VSFrame), ("srcMalloced", c_short), ("interpolate", c_void_p),
                ("conf", VSTransformConfig), ("lensMode", c_int), ("lensActive", c_int), ("lensK", c_double),
                ("lensMapK", c_double), ("lensMaps", VSLensPlaneMap * 3), ("initialized", c_int)]


class VSLensEstimateConfig(ctypes.Structure):
    _fields_ = [("kMin", c_double), ("kMax", c_double), ("tolerance", c_double), ("maxIterations", c_int),
                ("gaussNewtonSteps", c_int), ("maxUncertainty", c_double), ("rejectOutliers", c_int),
                ("outlierStddevs", c_double), ("outlierPasses", c_int), ("f", c_double)]


class VSLensEstimate(ctypes.Structure):
    _fields_ = [("k", c_double), ("residual", c_double), ("curvature", c_double), ("uncertainty", c_double),
                ("iterations", c_int), ("rejected", c_int), ("used", c_int), ("determined", c_int)]


class VSGStageDesc(ctypes.Structure):
    _fields_ = [("fieldSize", c_int), ("maxShift", c_int), ("stepSize", c_int),
                ("nfields", c_int), ("fieldRows", c_int), ("maxFields", c_int),
                ("contrastThreshold", ctypes.c_double), ("fields", ctypes.POINTER(Field))]


class VSGJob(ctypes.Structure):
    _fields_ = [("slot", ctypes.c_int32), ("prev", ctypes.c_int32), ("field", ctypes.c_int32),
                ("ox", ctypes.c_int16), ("oy", ctypes.c_int16)]


class VSGResult(ctypes.Structure):
    _fields_ = [("tx", ctypes.c_int16), ("ty", ctypes.c_int16), ("minerror", ctypes.c_uint32)]


_lib = None


def lib():
    """Loads the native library; raises if it is not built (no fallback path)."""
    global _lib
    if _lib is None:
        L = ctypes.CDLL(lib_path("libvidstab_amd.so"), mode=ctypes.RTLD_GLOBAL)
        L.vsFrameInfoInit.argtypes = [ctypes.POINTER(VSFrameInfo), c_int, c_int, c_int]
        L.vsFrameInfoInit.restype = c_int
        L.vsMotionDetectGetDefaultConfig.argtypes = [ctypes.c_char_p]
        L.vsMotionDetectGetDefaultConfig.restype = VSMotionDetectConfig
        L.vsMotionDetectInit.argtypes = [ctypes.POINTER(VSMotionDetect),
                                         ctypes.POINTER(VSMotionDetectConfig),
                                         ctypes.POINTER(VSFrameInfo)]
        L.vsMotionDetectInit.restype = c_int
        L.vsMotionDetection.argtypes = [ctypes.POINTER(VSMotionDetect), ctypes.POINTER(VSVector),
                                        ctypes.POINTER(VSFrame)]
        L.vsMotionDetection.restype = c_int
        L.vsMotionDetectionBatch.argtypes = [ctypes.POINTER(VSMotionDetect), ctypes.POINTER(VSVector),
                                             ctypes.POINTER(VSFrame), c_int]
        L.vsMotionDetectionBatch.restype = c_int
        L.vsMotionDetectionDevice.argtypes = [ctypes.POINTER(VSMotionDetect), ctypes.POINTER(VSVector),
                                              ctypes.POINTER(c_void_p), ctypes.POINTER(c_int), c_int]
        L.vsMotionDetectionDevice.restype = c_int
        L.vsMotionDetectGetBlurred.argtypes = [ctypes.POINTER(VSMotionDetect), c_void_p, c_int]
        L.vsMotionDetectGetBlurred.restype = c_int
        L.vsMotionDetectGetBlurredAt.argtypes = [ctypes.POINTER(VSMotionDetect), c_int, c_void_p, c_int]
        L.vsMotionDetectGetBlurredAt.restype = c_int
        L.vsMotionDetectionCleanup.argtypes = [ctypes.POINTER(VSMotionDetect)]
        L.vsMotionDetectionCleanup.restype = None
        L.vs_vector_del.argtypes = [ctypes.POINTER(VSVector)]
        L.vsLocalMotionsArrayDel.argtypes = [ctypes.POINTER(VSVector), c_int]
        L.vsMotionDetectMulti.argtypes = [ctypes.POINTER(VSMotionDetectConfig), ctypes.POINTER(VSFrameInfo),
                                          c_int, ctypes.POINTER(c_int), c_void_p, ctypes.POINTER(c_int),
                                          c_void_p]
        L.vsMotionDetectMulti.restype = c_int
        L.vsLocalMotionsArrayDel.restype = None
        L.vsLocalMotionsReleaseWait.argtypes = []
        L.vsLocalMotionsReleaseWait.restype = None
        L.vsSimpleMotionsToTransform.argtypes = [VSFrameInfo, ctypes.c_char_p, ctypes.POINTER(VSVector)]
        L.vsSimpleMotionsToTransform.restype = VSTransform
        L.vsMotionsToTransformBatch.argtypes = [ctypes.POINTER(VSFrameInfo), ctypes.POINTER(VSVector), c_int, c_int,
                                                ctypes.POINTER(VSTransform)]
        L.vsMotionDetectSetProposal.argtypes = [ctypes.POINTER(VSMotionDetect), ctypes.c_char_p]
        L.vsMotionDetectProposalStats.argtypes = [ctypes.POINTER(VSMotionDetect), ctypes.POINTER(ctypes.c_longlong),
                                                  ctypes.POINTER(ctypes.c_longlong)]
        L.vsMotionDetectProposalStageStats.argtypes = [ctypes.POINTER(VSMotionDetect), c_int,
                                                       ctypes.POINTER(ctypes.c_longlong),
                                                       ctypes.POINTER(ctypes.c_longlong)]
        L.vs_vector_size.argtypes = [ctypes.POINTER(VSVector)]
        L.vs_vector_get.argtypes = [ctypes.POINTER(VSVector), c_int]
        L.vs_vector_get.restype = c_void_p
        L.vsg_spiral_rank.argtypes = [c_int, c_int]
        L.vsg_spiral_rank.restype = c_int
        L.vsg_spiral_unrank.argtypes = [c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]
        L.vsg_last_error.restype = ctypes.c_char_p
        L.vsPrepareFile.argtypes = [ctypes.POINTER(VSMotionDetect), c_void_p]
        L.vsWriteToFile.argtypes = [ctypes.POINTER(VSMotionDetect), c_void_p, ctypes.POINTER(VSVector)]
        L.vsLocalmotions2Transforms.argtypes = [ctypes.POINTER(VSTransformData), ctypes.POINTER(VSVector),
                                                ctypes.POINTER(VSTransformations)]
        L.vsLocalmotions2Transforms.restype = c_int
        L.vsLensEstimateGetDefaultConfig.argtypes = []
        L.vsLensEstimateGetDefaultConfig.restype = VSLensEstimateConfig
        L.vsEstimateLensDistortion.argtypes = [ctypes.POINTER(VSFrameInfo), ctypes.POINTER(VSVector),
                                               ctypes.POINTER(VSLensEstimateConfig)]
        L.vsEstimateLensDistortion.restype = VSLensEstimate
        L.focal_from_fov.argtypes = [c_double, c_int]
        L.focal_from_fov.restype = c_double
        # the library's release hook (a function pointer variable, as in the reference)
        L.vs_free_fn = ctypes.CFUNCTYPE(None, c_void_p).in_dll(L, "vs_free")
        _lib = L
    return _lib


def _bits(x):
    return int(np.array([x], dtype=np.float64).view(np.uint64)[0])


def motions_of(vec):
    """LocalMotions -> tuples (vx, vy, fx, fy, size, contrast_bits, match_bits)"""
    L = lib()
    out = []
    for i in range(vec.nelems):
        m = ctypes.cast(vec.data[i], ctypes.POINTER(LocalMotion)).contents
        out.append((m.v.x, m.v.y, m.f.x, m.f.y, m.f.size, _bits(m.contrast), _bits(m.match)))
    return out


class Detector:
    """One vsMotionDetect stream on the GPU, driven through the C API."""

    def __init__(self, width, height, shakiness=5, accuracy=15, stepsize=6, mincontrast=0.25,
                 tripod=0, pformat=PF_YUV420P, mod=b"vidstab_amd", show=0):
        L = lib()
        self.L = L
        self.fi = VSFrameInfo()
        if not L.vsFrameInfoInit(ctypes.byref(self.fi), width, height, pformat):
            raise ValueError("vsFrameInfoInit failed")
        self.md = VSMotionDetect()
        self._mod = mod
        conf = L.vsMotionDetectGetDefaultConfig(mod)
        conf.shakiness, conf.accuracy, conf.stepSize = shakiness, accuracy, stepsize
        conf.contrastThreshold = float(mincontrast)
        conf.virtualTripod = tripod
        conf.show = show
        if L.vsMotionDetectInit(ctypes.byref(self.md), ctypes.byref(conf), ctypes.byref(self.fi)) != 0:
            raise RuntimeError("vsMotionDetectInit failed: " + L.vsg_last_error().decode())
        self.width, self.height = width, height
        self.show = show
        self.packed = pformat > PF_PACKED
        self.bpp = self.fi.bytesPerPixel if self.packed else 1
        self._chroma = np.full((height // 2) * (width // 2), 128, dtype=np.uint8)

    def _frame(self, luma):
        """luma: the luma plane, or plane 0 of a packed frame (width*bpp bytes a row)"""
        fr = VSFrame()
        fr.data[0] = luma.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        fr.linesize[0] = luma.shape[1]
        if not self.packed:
            fr.data[1] = self._chroma.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
            fr.data[2] = fr.data[1]
            fr.linesize[1] = fr.linesize[2] = self.width // 2
        return fr

    def detect(self, luma):
        """one vsMotionDetection call; with show > 0 the overlays are drawn into
        luma, which must then be a C-contiguous uint8 array"""
        if self.show:
            assert luma.dtype == np.uint8 and luma.flags.c_contiguous
        luma = np.ascontiguousarray(luma, dtype=np.uint8)
        fr = self._frame(luma)
        vec = VSVector()
        if self.L.vsMotionDetection(ctypes.byref(self.md), ctypes.byref(vec), ctypes.byref(fr)) != 0:
            raise RuntimeError("vsMotionDetection failed: " + self.L.vsg_last_error().decode())
        out = motions_of(vec)
        self.L.vs_vector_del(ctypes.byref(vec))
        return out

    def detect_batch(self, lumas, keep=True):
        if self.show:
            assert all(x.dtype == np.uint8 and x.flags.c_contiguous for x in lumas)
        lumas = [np.ascontiguousarray(x, dtype=np.uint8) for x in lumas]
        frames = (VSFrame * len(lumas))(*[self._frame(x) for x in lumas])
        vecs = (VSVector * len(lumas))()
        if self.L.vsMotionDetectionBatch(ctypes.byref(self.md), vecs, frames, len(lumas)) != 0:
            raise RuntimeError("vsMotionDetectionBatch failed: " + self.L.vsg_last_error().decode())
        out = [motions_of(v) for v in vecs] if keep else [v.nelems for v in vecs]
        self.L.vsLocalMotionsArrayDel(vecs, len(lumas))
        return out

    def detect_device(self, ptrs, linesizes, keep=False):
        """frames already in device memory (integer device addresses)"""
        n = len(ptrs)
        p = (c_void_p * n)(*ptrs)
        ls = (c_int * n)(*linesizes)
        vecs = (VSVector * n)()
        if self.L.vsMotionDetectionDevice(ctypes.byref(self.md), vecs, p, ls, n) != 0:
            raise RuntimeError("vsMotionDetectionDevice failed: " + self.L.vsg_last_error().decode())
        out = [motions_of(v) for v in vecs] if keep else [v.nelems for v in vecs]
        self.L.vsLocalMotionsArrayDel(vecs, n)
        return out

    def set_proposal(self, path):
        """switch to the learned-proposal form (vsMotionDetectSetProposal)"""
        if self.L.vsMotionDetectSetProposal(ctypes.byref(self.md), str(path).encode()) != 0:
            raise RuntimeError("vsMotionDetectSetProposal failed: " + self.L.vsg_last_error().decode())

    def proposal_stats(self, stage=0):
        """(fields decided by the learned form, fallbacks) of stage 0 (coarse) or 1 (fine)"""
        f, b = ctypes.c_longlong(0), ctypes.c_longlong(0)
        if self.L.vsMotionDetectProposalStageStats(ctypes.byref(self.md), stage, ctypes.byref(f),
                                                   ctypes.byref(b)) != 0:
            raise RuntimeError("no proposal network")
        return f.value, b.value

    def blurred(self, back=0):
        """the working plane of the last frame (or of the frame `back` frames
        before it, while its device batch is kept): blurred luma, or a packed
        frame's unblurred copy"""
        wb = self.width * self.bpp
        out = np.zeros((self.height, wb), dtype=np.uint8)
        if self.L.vsMotionDetectGetBlurredAt(ctypes.byref(self.md), back, out.ctypes.data, wb) != 0:
            raise RuntimeError("vsMotionDetectGetBlurredAt failed")
        return out

    def close(self):
        if self.md.initialized:
            self.L.vsMotionDetectionCleanup(ctypes.byref(self.md))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def detect_multi(clips, keep=True, mod=b"vidstab_amd"):
    """The batched multi-clip form (vsMotionDetectMulti).

    clips: sequence of dicts with width, height, shakiness, accuracy,
    stepsize, mincontrast, ptrs (device addresses of the luma planes, one per
    frame), linesize and optionally tripod (virtualTripod, default 0) and
    show (overlays drawn into the device frames, default 0).
    Returns per clip the per-frame LocalMotions
    (tuples, as motions_of) or, with keep=False, the per-frame counts."""
    L = lib()
    n = len(clips)
    confs = (VSMotionDetectConfig * n)()
    fis = (VSFrameInfo * n)()
    nframes = (c_int * n)()
    ls = (c_int * n)()
    rows = (c_void_p * n)()
    outs = (c_void_p * n)()
    keepalive = []
    for c, cl in enumerate(clips):
        if not L.vsFrameInfoInit(ctypes.byref(fis[c]), cl["width"], cl["height"], PF_YUV420P):
            raise ValueError("vsFrameInfoInit failed")
        conf = L.vsMotionDetectGetDefaultConfig(mod)
        conf.shakiness, conf.accuracy, conf.stepSize = cl["shakiness"], cl["accuracy"], cl["stepsize"]
        conf.contrastThreshold = float(cl["mincontrast"])
        conf.virtualTripod = int(cl.get("tripod", 0))
        conf.show = int(cl.get("show", 0))
        confs[c] = conf
        nf = len(cl["ptrs"])
        nframes[c] = nf
        ls[c] = cl["linesize"]
        p = (c_void_p * max(nf, 1))(*cl["ptrs"])
        v = (VSVector * max(nf, 1))()
        keepalive += [p, v]
        rows[c] = ctypes.cast(p, c_void_p)
        outs[c] = ctypes.cast(v, c_void_p)
    if L.vsMotionDetectMulti(confs, fis, n, nframes, rows, ls, outs) != 0:
        raise RuntimeError("vsMotionDetectMulti failed: " + L.vsg_last_error().decode())
    res = []
    for c in range(n):
        v = keepalive[2 * c + 1]
        nf = nframes[c]
        res.append([motions_of(v[t]) for t in range(nf)] if keep else [v[t].nelems for t in range(nf)])
        L.vsLocalMotionsArrayDel(v, nf)
    return res


def simple_transform(width, height, motions):
    """vsSimpleMotionsToTransform of a frame's LocalMotions (tuples as
    motions_of returns them) -> (x, y, alpha)"""
    L = lib()
    fi = VSFrameInfo()
    if not L.vsFrameInfoInit(ctypes.byref(fi), width, height, PF_YUV420P):
        raise ValueError("vsFrameInfoInit failed")
    n = len(motions)
    lms = (LocalMotion * max(n, 1))()
    ptrs = (c_void_p * max(n, 1))()
    for i, (vx, vy, fx, fy, size, cbits, mbits) in enumerate(motions):
        lm = lms[i]
        lm.v.x, lm.v.y = vx, vy
        lm.f.x, lm.f.y, lm.f.size = fx, fy, size
        lm.contrast = float(np.array([cbits], dtype=np.uint64).view(np.float64)[0])
        lm.match = float(np.array([mbits], dtype=np.uint64).view(np.float64)[0])
        ptrs[i] = ctypes.cast(ctypes.byref(lm), c_void_p)
    vec = VSVector()
    vec.data = ctypes.cast(ptrs, ctypes.POINTER(c_void_p))
    vec.buffersize = max(n, 1)
    vec.nelems = n
    t = L.vsSimpleMotionsToTransform(fi, b"vidstab_amd", ctypes.byref(vec))
    return t.x, t.y, t.alpha


def _vector_of(motions, keep):
    """a VSVector over LocalMotion structs built from motion tuples; the
    backing arrays go into `keep`"""
    n = len(motions)
    lms = (LocalMotion * max(n, 1))()
    ptrs = (c_void_p * max(n, 1))()
    for i, (vx, vy, fx, fy, size, cbits, mbits) in enumerate(motions):
        lm = lms[i]
        lm.v.x, lm.v.y = vx, vy
        lm.f.x, lm.f.y, lm.f.size = fx, fy, size
        lm.contrast = float(np.array([cbits], dtype=np.uint64).view(np.float64)[0])
        lm.match = float(np.array([mbits], dtype=np.uint64).view(np.float64)[0])
        ptrs[i] = ctypes.cast(ctypes.byref(lm), c_void_p)
    keep.append((lms, ptrs))
    vec = VSVector()
    vec.data = ctypes.cast(ptrs, ctypes.POINTER(c_void_p))
    vec.buffersize = max(n, 1)
    vec.nelems = n
    return vec


class ClipMotions:
    """a clip's LocalMotions as the C API's LocalMotions vectors (one per
    frame), built once from motion tuples"""

    def __init__(self, frames):
        self.n = len(frames)
        self._keep = []
        self.vecs = (VSVector * max(self.n, 1))()
        for f, motions in enumerate(frames):
            self.vecs[f] = _vector_of(motions, self._keep)


def motions_to_transforms(width, height, frames, smooth_zoom=0):
    """the second pass's per-frame global transforms of a clip on the GPU
    (vsMotionsToTransformBatch): frames = one list of motion tuples per frame,
    or a ClipMotions -> list of (x, y, alpha, zoom, extra)"""
    L = lib()
    fi = VSFrameInfo()
    if not L.vsFrameInfoInit(ctypes.byref(fi), width, height, PF_YUV420P):
        raise ValueError("vsFrameInfoInit failed")
    clip = frames if isinstance(frames, ClipMotions) else ClipMotions(frames)
    out = (VSTransform * max(clip.n, 1))()
    if L.vsMotionsToTransformBatch(ctypes.byref(fi), clip.vecs, clip.n, smooth_zoom, out) != 0:
        raise RuntimeError("vsMotionsToTransformBatch failed: " + L.vsg_last_error().decode())
    return [(t.x, t.y, t.alpha, t.zoom, t.extra) for t in out[:clip.n]]


def _many(clip):
    """a VSManyLocalMotions (VSVector of LocalMotions pointers) over a ClipMotions"""
    ptrs = (c_void_p * max(clip.n, 1))()
    for f in range(clip.n):
        ptrs[f] = ctypes.cast(ctypes.byref(clip.vecs[f]), c_void_p)
    many = VSVector()
    many.data = ctypes.cast(ptrs, ctypes.POINTER(c_void_p))
    many.buffersize = max(clip.n, 1)
    many.nelems = clip.n
    clip._keep.append(ptrs)
    return many


def localmotions_to_transforms(width, height, frames, fov=0.0, estimate=1, simple=0, smooth_zoom=0,
                               lens_k=0.0, store=False, mod=b"vidstab_amd"):
    """vsLocalmotions2Transforms on a clip's LocalMotions (motion tuples per
    frame, or a ClipMotions) with a transform config holding the given fov,
    estimateLensDistortion, simpleMotionCalculation, smoothZoom, lensK and
    storeTransforms -> (list of (x, y, alpha, zoom, barrel, extra), the lens
    strength the call left in the transform data)"""
    L = lib()
    fi = VSFrameInfo()
    if not L.vsFrameInfoInit(ctypes.byref(fi), width, height, PF_YUV420P):
        raise ValueError("vsFrameInfoInit failed")
    clip = frames if isinstance(frames, ClipMotions) else ClipMotions(frames)
    many = _many(clip)
    td = VSTransformData()
    td.fiSrc = fi
    td.fiDest = fi
    td.conf.modName = mod
    td.conf.relative = 1
    td.conf.smoothing = 15
    td.conf.optZoom = 1
    td.conf.zoomSpeed = 0.25
    td.conf.interpolType = 2
    td.conf.maxShift = -1
    td.conf.maxAngle = -1.0
    td.conf.simpleMotionCalculation = simple
    td.conf.storeTransforms = 1 if store else 0
    td.conf.smoothZoom = smooth_zoom
    td.conf.estimateLensDistortion = estimate
    td.conf.lensCorrection = 1
    td.conf.lensK = lens_k
    td.conf.fov = fov
    trans = VSTransformations()
    if L.vsLocalmotions2Transforms(ctypes.byref(td), ctypes.byref(many), ctypes.byref(trans)) != 0:
        raise RuntimeError("vsLocalmotions2Transforms failed: " + L.vsg_last_error().decode())
    out = [(t.x, t.y, t.alpha, t.zoom, t.barrel, t.extra) for t in trans.ts[:trans.len]]
    if trans.ts:
        L.vs_free_fn(ctypes.cast(trans.ts, c_void_p))
    return out, td.lensK


def release_wait():
    """vsLocalMotionsReleaseWait: returns once every LocalMotions array
    released by earlier calls has gone back through vs_free"""
    lib().vsLocalMotionsReleaseWait()


def estimate_lens(width, height, frames, focal=0.0):
    """vsEstimateLensDistortion of a clip's LocalMotions with the default
    search settings and the given focal length -> dict of the estimate"""
    L = lib()
    fi = VSFrameInfo()
    if not L.vsFrameInfoInit(ctypes.byref(fi), width, height, PF_YUV420P):
        raise ValueError("vsFrameInfoInit failed")
    clip = frames if isinstance(frames, ClipMotions) else ClipMotions(frames)
    many = _many(clip)
    cfg = L.vsLensEstimateGetDefaultConfig()
    cfg.f = focal
    e = L.vsEstimateLensDistortion(ctypes.byref(fi), ctypes.byref(many), ctypes.byref(cfg))
    return {n: getattr(e, n) for n, _ in VSLensEstimate._fields_}


def transform_distance(width, height, a, b):
    """largest displacement difference (px) of the frame corners between two
    (x, y, alpha) transforms about the frame centre"""
    cx, cy = width / 2.0, height / 2.0
    worst = 0.0
    for qx, qy in ((0, 0), (width, 0), (0, height), (width, height)):
        dx, dy = qx - cx, qy - cy
        pa = (np.cos(a[2]) * dx - np.sin(a[2]) * dy + a[0], np.sin(a[2]) * dx + np.cos(a[2]) * dy + a[1])
        pb = (np.cos(b[2]) * dx - np.sin(b[2]) * dy + b[0], np.sin(b[2]) * dx + np.cos(b[2]) * dy + b[1])
        worst = max(worst, float(np.hypot(pa[0] - pb[0], pa[1] - pb[1])))
    return worst


def compare_forms(plain, learned, coarse_size):
    """per-field agreement of two runs' LocalMotions on the same frames:
    (coarse fields compared, coarse equal, all fields compared, all equal);
    a field is identified by its position and size"""
    cn = ce = an = ae = 0
    for fp, fl in zip(plain, learned):
        dp = {(m[2], m[3], m[4]): m for m in fp}
        dl = {(m[2], m[3], m[4]): m for m in fl}
        for key, m in dp.items():
            same = dl.get(key) == m
            an += 1
            ae += same
            if key[2] == coarse_size:
                cn += 1
                ce += same
    return cn, ce, an, ae
