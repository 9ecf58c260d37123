"""CPU checks of the container and frame helpers of the detect ABI
(include/vsvector.h, include/frameinfo.h), through the C library.

The LocalMotions a caller receives are a VSVector of separately allocated
elements that vs_vector_del hands back to vs_free; set fills gaps with NULL;
filter and concat share element pointers.  Frame helpers: vsFrameAllocate
gives zero-filled planes with packed rows, vsFrameCopy copies row by row when
the strides differ.  (Reference behaviour: src/vsvector.c, src/frameinfo.c.)
"""
import ctypes

import numpy as np

from vidstab_amd import api

c_void_p, c_int = ctypes.c_void_p, ctypes.c_int


def _L():
    L = api.lib()
    L.vs_vector_init.argtypes = [ctypes.POINTER(api.VSVector), c_int]
    L.vs_vector_append_dup.argtypes = [ctypes.POINTER(api.VSVector), c_void_p, c_int]
    L.vs_vector_append.argtypes = [ctypes.POINTER(api.VSVector), c_void_p]
    L.vs_vector_set.argtypes = [ctypes.POINTER(api.VSVector), c_int, c_void_p, ctypes.POINTER(c_void_p)]
    L.vs_vector_set_dup.argtypes = [ctypes.POINTER(api.VSVector), c_int, c_void_p, c_int, ctypes.POINTER(c_void_p)]
    L.vs_vector_concat.argtypes = [ctypes.POINTER(api.VSVector), ctypes.POINTER(api.VSVector)]
    L.vs_vector_concat.restype = api.VSVector
    L.vs_vector_fini.argtypes = [ctypes.POINTER(api.VSVector)]
    L.vs_vector_zero.argtypes = [ctypes.POINTER(api.VSVector)]
    L.vsFrameAllocate.argtypes = [ctypes.POINTER(api.VSFrame), ctypes.POINTER(api.VSFrameInfo)]
    L.vsFrameCopy.argtypes = [ctypes.POINTER(api.VSFrame), ctypes.POINTER(api.VSFrame), ctypes.POINTER(api.VSFrameInfo)]
    L.vsFrameFree.argtypes = [ctypes.POINTER(api.VSFrame)]
    L.vsFrameFillFromBuffer.argtypes = [ctypes.POINTER(api.VSFrame), c_void_p, ctypes.POINTER(api.VSFrameInfo)]
    L.vsFrameIsNull.argtypes = [ctypes.POINTER(api.VSFrame)]
    return L


def _ints(vec, n):
    return [ctypes.cast(vec.data[i], ctypes.POINTER(c_int)).contents.value if vec.data[i] else None
            for i in range(n)]


def test_vector_append_grow_get_and_delete():
    L = _L()
    v = api.VSVector()
    assert L.vs_vector_init(ctypes.byref(v), 0) == 0 and not v.data and v.buffersize == 0
    for k in range(100):
        x = c_int(k * 3)
        assert L.vs_vector_append_dup(ctypes.byref(v), ctypes.byref(x), 4) == 0
    assert L.vs_vector_size(ctypes.byref(v)) == 100 and v.buffersize >= 100
    assert _ints(v, 100) == [3 * k for k in range(100)]
    assert L.vs_vector_get(ctypes.byref(v), 100) is None and L.vs_vector_get(ctypes.byref(v), -1) is None
    assert L.vs_vector_del(ctypes.byref(v)) == 0
    assert v.nelems == 0 and v.buffersize == 0 and not v.data


def test_vector_set_fills_gaps_and_returns_the_old_element():
    L = _L()
    v = api.VSVector()
    L.vs_vector_init(ctypes.byref(v), 2)
    x, y = c_int(7), c_int(9)
    old = c_void_p()
    assert L.vs_vector_set_dup(ctypes.byref(v), 5, ctypes.byref(x), 4, ctypes.byref(old)) == 0
    assert old.value is None and v.nelems == 6
    assert _ints(v, 6) == [None] * 5 + [7]
    first = v.data[5]
    assert L.vs_vector_set_dup(ctypes.byref(v), 5, ctypes.byref(y), 4, ctypes.byref(old)) == 0
    assert old.value == first and _ints(v, 6)[5] == 9
    L.vs_vector_del(ctypes.byref(v))   # frees the non-NULL elements only


def test_vector_concat_shares_elements_in_order():
    L = _L()
    a, b = api.VSVector(), api.VSVector()
    L.vs_vector_init(ctypes.byref(a), 0)
    L.vs_vector_init(ctypes.byref(b), 0)
    for k in range(3):
        x = c_int(k)
        L.vs_vector_append_dup(ctypes.byref(a), ctypes.byref(x), 4)
    for k in range(10, 12):
        x = c_int(k)
        L.vs_vector_append_dup(ctypes.byref(b), ctypes.byref(x), 4)
    c = L.vs_vector_concat(ctypes.byref(a), ctypes.byref(b))
    assert c.nelems == 5 and _ints(c, 5) == [0, 1, 2, 10, 11]
    assert c.data[0] == a.data[0] and c.data[4] == b.data[1]
    L.vs_vector_fini(ctypes.byref(a))   # pointer arrays only: c owns the elements now
    L.vs_vector_fini(ctypes.byref(b))
    L.vs_vector_del(ctypes.byref(c))


def test_frame_allocate_copy_and_fill():
    L = _L()
    fi = api.VSFrameInfo()
    assert L.vsFrameInfoInit(ctypes.byref(fi), 64, 32, api.PF_YUV420P) == 1
    fr = api.VSFrame()
    L.vsFrameAllocate(ctypes.byref(fr), ctypes.byref(fi))
    assert list(fr.linesize) == [64, 32, 32, 0] and not L.vsFrameIsNull(ctypes.byref(fr))
    assert not np.ctypeslib.as_array(fr.data[0], (32 * 64,)).any()
    # a padded source (stride 80) copied into the packed frame, row by row
    src = np.arange(32 * 80, dtype=np.uint32).astype(np.uint8).reshape(32, 80)
    ch = np.full((16, 40), 200, dtype=np.uint8)
    s = api.VSFrame()
    s.data[0] = src.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    s.data[1] = s.data[2] = ch.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    s.linesize[0], s.linesize[1], s.linesize[2] = 80, 40, 40
    L.vsFrameCopy(ctypes.byref(fr), ctypes.byref(s), ctypes.byref(fi))
    assert np.array_equal(np.ctypeslib.as_array(fr.data[0], (32 * 64,)).reshape(32, 64), src[:, :64])
    assert (np.ctypeslib.as_array(fr.data[2], (16 * 32,)) == 200).all()
    L.vsFrameFree(ctypes.byref(fr))
    assert L.vsFrameIsNull(ctypes.byref(fr)) and list(fr.linesize) == [0, 0, 0, 0]
    # packed RGB: one plane of width * 3 bytes a row; planes of one buffer back to back
    assert L.vsFrameInfoInit(ctypes.byref(fi), 10, 4, api.PF_RGB24) == 1 and fi.planes == 1
    buf = np.zeros(120, dtype=np.uint8)
    L.vsFrameFillFromBuffer(ctypes.byref(fr), buf.ctypes.data, ctypes.byref(fi))
    assert fr.linesize[0] == 30 and not fr.data[1]
    assert L.vsFrameInfoInit(ctypes.byref(fi), 64, 32, 4) == 1   # PF_YUV410P
    L.vsFrameFillFromBuffer(ctypes.byref(fr), buf.ctypes.data, ctypes.byref(fi))
    base = ctypes.cast(fr.data[0], c_void_p).value
    assert ctypes.cast(fr.data[1], c_void_p).value - base == 64 * 32
    assert ctypes.cast(fr.data[2], c_void_p).value - base == 64 * 32 + 16 * 8


def test_frameinfo_rejects_and_resets_the_format():
    L = _L()
    fi = api.VSFrameInfo()
    assert L.vsFrameInfoInit(ctypes.byref(fi), 63, 32, api.PF_YUV422P) == 0 and fi.pFormat == 0
    assert L.vsFrameInfoInit(ctypes.byref(fi), 64, 33, api.PF_YUV422P) == 1
    assert L.vsFrameInfoInit(ctypes.byref(fi), 64, 32, 99) == 0 and fi.pFormat == 0 and fi.planes == 3
    assert L.vsFrameInfoInit(ctypes.byref(fi), 0, 32, api.PF_GRAY8) == 0
