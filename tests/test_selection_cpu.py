"""The device selection's formulation (each field's place = number of fields
that precede it under (contrast desc, index asc)) against the reference's
qsort-based selectfields as restated by the oracle, on inputs full of ties."""
import ctypes

import numpy as np
import pytest

import pyoracle


def rank_select(c, rows, max_fields):
    """mirror of k_select's rule, in plain Python"""
    n = len(c)
    numsegms = rows + 1
    segmlen = n // numsegms + 1
    per_seg = max_fields // numsegms
    pick = [0] * n
    seg_cnt = [0] * numsegms
    for i in range(n):
        if not c[i] > 0:
            continue
        s = i // segmlen
        start, end = s * segmlen, min(s * segmlen + segmlen, n)
        pos = sum(1 for j in range(start, end) if c[j] > c[i] or (c[j] == c[i] and j < i))
        if pos < per_seg:
            pick[i] = pos + 1
            seg_cnt[s] += 1
    base = np.concatenate([[0], np.cumsum(seg_cnt)[:-1]]).astype(int)
    picked = sum(seg_cnt)
    out = [None] * n
    remaining = max_fields - picked
    nleft = 0
    for i in range(n):
        if pick[i]:
            out[base[i // segmlen] + pick[i] - 1] = i
        elif remaining > 0 and c[i] > 0:
            c2 = [0.0 if pick[j] else c[j] for j in range(n)]
            rank = sum(1 for j in range(n) if c2[j] > c[i] or (c2[j] == c[i] and j < i))
            if rank < remaining:
                out[picked + rank] = i
                nleft += 1
    return out[:picked + nleft]


@pytest.mark.parametrize("seed", range(12))
def test_rank_selection_equals_stable_sort_selection(seed):
    rng = np.random.default_rng(seed)
    rows = int(rng.integers(2, 8))
    cols = int(rng.integers(3, 12))
    n = rows * cols
    acc = int(rng.integers(1, 16))
    max_fields = acc * n // 15
    ct = float(rng.choice([0.0, 0.1, 0.25, 0.3]))
    # few distinct (min, max) pairs -> many exactly equal contrasts
    pairs = rng.integers(0, 256, size=(4, 2))
    mm = pairs[rng.integers(0, 4, size=n)]
    mn, mx = np.minimum(mm[:, 0], mm[:, 1]), np.maximum(mm[:, 0], mm[:, 1])
    L = pyoracle.lib()
    contrast = [L.or_contrast(int(a), int(b)) for a, b in zip(mn, mx)]
    fs = pyoracle.OrFields()
    fs.fieldNum, fs.fieldRows, fs.maxFields, fs.contrastThreshold = n, rows, max_fields, ct
    cbuf = (ctypes.c_double * n)(*contrast)
    idx = (ctypes.c_int * n)()
    sc = (ctypes.c_double * n)()
    cnt = L.or_select(ctypes.byref(fs), cbuf, idx, sc)
    want = [idx[i] for i in range(cnt)]
    thr = [0.0 if v < ct else v for v in contrast]
    assert rank_select(thr, rows, max_fields) == want
