"""The library's host clustering (bsam_cluster_tags, C++) against the host
layer's restatement of ebTagCluster.m:92-380 (br_cluster, clip_allowed,
grow_min, merge_or_drop_thin), which the published hierarchies pin (README
elliptic and starfish patch counts).  Stated tolerance: identical box lists,
in order, and the same number of dropped slivers."""
import warnings

import numpy as np
import pytest

from bsamgpu import tagcluster as tc


def _python(T, allowed, minw, fill):
    n = T.shape
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        b = tc.br_cluster(T, minw, fill)
        b = tc.clip_allowed(b, T, allowed)
        b = tc.grow_min(b, allowed, minw, n)
        b = tc.merge_or_drop_thin(b, allowed)
    return [[int(v) for v in bb] for bb in b], len(w)


def _native(T, allowed, minw, fill):
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        b = tc.cluster_native(T, allowed, minw, fill)
    return b, len(w)


def _cases():
    rng = np.random.default_rng(17)
    out = []
    for n1, n2 in ((16, 16), (32, 24), (64, 64), (40, 72), (128, 96)):
        cover = np.zeros((n1, n2), dtype=bool)
        # a coverage of a few rectangles, so that the nesting clip bites
        for _ in range(3):
            i0, j0 = rng.integers(0, n1 // 2), rng.integers(0, n2 // 2)
            cover[i0:i0 + rng.integers(n1 // 4, n1), j0:j0 + rng.integers(n2 // 4, n2)] = True
        if rng.random() < 0.4:
            cover[:] = True
        allowed = tc.erode8(cover)
        for dens in (0.02, 0.1, 0.35):
            raw = rng.random((n1, n2)) < dens
            blobs = np.zeros_like(raw)
            for _ in range(4):  # a few dense blobs on top of the noise
                ci, cj, r = rng.integers(0, n1), rng.integers(0, n2), rng.integers(2, 9)
                blobs[max(ci - r, 0):ci + r, max(cj - r, 0):cj + r] = True
            T = tc.dilate(raw & (rng.random((n1, n2)) < 0.3) | blobs, int(rng.integers(0, 3))) & allowed
            for minw, fill in ((4, 0.7), (2, 0.8), (8, 0.5), (1, 0.9)):
                out.append((T, allowed, minw, fill))
    return out


def test_native_clustering_equals_restatement():
    nboxes = ndrop = 0
    for T, allowed, minw, fill in _cases():
        if not T.any():
            continue
        bp, dp = _python(T, allowed, minw, fill)
        bn, dn = _native(T, allowed, minw, fill)
        assert bn == bp, (T.shape, minw, fill)
        assert dn == dp
        nboxes += len(bp)
        ndrop += dp
    print(f"{nboxes} boxes, {ndrop} slivers dropped", flush=True)
    assert nboxes > 100


def test_native_clustering_errors_and_small_buffer():
    import ctypes

    from bsamgpu import _lib

    lib = _lib.load()
    T = np.zeros((8, 8), dtype=np.uint8, order="F")
    T[1:3, 1:3] = 1
    T[5:7, 5:7] = 1
    A = np.ones((8, 8), dtype=np.uint8, order="F")
    buf = np.zeros(4, dtype=np.int32)
    nb = ctypes.c_int()
    rc = lib.bsam_cluster_tags(T.ctypes.data, A.ctypes.data, 8, 8, 2, 0.7, buf.ctypes.data_as(_lib.P_int), 1,
                               ctypes.byref(nb), None)
    assert rc == -1 and nb.value == 2
    rc = lib.bsam_cluster_tags(T.ctypes.data, A.ctypes.data, 8, 8, 0, 0.7, buf.ctypes.data_as(_lib.P_int), 1,
                               ctypes.byref(nb), None)
    assert rc == -1
    boxes = tc.cluster_native(T.astype(bool), A.astype(bool), 2, 0.7)
    assert len(boxes) == 2 and boxes == _python(T.astype(bool), A.astype(bool), 2, 0.7)[0]
