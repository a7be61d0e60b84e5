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


def _parent_cases():
    """Level maps covered by a few disjoint parent patches (a tiling of
    vertical bands, or scattered rectangles), tags dilated across patch edges
    as the global buffer does, clipped to the nesting region of the cover."""
    rng = np.random.default_rng(29)
    out = []
    for n1, n2 in ((32, 32), (64, 48), (96, 96)):
        for layout in ("bands", "scattered"):
            if layout == "bands":
                cuts = np.sort(rng.choice(np.arange(8, n1 - 8, 2), 2, replace=False))
                edges = [0, *cuts.tolist(), n1]
                parents = [[edges[q] + 1, edges[q + 1], 1, n2] for q in range(3)]
            else:
                parents = []
                for i0 in range(0, n1 - 12, n1 // 2):
                    for j0 in range(0, n2 - 12, n2 // 2):
                        parents.append([i0 + 3, i0 + n1 // 2 - 2, j0 + 3, j0 + n2 // 2 - 2])
            cover = np.zeros((n1, n2), dtype=bool)
            for ilo, ihi, jlo, jhi in parents:
                cover[ilo - 1:ihi, jlo - 1:jhi] = True
            allowed = tc.erode8(cover)
            for dens in (0.03, 0.2):
                raw = rng.random((n1, n2)) < dens
                for _ in range(3):
                    ci, cj, r = rng.integers(0, n1), rng.integers(0, n2), rng.integers(2, 8)
                    raw[max(ci - r, 0):ci + r, max(cj - r, 0):cj + r] = True
                T = tc.dilate(raw, int(rng.integers(0, 3))) & allowed
                for minw, fill in ((4, 0.7), (2, 0.8), (8, 0.5)):
                    out.append((T, allowed, parents, minw, fill))
    return out


def test_native_per_parent_clustering_equals_restatement():
    """ebtTagCluster.m:93-128: the library's per-parent clustering equals the
    restatement box for box, every box lies inside its own parent, and the
    boxes of one parent do not overlap."""
    nboxes = nparents = 0
    for T, allowed, parents, minw, fill in _parent_cases():
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            want = tc.cluster_per_parent(T, allowed, parents, minw, fill)
            nw = len(w)
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            got = tc.cluster_per_parent_native(T, allowed, parents, minw, fill)
            ng = len(w)
        assert got == want
        assert ng == nw
        for q, ilo, ihi, jlo, jhi in got:
            p = parents[q - 1]
            assert p[0] <= ilo <= ihi <= p[1] and p[2] <= jlo <= jhi <= p[3]
        for a in got:
            for b in got:
                if a is not b and a[0] == b[0]:
                    assert a[2] < b[1] or b[2] < a[1] or a[4] < b[3] or b[4] < a[3]
        nboxes += len(got)
        nparents += len({b[0] for b in got})
    print(f"{nboxes} boxes over {nparents} parents", flush=True)
    assert nboxes > 50


def test_per_parent_clustering_errors():
    from bsamgpu.problem import BsamError

    T = np.ones((8, 8), dtype=bool)
    with pytest.raises(BsamError):
        tc.cluster_per_parent_native(T, T, [[0, 4, 1, 8]], 2, 0.7)  # outside the level
    with pytest.raises(BsamError):
        tc.cluster_per_parent_native(T, T, [[1, 8, 1, 8]], 0, 0.7)  # MinBoxWidth
    assert tc.cluster_per_parent_native(np.zeros((8, 8), dtype=bool), T, [[1, 8, 1, 8]], 2, 0.7) == []
