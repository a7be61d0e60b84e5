"""Tagging and Berger-Rigoutsos clustering on the host.

Restates EllipticBSAM/src/ebTagCluster.m:1-380 with MATLAB's tie-breaking
(first index on min/max, ascending ``find``, LIFO stack popped from the end,
column-major scans).  Boxes are 1-based inclusive ``[ilo ihi jlo jhi]`` in the
index space of the current finest level.  Tagging needs the finest level's
solution, which the caller downloads from the device before calling here.
"""
from __future__ import annotations

import warnings

import numpy as np

from .hierarchy import restrict2
from .problem import BsamError, call_tagfun


def dilate(M, b):
    """ebTagCluster.m:113-127 -- dilation by a (2b+1)^2 square."""
    if b <= 0:
        return M
    n1, n2 = M.shape
    P = np.zeros((n1 + 2 * b, n2 + 2 * b), dtype=bool)
    P[b:b + n1, b:b + n2] = M
    acc = np.zeros((n1, n2), dtype=bool)
    for dx in range(-b, b + 1):
        for dy in range(-b, b + 1):
            acc |= P[b + dx:b + n1 + dx, b + dy:b + n2 + dy]
    return acc


def erode8(C):
    """ebTagCluster.m:130-141 -- cell and all in-domain 8-neighbours covered."""
    n1, n2 = C.shape
    P = np.ones((n1 + 2, n2 + 2), dtype=bool)
    P[1:n1 + 1, 1:n2 + 1] = C
    A = np.ones((n1, n2), dtype=bool)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            A &= P[1 + dx:n1 + 1 + dx, 1 + dy:n2 + 1 + dy]
    return A


def _sub(T, bb):
    return T[bb[0] - 1:bb[1], bb[2] - 1:bb[3]]


def tight_bbox(T, bb):
    """ebTagCluster.m:201-208."""
    sub = _sub(T, bb)
    rows = np.any(sub, axis=1)
    cols = np.any(sub, axis=0)
    if not rows.any():
        return None
    ri = np.flatnonzero(rows)
    ci = np.flatnonzero(cols)
    return [bb[0] + ri[0], bb[0] + ri[-1], bb[2] + ci[0], bb[2] + ci[-1]]


def pick_split(sig, minW):
    """ebTagCluster.m:211-236: (s, quality) with 2 = hole, 1 = inflection."""
    n = sig.size
    s = 0
    quality = 0
    holes = np.flatnonzero(sig == 0) + 1  # 1-based positions
    holes = holes[(holes > minW) & (holes <= n - minW)]
    if holes.size:
        ii = int(np.argmin(np.abs(holes - n / 2)))
        return int(holes[ii]), 2
    if n < 2 * minW:
        return s, quality
    # d2(m), m = 1..n-2, lives at position m+1
    d2 = sig[2:n] - 2 * sig[1:n - 1] + sig[0:n - 2]
    best = -1.0
    for i in range(2, n - 1):  # i = 2..n-2
        j = abs(d2[i - 1] - d2[i - 2])
        cut = i
        if cut >= minW and (n - cut) >= minW and j > best:
            best = j
            s = cut
    if s > 0:
        quality = 1
    return s, quality


def br_cluster(T, minW, eff):
    """ebTagCluster.m:144-198 (explicit LIFO stack)."""
    out = []
    stack = [tight_bbox(T, [1, T.shape[0], 1, T.shape[1]])]
    while stack:
        bb = stack.pop()
        if bb is None:
            continue
        bb = tight_bbox(T, bb)
        if bb is None:
            continue
        w = (bb[1] - bb[0] + 1, bb[3] - bb[2] + 1)
        sub = _sub(T, bb)
        fill = np.count_nonzero(sub) / sub.size
        if fill >= eff or (w[0] <= minW and w[1] <= minW):
            out.append(bb)
            continue
        sigx = sub.sum(axis=1).astype(float)
        sigy = sub.sum(axis=0).astype(float)
        sx, qx = pick_split(sigx, minW)
        sy, qy = pick_split(sigy, minW)
        if qx == 0 and qy == 0:
            if w[0] >= 2 * minW or w[1] >= 2 * minW:
                if w[0] >= w[1]:
                    s = w[0] // 2
                    stack.append([bb[0], bb[0] + s - 1, bb[2], bb[3]])
                    stack.append([bb[0] + s, bb[1], bb[2], bb[3]])
                else:
                    s = w[1] // 2
                    stack.append([bb[0], bb[1], bb[2], bb[2] + s - 1])
                    stack.append([bb[0], bb[1], bb[2] + s, bb[3]])
            else:
                out.append(bb)
            continue
        usex = qx > qy or (qx == qy and w[0] >= w[1])
        if qx == 0:
            usex = False
        if qy == 0:
            usex = True
        if usex:
            stack.append([bb[0], bb[0] + sx - 1, bb[2], bb[3]])
            stack.append([bb[0] + sx, bb[1], bb[2], bb[3]])
        else:
            stack.append([bb[0], bb[1], bb[2], bb[2] + sy - 1])
            stack.append([bb[0], bb[1], bb[2] + sy, bb[3]])
    return out


def clip_allowed(boxes, T, allowed):
    """ebTagCluster.m:239-264."""
    out = []
    stack = [list(b) for b in boxes]
    while stack:
        bb = stack.pop()
        if not _sub(T, bb).any():
            continue
        if _sub(allowed, bb).all():
            out.append(bb)
            continue
        w = (bb[1] - bb[0] + 1, bb[3] - bb[2] + 1)
        if w[0] >= w[1]:
            s = w[0] // 2
            stack.append([bb[0], bb[0] + s - 1, bb[2], bb[3]])
            stack.append([bb[0] + s, bb[1], bb[2], bb[3]])
        else:
            s = w[1] // 2
            stack.append([bb[0], bb[1], bb[2], bb[2] + s - 1])
            stack.append([bb[0], bb[1], bb[2] + s, bb[3]])
    return out


def _strip_ok(boxes, allowed, b, dim, pos):
    """ebTagCluster.m:359-380."""
    bb = boxes[b]
    if dim == 1:
        strip = allowed[pos - 1, bb[2] - 1:bb[3]]
        s = (pos, pos, bb[2], bb[3])
    else:
        strip = allowed[bb[0] - 1:bb[1], pos - 1]
        s = (bb[0], bb[1], pos, pos)
    if not strip.all():
        return False
    for o in range(len(boxes)):
        if o == b:
            continue
        ob = boxes[o]
        if s[0] <= ob[1] and s[1] >= ob[0] and s[2] <= ob[3] and s[3] >= ob[2]:
            return False
    return True


def grow_min(boxes, allowed, minW, n):
    """ebTagCluster.m:267-297 (widen undersized boxes one strip at a time)."""
    boxes = [list(b) for b in boxes]
    for b in range(len(boxes)):
        for dim in (1, 2):
            lo = boxes[b][2 * dim - 2]
            hi = boxes[b][2 * dim - 1]
            nd = n[dim - 1]
            side = 0
            while hi - lo + 1 < minW:
                grown = False
                for tryside in (side, 1 - side):
                    if tryside == 0 and lo > 1:
                        if _strip_ok(boxes, allowed, b, dim, lo - 1):
                            lo -= 1
                            grown = True
                            break
                    elif tryside == 1 and hi < nd:
                        if _strip_ok(boxes, allowed, b, dim, hi + 1):
                            hi += 1
                            grown = True
                            break
                if not grown:
                    break
                side = 1 - side
            boxes[b][2 * dim - 2] = lo
            boxes[b][2 * dim - 1] = hi
    return boxes


def merge_or_drop_thin(boxes, allowed):
    """ebTagCluster.m:299-356."""
    boxes = [list(b) for b in boxes]

    def thin(bb):
        return (bb[1] - bb[0] + 1 < 2) or (bb[3] - bb[2] + 1 < 2)

    b = 0
    while b < len(boxes):
        if not thin(boxes[b]):
            b += 1
            continue
        merged = False
        nb = len(boxes)
        bx = boxes[b]
        areab = (bx[1] - bx[0] + 1) * (bx[3] - bx[2] + 1)
        ua = np.full(nb, np.inf)
        for o in range(nb):
            if o == b:
                continue
            ob = boxes[o]
            u = [min(bx[0], ob[0]), max(bx[1], ob[1]), min(bx[2], ob[2]), max(bx[3], ob[3])]
            ua[o] = (u[1] - u[0] + 1) * (u[3] - u[2] + 1)
        order = np.argsort(ua, kind="stable")
        for o in order:
            if o == b or not np.isfinite(ua[o]):
                continue
            sb = boxes[b]
            ob = boxes[o]
            if not (sb[0] - 1 <= ob[1] and sb[1] + 1 >= ob[0] and sb[2] - 1 <= ob[3] and sb[3] + 1 >= ob[2]):
                continue
            u = [min(sb[0], ob[0]), max(sb[1], ob[1]), min(sb[2], ob[2]), max(sb[3], ob[3])]
            areao = (ob[1] - ob[0] + 1) * (ob[3] - ob[2] + 1)
            if ua[o] > 1.5 * (areab + areao):
                continue
            if not _sub(allowed, u).all():
                continue
            clash = False
            for t in range(nb):
                if t == b or t == o:
                    continue
                tb = boxes[t]
                if u[0] <= tb[1] and u[1] >= tb[0] and u[2] <= tb[3] and u[3] >= tb[2]:
                    clash = True
                    break
            if clash:
                continue
            boxes[o] = u
            del boxes[b]
            merged = True
            break
        if not merged:
            warnings.warn("eb_tag_cluster:dropthin: dropping an ungrowable sliver cluster box")
            del boxes[b]
    return boxes


def ulap_indicator(Q):
    """Undivided Laplacian magnitude (ebTagCluster.m:50-55)."""
    mx, my = Q.shape[0] - 2, Q.shape[1] - 2
    sdx = Q[2:mx + 2, 1:my + 1] - 2 * Q[1:mx + 1, 1:my + 1] + Q[0:mx, 1:my + 1]
    sdy = Q[1:mx + 1, 2:my + 2] - 2 * Q[1:mx + 1, 1:my + 1] + Q[1:mx + 1, 0:my]
    return np.sqrt(sdx ** 2 + sdy ** 2)


def _op2d(Q, DeW, DnS, CC, hx, hy):
    """L(u) = -div(D grad u) + C u on a padded array (ebKernels.m:13-25)."""
    mx, my = CC.shape
    feW = DeW * (Q[1:mx + 2, 1:my + 1] - Q[0:mx + 1, 1:my + 1])
    fnS = DnS * (Q[1:mx + 1, 1:my + 2] - Q[1:mx + 1, 0:my + 1])
    return -(feW[1:, :] - feW[:-1, :]) / (hx * hx) - (fnS[:, 1:] - fnS[:, :-1]) / (hy * hy) + CC * Q[1:mx + 1, 1:my + 1]


def rte_indicator(Q, DeW, DnS, CC, hx, hy):
    """Relative truncation error |L_2h(R u) - R(L_h u)| per fine cell
    (ebRte.m:14-42).  R u carries a ghost ring averaged from the fine ghost
    pairs; the 2h operator uses face-averaged D and the 4-cell mean of C."""
    mx, my = Q.shape[0] - 2, Q.shape[1] - 2
    if mx % 2 or my % 2:
        raise BsamError("eb_tag_cluster:rte", "the rte indicator needs even patch widths")
    cmx, cmy = mx // 2, my // 2
    rlh = restrict2(_op2d(Q, DeW, DnS, CC, hx, hy))
    R = np.zeros((cmx + 2, cmy + 2))
    R[1:-1, 1:-1] = restrict2(Q[1:-1, 1:-1])
    j = np.arange(1, cmy + 1)
    R[0, j] = 0.5 * (Q[0, 2 * j - 1] + Q[0, 2 * j])
    R[cmx + 1, j] = 0.5 * (Q[mx + 1, 2 * j - 1] + Q[mx + 1, 2 * j])
    i = np.arange(1, cmx + 1)
    R[i, 0] = 0.5 * (Q[2 * i - 1, 0] + Q[2 * i, 0])
    R[i, cmy + 1] = 0.5 * (Q[2 * i - 1, my + 1] + Q[2 * i, my + 1])
    cDeW = 0.5 * (DeW[0:mx + 1:2, 0:my - 1:2] + DeW[0:mx + 1:2, 1:my:2])
    cDnS = 0.5 * (DnS[0:mx - 1:2, 0:my + 1:2] + DnS[1:mx:2, 0:my + 1:2])
    tau = _op2d(R, cDeW, cDnS, restrict2(CC), 2 * hx, 2 * hy) - rlh
    return np.repeat(np.repeat(np.abs(tau), 2, axis=0), 2, axis=1)


def _threshold(opts, level):
    thr = opts.TagThreshold
    return thr(level) if callable(thr) else thr


def tag_cluster(H, with_info=False, device=None):
    """ebTagCluster.m:1-110 -> boxes (list of [ilo ihi jlo jhi], level-k space).

    ``device``: a GpuHierarchy holding the same solution as ``H``; the 'ulap'
    and 'rte' indicators, the buffer and the nesting clip then run on the
    device (bsam_tag_level) and only the tag map comes back."""
    opts = H.opts
    k = H.nlev
    lev = H.L(k)
    n = lev.n
    dom = H.prob.domain
    info = {"indicator": [], "Traw": None, "T": None, "allowed": None}
    ind = opts.Indicator.lower()
    cover = np.zeros(n, dtype=bool)
    for p in range(lev.np):
        b = lev.box[p]
        cover[b[0] - 1:b[1], b[2] - 1:b[3]] = True
    allowed = erode8(cover)
    if device is not None and ind in ("ulap", "rte"):
        T, info["Traw"] = device.tag_level(k, ind, _threshold(opts, k - H.rootIdx + 1), opts.TagBuffer)
    else:
        T = np.zeros(n, dtype=bool)
        for p in range(lev.np):
            b = lev.box[p]
            if ind == "fun":
                if opts.TagFunction is None:
                    raise BsamError("eb_tag_cluster:tagfun", "Indicator=fun needs TagFunction.")
                XC, YC = lev.cell_centers(dom, p)
                tp = call_tagfun(opts.TagFunction, XC, YC, k - H.rootIdx + 1)
            elif ind in ("ulap", "rte"):
                if ind == "ulap":
                    s = ulap_indicator(lev.Q[p])
                else:
                    s = rte_indicator(lev.Q[p], lev.DeW[p], lev.DnS[p], lev.CC[p], lev.h[0], lev.h[1])
                tp = s > _threshold(opts, k - H.rootIdx + 1)
                if with_info:
                    info["indicator"].append(s)
            else:
                raise BsamError("eb_tag_cluster:ind", f"unknown Indicator {opts.Indicator}")
            T[b[0] - 1:b[1], b[2] - 1:b[3]] |= tp
        info["Traw"] = T.copy()
        T = dilate(T, opts.TagBuffer)
        T = T & allowed
    info["T"] = T
    info["allowed"] = allowed
    if not T.any():
        return ([], info) if with_info else []
    if device is not None:
        boxes = cluster_native(T, allowed, opts.MinBoxWidth, opts.MinFillRatio)
    else:
        boxes = br_cluster(T, opts.MinBoxWidth, opts.MinFillRatio)
        boxes = clip_allowed(boxes, T, allowed)
        boxes = grow_min(boxes, allowed, opts.MinBoxWidth, n)
        boxes = merge_or_drop_thin(boxes, allowed)
    boxes = [[int(v) for v in bb] for bb in boxes]
    return (boxes, info) if with_info else boxes


def cluster_native(T, allowed, min_width, fill_ratio):
    """The library's clustering (bsam_cluster_tags, host C++): the same steps
    as br_cluster -> clip_allowed -> grow_min -> merge_or_drop_thin."""
    import ctypes

    from . import _lib

    lib = _lib.load()
    t = np.asfortranarray(T, dtype=np.uint8)
    a = np.asfortranarray(allowed, dtype=np.uint8)
    n1, n2 = t.shape
    nb, nd = ctypes.c_int(), ctypes.c_int()
    cap = 64
    while True:
        buf = np.zeros(4 * cap, dtype=np.int32)
        rc = lib.bsam_cluster_tags(t.ctypes.data, a.ctypes.data, n1, n2, int(min_width), float(fill_ratio),
                                   buf.ctypes.data_as(_lib.P_int), cap, ctypes.byref(nb), ctypes.byref(nd))
        if rc != 0 and nb.value > cap:
            cap = nb.value
            continue
        _lib.check(rc, "bsam_cluster_tags")
        break
    for _ in range(nd.value):
        warnings.warn("eb_tag_cluster:dropthin: dropping an ungrowable sliver cluster box")
    return buf[:4 * nb.value].reshape(-1, 4).tolist()


def cluster_per_parent(T, allowed, parents, min_width, fill_ratio):
    """ebtTagCluster.m:93-128 (tree variant): br_cluster -> clip_allowed ->
    grow_min -> merge_or_drop_thin inside every parent box [ilo ihi jlo jhi]
    (1-based) on that parent's own tags and allowed region; rows
    [parent, ilo, ihi, jlo, jhi] in level coordinates."""
    out = []
    for q, (ilo, ihi, jlo, jhi) in enumerate(parents, 1):
        Tq = T[ilo - 1:ihi, jlo - 1:jhi]
        if not Tq.any():
            continue
        Aq = allowed[ilo - 1:ihi, jlo - 1:jhi]
        b = br_cluster(Tq, min_width, fill_ratio)
        b = clip_allowed(b, Tq, Aq)
        b = grow_min(b, Aq, min_width, Tq.shape)
        b = merge_or_drop_thin(b, Aq)
        out += [[q, int(bb[0]) + ilo - 1, int(bb[1]) + ilo - 1, int(bb[2]) + jlo - 1, int(bb[3]) + jlo - 1]
                for bb in b]
    return out


def cluster_per_parent_native(T, allowed, parents, min_width, fill_ratio):
    """The library's per-parent clustering (bsam_cluster_tags_per_parent)."""
    import ctypes

    from . import _lib

    lib = _lib.load()
    t = np.asfortranarray(T, dtype=np.uint8)
    a = np.asfortranarray(allowed, dtype=np.uint8)
    par = np.ascontiguousarray(np.asarray(parents, dtype=np.int32).reshape(-1))
    n1, n2 = t.shape
    nb, nd = ctypes.c_int(), ctypes.c_int()
    cap = 64
    while True:
        buf = np.zeros(5 * cap, dtype=np.int32)
        rc = lib.bsam_cluster_tags_per_parent(t.ctypes.data, a.ctypes.data, n1, n2, par.ctypes.data_as(_lib.P_int),
                                              len(parents), int(min_width), float(fill_ratio),
                                              buf.ctypes.data_as(_lib.P_int), cap, ctypes.byref(nb),
                                              ctypes.byref(nd))
        if rc != 0 and nb.value > cap:
            cap = nb.value
            continue
        _lib.check(rc, "bsam_cluster_tags_per_parent")
        break
    for _ in range(nd.value):
        warnings.warn("ebt_tag_cluster:dropthin: dropping an ungrowable sliver cluster box")
    return buf[:5 * nb.value].reshape(-1, 5).tolist()
