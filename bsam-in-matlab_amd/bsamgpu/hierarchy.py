"""Host-side patch hierarchy: the value struct ``H`` the solve path consumes.

Restates the reference's host layer (kept on the host by design; the device
library receives the sampled arrays through the C ABI):
  * ``make_level``      <- EllipticBSAM/src/ebMakeLevel.m:1-143
  * ``bc_fold``         <- EllipticBSAM/src/ebBcFold.m:1-42
  * ``inv_diag``        <- EllipticBSAM/src/ebKernels.m:28-34 (invdiag2d)
  * ``coeff_filldown``  <- EllipticBSAM/src/ebCoeffFilldown.m:1-62
  * ``covered_masks``   <- EllipticBSAM/src/ebBuildOps.m:171-199
  * ``setup``           <- EllipticBSAM/src/ebSetup.m:1-69 (host part)
  * ``add_level``       <- EllipticBSAM/src/ebAddLevel.m:1-40 (host part)
  * ``compat_project``  <- EllipticBSAM/src/ebCompatProject.m:1-111
  * ``is_singular``     <- EllipticBSAM/src/ebCoarsest.m:58-66

Array convention: per-patch 2-D arrays indexed ``A[i, j]`` with ``i`` along x
(MATLAB rows) and ``j`` along y; boxes are 1-based inclusive
``[ilo ihi jlo jhi]`` exactly as in the reference.  Arrays are packed in
column-major (``order='F'``) order when they cross the C ABI.
"""
from __future__ import annotations

import numpy as np

from .problem import BsamError, Options, Problem


def ndgrid(x, y):
    return np.meshgrid(np.asarray(x, float), np.asarray(y, float), indexing="ij")


def _evalgrid(fh, X, Y):
    # ebMakeLevel.m:95-105
    v = fh(X, Y)
    if np.isscalar(v) or np.ndim(v) == 0:
        v = float(v) * np.ones(X.shape)
    v = np.asarray(v, dtype=float)
    if v.shape != X.shape:
        if v.size == X.size:
            v = v.reshape(X.shape)
        else:
            raise BsamError("eb_make_level:eval", "coefficient function returned wrong size.")
    if not np.all(np.isfinite(v)):
        raise BsamError("eb_make_level:eval", "coefficient function returned non-finite values.")
    return v


def _srcquad(fh, XC, YC, hx, hy, nq):
    # ebMakeLevel.m:108-143 (midpoint, or tensor Gauss-Legendre cell average)
    if nq <= 1:
        return _evalgrid(fh, XC, YC)
    xi, w = np.polynomial.legendre.leggauss(nq)
    w = w / 2
    F = np.zeros(XC.shape)
    for i in range(nq):
        for j in range(nq):
            F = F + (w[i] * w[j]) * _evalgrid(fh, XC + xi[i] * hx / 2, YC + xi[j] * hy / 2)
    return F


def _bc_side(prob: Problem, side: int, xv, yv, isvert: bool):
    # ebMakeLevel.m:78-92
    bcs = prob.bc[side]
    xv = np.asarray(xv, float)
    yv = np.asarray(yv, float)
    if isvert:
        g = bcs.g(xv * np.ones(yv.shape), yv)
        n = yv.size
    else:
        g = bcs.g(xv, yv * np.ones(xv.shape))
        n = xv.size
    g = np.asarray(g, float)
    if g.size == 1:
        g = float(g.ravel()[0]) * np.ones(n)
    g = g.reshape(-1)
    code = 1 if bcs.type == "dirichlet" else 2
    return code, g


def inv_diag(DeW, DnS, CC, hx, hy):
    """invdiag2d (ebKernels.m:28-34): 1/(sum D/h^2 + C)."""
    mx, my = CC.shape
    den = (DeW[1:mx + 1, :] + DeW[0:mx, :]) / (hx * hx) \
        + (DnS[:, 1:my + 1] + DnS[:, 0:my]) / (hy * hy) + CC
    return 1.0 / den


def bc_fold(lev, p):
    """Diagonal increment of the physical-ghost elimination (ebBcFold.m:12-42)."""
    codes = lev.bc_code[p]
    DeW = lev.DeW[p]
    DnS = lev.DnS[p]
    mx, my = lev.CC[p].shape
    hx2 = lev.h[0] ** 2
    hy2 = lev.h[1] ** 2
    add = np.zeros((mx, my))
    if codes[0] == 1:
        add[0, :] = add[0, :] + 2 * DeW[0, :] / hx2
    elif codes[0] == 2:
        add[0, :] = add[0, :] - DeW[0, :] / hx2
    if codes[1] == 1:
        add[mx - 1, :] = add[mx - 1, :] + 2 * DeW[mx, :] / hx2
    elif codes[1] == 2:
        add[mx - 1, :] = add[mx - 1, :] - DeW[mx, :] / hx2
    if codes[2] == 1:
        add[:, 0] = add[:, 0] + 2 * DnS[:, 0] / hy2
    elif codes[2] == 2:
        add[:, 0] = add[:, 0] - DnS[:, 0] / hy2
    if codes[3] == 1:
        add[:, my - 1] = add[:, my - 1] + 2 * DnS[:, my] / hy2
    elif codes[3] == 2:
        add[:, my - 1] = add[:, my - 1] - DnS[:, my] / hy2
    return add


def refresh_invdiag(lev, p):
    ID = inv_diag(lev.DeW[p], lev.DnS[p], lev.CC[p], lev.h[0], lev.h[1])
    lev.InvDiagS[p] = 1.0 / (1.0 / ID + bc_fold(lev, p))


class Level:
    """One level of the hierarchy (fields of ebMakeLevel.m:78-143)."""

    def __init__(self):
        self.n = (0, 0)
        self.h = (0.0, 0.0)
        self.np = 0
        self.box = np.zeros((0, 4), dtype=np.int64)
        self.Q, self.FTrue, self.DeW, self.DnS, self.CC, self.InvDiagS = [], [], [], [], [], []
        self.bc_code, self.bc_g = [], []   # per patch: [W E S N] codes and g vectors
        self.par0, self.faceInternal, self.covered = [], [], []
        self.sourceShift = 0.0

    def dims(self, p):
        b = self.box[p]
        return int(b[1] - b[0] + 1), int(b[3] - b[2] + 1)

    def cell_centers(self, dom, p):
        b = self.box[p]
        xc = dom[0] + (np.arange(b[0], b[1] + 1) - 0.5) * self.h[0]
        yc = dom[2] + (np.arange(b[2], b[3] + 1) - 0.5) * self.h[1]
        return ndgrid(xc, yc)


def make_level(prob: Problem, opts: Options, k: int, n, boxes) -> Level:
    """ebMakeLevel.m:1-143 (k is the 1-based level number)."""
    dom = prob.domain
    lev = Level()
    lev.n = (int(n[0]), int(n[1]))
    lev.h = ((dom[1] - dom[0]) / lev.n[0], (dom[3] - dom[2]) / lev.n[1])
    boxes = np.asarray(boxes, dtype=np.int64).reshape(-1, 4)
    lev.np = boxes.shape[0]
    lev.box = boxes.copy()
    hx, hy = lev.h
    for p in range(lev.np):
        b = boxes[p]
        if (b[0] < 1 or b[0] > n[0] or b[1] < 1 or b[1] > n[0] or b[2] < 1 or b[2] > n[1]
                or b[3] < 1 or b[3] > n[1] or b[0] > b[1] or b[2] > b[3]):
            raise BsamError("eb_make_level:box", "invalid patch box.")
        mx = int(b[1] - b[0] + 1)
        my = int(b[3] - b[2] + 1)
        lev.Q.append(np.zeros((mx + 2, my + 2)))
        xf = dom[0] + np.arange(b[0] - 1, b[1] + 1, dtype=float) * hx
        yc = dom[2] + (np.arange(b[2], b[3] + 1, dtype=float) - 0.5) * hy
        xc = dom[0] + (np.arange(b[0], b[1] + 1, dtype=float) - 0.5) * hx
        yf = dom[2] + np.arange(b[2] - 1, b[3] + 1, dtype=float) * hy
        XF, YCF = ndgrid(xf, yc)
        lev.DeW.append(_evalgrid(prob.D, XF, YCF))
        XCF, YF = ndgrid(xc, yf)
        lev.DnS.append(_evalgrid(prob.D, XCF, YF))
        XC, YC = ndgrid(xc, yc)
        lev.CC.append(_evalgrid(prob.C, XC, YC))
        lev.FTrue.append(_srcquad(prob.f, XC, YC, hx, hy, opts.SourceQuad))
        if np.any(lev.DeW[p] <= 0) or np.any(lev.DnS[p] <= 0):
            raise BsamError("eb_make_level:D", "D(x,y) must be strictly positive at faces.")
        lev.par0.append(int((b[0] + b[2]) % 2))
        fint = [bool(b[0] > 1), bool(b[1] < n[0]), bool(b[2] > 1), bool(b[3] < n[1])]
        lev.faceInternal.append(fint)
        codes = [0, 0, 0, 0]
        gs = [np.zeros(my), np.zeros(my), np.zeros(mx), np.zeros(mx)]
        if not fint[0]:
            codes[0], gs[0] = _bc_side(prob, 0, dom[0], yc, True)
        if not fint[1]:
            codes[1], gs[1] = _bc_side(prob, 1, dom[1], yc, True)
        if not fint[2]:
            codes[2], gs[2] = _bc_side(prob, 2, xc, dom[2], False)
        if not fint[3]:
            codes[3], gs[3] = _bc_side(prob, 3, xc, dom[3], False)
        lev.bc_code.append(codes)
        lev.bc_g.append(gs)
        lev.InvDiagS.append(None)
        refresh_invdiag(lev, p)
        lev.covered.append(np.zeros((mx, my), dtype=bool))
    return lev


def _rect_isect(a, b):
    r = [max(a[0], b[0]), min(a[1], b[1]), max(a[2], b[2]), min(a[3], b[3])]
    if r[0] > r[1] or r[2] > r[3]:
        return None
    return r


def restrict2(A):
    """Conservative 4-point average (ebKernels.m:98-102)."""
    return 0.25 * (A[0:-1:2, 0:-1:2] + A[1::2, 0:-1:2] + A[0:-1:2, 1::2] + A[1::2, 1::2])


class Hierarchy:
    """The ``H`` value (ebSetup.m:10-68): levels 1..nlev, rootIdx, options."""

    def __init__(self, prob: Problem, opts: Options):
        self.prob = prob
        self.opts = opts
        self.lev: list[Level] = []  # lev[k-1] is level k
        self.rootIdx = 0
        self.stats = {"stages": []}
        self.singular = False

    @property
    def nlev(self):
        return len(self.lev)

    def L(self, k) -> Level:
        return self.lev[k - 1]


def coeff_filldown(H: Hierarchy):
    """ebCoeffFilldown.m:1-62: average fine D/C onto covered coarse faces/cells."""
    for k in range(H.nlev, 1, -1):
        levF = H.L(k)
        levC = H.L(k - 1)
        for p in range(levF.np):
            b = levF.box[p]
            Il, Ih, Jl, Jh = (b[0] + 1) // 2, b[1] // 2, (b[2] + 1) // 2, b[3] // 2
            mx, my = levF.dims(p)
            DeW = levF.DeW[p]
            DnS = levF.DnS[p]
            cDeW = 0.5 * (DeW[0:mx + 1:2, 0:my - 1:2] + DeW[0:mx + 1:2, 1:my:2])
            cDnS = 0.5 * (DnS[0:mx - 1:2, 0:my + 1:2] + DnS[1:mx:2, 0:my + 1:2])
            cCC = restrict2(levF.CC[p])
            for q in range(levC.np):
                qb = levC.box[q]
                r = _rect_isect([Il, Ih, Jl, Jh], qb)
                if r is not None:
                    levC.CC[q][r[0] - qb[0]:r[1] - qb[0] + 1, r[2] - qb[2]:r[3] - qb[2] + 1] = \
                        cCC[r[0] - Il:r[1] - Il + 1, r[2] - Jl:r[3] - Jl + 1]
                rf = _rect_isect([Il, Ih + 1, Jl, Jh], [qb[0], qb[1] + 1, qb[2], qb[3]])
                if rf is not None:
                    levC.DeW[q][rf[0] - qb[0]:rf[1] - qb[0] + 1, rf[2] - qb[2]:rf[3] - qb[2] + 1] = \
                        cDeW[rf[0] - Il:rf[1] - Il + 1, rf[2] - Jl:rf[3] - Jl + 1]
                rf = _rect_isect([Il, Ih, Jl, Jh + 1], [qb[0], qb[1], qb[2], qb[3] + 1])
                if rf is not None:
                    levC.DnS[q][rf[0] - qb[0]:rf[1] - qb[0] + 1, rf[2] - qb[2]:rf[3] - qb[2] + 1] = \
                        cDnS[rf[0] - Il:rf[1] - Il + 1, rf[2] - Jl:rf[3] - Jl + 1]
    for k in range(1, H.nlev + 1):
        lev = H.L(k)
        for p in range(lev.np):
            refresh_invdiag(lev, p)


def covered_masks(H: Hierarchy, k: int):
    """covered{q} of level k-1 from the boxes of level k (ebBuildOps.m:171-189)."""
    levF = H.L(k)
    levC = H.L(k - 1)
    for q in range(levC.np):
        qb = levC.box[q]
        cov = np.zeros((qb[1] - qb[0] + 1, qb[3] - qb[2] + 1), dtype=bool)
        for p in range(levF.np):
            b = levF.box[p]
            fp = [(b[0] + 1) // 2, b[1] // 2, (b[2] + 1) // 2, b[3] // 2]
            r = _rect_isect(fp, qb)
            if r is not None:
                cov[r[0] - qb[0]:r[1] - qb[0] + 1, r[2] - qb[2]:r[3] - qb[2] + 1] = True
        levC.covered[q] = cov


def check_alignment(H: Hierarchy, k: int):
    """Alignment rule of ebBuildOps.m:38-42 (checked on the host before upload)."""
    for b in H.L(k).box:
        if b[0] % 2 != 1 or b[1] % 2 != 0 or b[2] % 2 != 1 or b[3] % 2 != 0:
            raise BsamError("eb_build_ops:align", "patch not aligned to parent cells.")


def is_singular(H: Hierarchy) -> bool:
    """ebCoarsest.m:58-66: all-Neumann level 1 and C == 0 everywhere."""
    allneu = all(c != 1 for c in H.L(1).bc_code[0])
    cmax = 0.0
    for lev in H.lev:
        for p in range(lev.np):
            cmax = max(cmax, float(np.max(np.abs(lev.CC[p]))))
    return bool(allneu and cmax == 0)


def compat_project(H: Hierarchy):
    """ebCompatProject.m:1-111: project the source for the singular Neumann case."""
    if not H.singular:
        for k in range(1, H.nlev + 1):
            lev = H.L(k)
            old = lev.sourceShift
            if old == 0:
                continue
            for p in range(lev.np):
                lev.FTrue[p] = lev.FTrue[p] + old
            lev.sourceShift = 0.0
        H.compatibility = {"sourceShift": 0.0, "unprojectedDefect": 0.0}
        return
    S = 0.0
    Av = 0.0
    R = 0.0
    for k in range(H.rootIdx, H.nlev + 1):
        lev = H.L(k)
        old = lev.sourceShift
        hx, hy = lev.h
        da = hx * hy
        for p in range(lev.np):
            cov = lev.covered[p]
            vis = ~cov
            S = S + (np.sum(lev.FTrue[p].T[vis.T]) + old * np.count_nonzero(vis)) * da
            Av = Av + np.count_nonzero(vis) * da
            mx, my = cov.shape
            codes = lev.bc_code[p]
            g = lev.bc_g[p]
            if codes[0] == 2:
                m = ~cov[0, :]
                fl = lev.DeW[p][0, :] * g[0]
                R = R + np.sum(fl[m]) * hy
            if codes[1] == 2:
                m = ~cov[mx - 1, :]
                fl = lev.DeW[p][mx, :] * g[1]
                R = R + np.sum(fl[m]) * hy
            if codes[2] == 2:
                m = ~cov[:, 0]
                fl = lev.DnS[p][:, 0] * g[2]
                R = R + np.sum(fl[m]) * hx
            if codes[3] == 2:
                m = ~cov[:, my - 1]
                fl = lev.DnS[p][:, my] * g[3]
                R = R + np.sum(fl[m]) * hx
    c = (S + R) / max(Av, np.finfo(float).tiny)
    for k in range(1, H.nlev + 1):
        lev = H.L(k)
        old = lev.sourceShift
        for p in range(lev.np):
            lev.FTrue[p] = lev.FTrue[p] - (c - old)
        lev.sourceShift = c
    H.compatibility = {"sourceShift": c, "unprojectedDefect": S + R}


def ladder_sizes(opts: Options):
    """Sub-root size chain (ebSetup.m:14-25)."""
    n0 = np.array(opts.BaseCells, dtype=np.int64)
    sizes = [n0.copy()]
    ns = n0.copy()
    while np.all(ns % 2 == 0) and ns.min() / 2 >= 4 and ns.min() > opts.CoarsestCells:
        ns = ns // 2
        sizes.insert(0, ns.copy())
    return sizes


def setup(prob: Problem, opts: Options) -> Hierarchy:
    """Host part of ebSetup.m:10-68 (ghost fill and the coarsest factorisation
    are done by the solve backend)."""
    H = Hierarchy(prob, opts)
    sizes = ladder_sizes(opts)
    H.rootIdx = len(sizes)
    for k in range(1, len(sizes) + 1):
        n = sizes[k - 1]
        H.lev.append(make_level(prob, opts, k, n, [[1, n[0], 1, n[1]]]))
    coeff_filldown(H)
    for k in range(2, H.nlev + 1):
        check_alignment(H, k)
        covered_masks(H, k)
    ig = opts.InitialGuess
    if ig is not None:
        for k in range(1, H.nlev + 1):
            lev = H.L(k)
            for p in range(lev.np):
                XC, YC = lev.cell_centers(prob.domain, p)
                lev.Q[p][1:-1, 1:-1] = _evalgrid(ig, XC, YC)
    H.singular = is_singular(H)
    compat_project(H)
    return H


def add_level(H: Hierarchy, boxes):
    """Host part of ebAddLevel.m:10-32: append the 2x refinement of the given
    parent-space boxes.  The prolongation initial guess is the backend's."""
    boxes = np.asarray(boxes, dtype=np.int64).reshape(-1, 4)
    kNew = H.nlev + 1
    nOld = H.L(H.nlev).n
    nNew = (nOld[0] * 2, nOld[1] * 2)
    fine = np.stack([2 * boxes[:, 0] - 1, 2 * boxes[:, 1], 2 * boxes[:, 2] - 1, 2 * boxes[:, 3]], axis=1)
    H.lev.append(make_level(H.prob, H.opts, kNew, nNew, fine))
    coeff_filldown(H)
    check_alignment(H, kNew)
    covered_masks(H, kNew)
    H.singular = is_singular(H)
    compat_project(H)
    return kNew


def total_cells(H: Hierarchy, from_root=True):
    k0 = H.rootIdx if from_root else 1
    return sum(int(np.prod(H.L(k).dims(p))) for k in range(k0, H.nlev + 1) for p in range(H.L(k).np))
