"""CPU oracle for the adaptive FAS V-cycle -- TEST INFRASTRUCTURE ONLY.

This module is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it; the product path (the
HIP library and its host layer) never does.

It restates, in plain NumPy FP64 with the reference's operation order (no
fused multiply-add; numpy evaluates elementwise expressions left to right), the
MATLAB functions of the solve path:

  op2d / invdiag2d / relax2d / restrict2 / prolongBilin
                                  EllipticBSAM/src/ebKernels.m:13-123
  fill_ghosts                     EllipticBSAM/src/ebFillGhosts.m:1-95
  interp_cf                       EllipticBSAM/src/ebInterpCf.m:1-154
  relax_level                     EllipticBSAM/src/ebRelaxLevel.m:1-39
  transfer (restrict/ring/correct/filldown)
                                  EllipticBSAM/src/ebTransfer.m:1-84
  gather_window                   EllipticBSAM/src/ebGatherWindow.m:1-19
  coarse_load                     EllipticBSAM/src/ebCoarseLoad.m:1-47
  parent_coeffs                   EllipticBSAM/src/ebParentCoeffs.m:1-31
  face_corrections                EllipticBSAM/src/ebFaceCorrections.m:1-61
  coarsest (assemble/solve)       EllipticBSAM/src/ebCoarsest.m:1-109
  comp_residual                   EllipticBSAM/src/ebCompResidual.m:1-85
  vcycle                          EllipticBSAM/src/ebVcycle.m:1-36
  solve_cycles                    EllipticBSAM/src/ebSolveCycles.m:1-80
  composite_mean                  EllipticBSAM/src/ebCompositeMean.m:1-24
  mass_check                      EllipticBSAM/src/ebMassCheck.m:1-67
  build_ops (tables)              EllipticBSAM/src/ebBuildOps.m:1-288
  gather_parent_coeffs (cfD)      EllipticBSAM/src/ebGatherParentCoeffs.m:1-69
  bdf2 right-hand side            ParabolicBSAM/src/ebpRhs.m:25-48

The coarsest level is solved with an explicit inverse obtained from LU with
partial pivoting (first maximum wins), applied as a dot product accumulated in
ascending column order -- the same arithmetic the device library performs, so
the two agree bit for bit.  MATLAB's own ``decomposition`` is not reproducible
here; the coarsest solve is pinned by its defining equation (residual at
round-off) and by the published V-cycle histories.

Pins (see tests/test_oracle.py): the published stage log of the README
elliptic example (verification/readme-examples.txt:2-6), kernel identities of
EllipticBSAM/tests/unitKernels.m:12-89, the discrete conservation identity
(suite-output.txt:64-68), and brute-force / closed-form checks.  Functions
with no such check are marked "parity unpinned" in their docstring.

Array convention: per-patch arrays ``A[i, j]`` (i along x), 0-based numpy
indices for the reference's 1-based padded indices.
"""
from __future__ import annotations

import copy

import numpy as np

A1, A2, A3 = -1.0 / 5.0, 2.0 / 3.0, 8.0 / 15.0       # ebInterpCf.m:41
B1, B2, B3 = 5.0 / 32.0, 15.0 / 16.0, -3.0 / 32.0     # ebInterpCf.m:42
C83, C13, C23 = 8.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0
REALMIN = np.finfo(float).tiny


# --------------------------------------------------------------------------
# ebKernels.m
def op2d(Q, DeW, DnS, CC, hx, hy):
    """ebKernels.m:13-25 -- flux-form L(u) on the interior."""
    mx, my = CC.shape
    feW = DeW * (Q[1:mx + 2, 1:my + 1] - Q[0:mx + 1, 1:my + 1])
    fnS = DnS * (Q[1:mx + 1, 1:my + 2] - Q[1:mx + 1, 0:my + 1])
    return -(feW[1:mx + 1, :] - feW[0:mx, :]) / (hx * hx) \
        - (fnS[:, 1:my + 1] - fnS[:, 0:my]) / (hy * hy) \
        + CC * Q[1:mx + 1, 1:my + 1]


def invdiag2d(DeW, DnS, CC, hx, hy):
    """ebKernels.m:28-34."""
    mx, my = CC.shape
    den = (DeW[1:mx + 1, :] + DeW[0:mx, :]) / (hx * hx) \
        + (DnS[:, 1:my + 1] + DnS[:, 0:my]) / (hy * hy) + CC
    return 1.0 / den


def colour_mask(mx, my, par0, color):
    """Cells updated by colour ``color`` (1 red / 2 black): mod(li+lj+par0,2)==color-1
    with 1-based local indices (ebKernels.m:80-84)."""
    li = np.arange(1, mx + 1)[:, None]
    lj = np.arange(1, my + 1)[None, :]
    return ((li + lj + par0) % 2) == (color - 1)


def relax2d(Q, F, DeW, DnS, InvDiag, codes, gs, par0, color, hx, hy, omega):
    """ebKernels.m:37-95 -- one coloured GS pass, in place on Q."""
    mx, my = F.shape
    if codes[0] == 1:
        Q[0, 1:my + 1] = C83 * gs[0] + C13 * Q[2, 1:my + 1]
    elif codes[0] == 2:
        Q[0, 1:my + 1] = hx * gs[0]
    if codes[1] == 1:
        Q[mx + 1, 1:my + 1] = C83 * gs[1] + C13 * Q[mx - 1, 1:my + 1]
    elif codes[1] == 2:
        Q[mx + 1, 1:my + 1] = hx * gs[1]
    if codes[2] == 1:
        Q[1:mx + 1, 0] = C83 * gs[2] + C13 * Q[1:mx + 1, 2]
    elif codes[2] == 2:
        Q[1:mx + 1, 0] = hy * gs[2]
    if codes[3] == 1:
        Q[1:mx + 1, my + 1] = C83 * gs[3] + C13 * Q[1:mx + 1, my - 1]
    elif codes[3] == 2:
        Q[1:mx + 1, my + 1] = hy * gs[3]
    hx2 = hx * hx
    hy2 = hy * hy
    tmp = (F
           + Q[2:mx + 2, 1:my + 1] * DeW[1:mx + 1, :] / hx2
           + Q[0:mx, 1:my + 1] * DeW[0:mx, :] / hx2
           + Q[1:mx + 1, 2:my + 2] * DnS[:, 1:my + 1] / hy2
           + Q[1:mx + 1, 0:my] * DnS[:, 0:my] / hy2) * InvDiag
    mask = colour_mask(mx, my, par0, color)
    inter = Q[1:mx + 1, 1:my + 1]
    if omega == 1:
        inter[mask] = tmp[mask]
    else:
        inter[mask] = omega * tmp[mask] + (1 - omega) * inter[mask]
    return Q


def restrict2(A):
    """ebKernels.m:98-102."""
    return 0.25 * (A[0:-1:2, 0:-1:2] + A[1::2, 0:-1:2] + A[0:-1:2, 1::2] + A[1::2, 1::2])


def prolong_bilin(W):
    """ebKernels.m:105-123 -- (9,3,3,1)/16 cell-centred bilinear."""
    cm = W.shape[0] - 2
    cn = W.shape[1] - 2
    C = W[1:cm + 1, 1:cn + 1]
    Wm = W[0:cm, 1:cn + 1]
    Wp = W[2:cm + 2, 1:cn + 1]
    Sm = W[1:cm + 1, 0:cn]
    Sp = W[1:cm + 1, 2:cn + 2]
    Dmm = W[0:cm, 0:cn]
    Dpm = W[2:cm + 2, 0:cn]
    Dmp = W[0:cm, 2:cn + 2]
    Dpp = W[2:cm + 2, 2:cn + 2]
    F = np.zeros((2 * cm, 2 * cn))
    F[0::2, 0::2] = (9 * C + 3 * Wm + 3 * Sm + Dmm) / 16
    F[1::2, 0::2] = (9 * C + 3 * Wp + 3 * Sm + Dpm) / 16
    F[0::2, 1::2] = (9 * C + 3 * Wm + 3 * Sp + Dmp) / 16
    F[1::2, 1::2] = (9 * C + 3 * Wp + 3 * Sp + Dpp) / 16
    return F


# --------------------------------------------------------------------------
# ebInterpCf.m
def interp_cf(Q, QC, fint, cCF, qif):
    """ebInterpCf.m:36-154 -- QIF/LIF ghosts from the QC window (in place)."""
    mx = Q.shape[0] - 2
    my = Q.shape[1] - 2
    cmx = mx // 2
    cmy = my // 2
    a1, a2, a3, b1, b2, b3 = A1, A2, A3, B1, B2, B3
    if fint[0]:  # WEST
        if qif:
            jj = np.arange(2, my - 1, 2)
            Q[0, jj] = a1 * Q[2, jj + 2] + a2 * Q[1, jj + 1] + a3 * QC[0, jj // 2]
            jj = np.arange(3, my, 2)
            Q[0, jj] = a1 * Q[2, jj - 2] + a2 * Q[1, jj - 1] + a3 * QC[0, (jj + 1) // 2]
            Q[0, my] = a1 * Q[2, my] + a2 * Q[1, my] + a3 * (b1 * QC[0, cmy + 1] + b2 * QC[0, cmy] + b3 * QC[0, cmy - 1])
            Q[0, 1] = a1 * Q[2, 1] + a2 * Q[1, 1] + a3 * (b1 * QC[0, 0] + b2 * QC[0, 1] + b3 * QC[0, 2])
        else:
            jj = np.arange(2, my + 1, 2)
            Q[0, jj] = C23 * QC[0, jj // 2] + C23 * Q[1, jj] - C13 * Q[1, jj - 1]
            jj = np.arange(1, my, 2)
            Q[0, jj] = C23 * QC[0, (jj + 1) // 2] + C23 * Q[1, jj] - C13 * Q[1, jj + 1]
    if fint[1]:  # EAST
        if qif:
            jj = np.arange(2, my - 1, 2)
            Q[mx + 1, jj] = a1 * Q[mx - 1, jj + 2] + a2 * Q[mx, jj + 1] + a3 * QC[cmx + 1, jj // 2]
            jj = np.arange(3, my, 2)
            Q[mx + 1, jj] = a1 * Q[mx - 1, jj - 2] + a2 * Q[mx, jj - 1] + a3 * QC[cmx + 1, (jj + 1) // 2]
            Q[mx + 1, my] = a1 * Q[mx - 1, my] + a2 * Q[mx, my] + a3 * (
                b1 * QC[cmx + 1, cmy + 1] + b2 * QC[cmx + 1, cmy] + b3 * QC[cmx + 1, cmy - 1])
            Q[mx + 1, 1] = a1 * Q[mx - 1, 1] + a2 * Q[mx, 1] + a3 * (
                b1 * QC[cmx + 1, 0] + b2 * QC[cmx + 1, 1] + b3 * QC[cmx + 1, 2])
        else:
            jj = np.arange(2, my + 1, 2)
            Q[mx + 1, jj] = C23 * QC[cmx + 1, jj // 2] + C23 * Q[mx, jj] - C13 * Q[mx, jj - 1]
            jj = np.arange(1, my, 2)
            Q[mx + 1, jj] = C23 * QC[cmx + 1, (jj + 1) // 2] + C23 * Q[mx, jj] - C13 * Q[mx, jj + 1]
    if fint[2]:  # SOUTH
        if qif:
            ii = np.arange(2, mx - 1, 2)
            Q[ii, 0] = a1 * Q[ii + 2, 2] + a2 * Q[ii + 1, 1] + a3 * QC[ii // 2, 0]
            ii = np.arange(3, mx, 2)
            Q[ii, 0] = a1 * Q[ii - 2, 2] + a2 * Q[ii - 1, 1] + a3 * QC[(ii + 1) // 2, 0]
            Q[mx, 0] = a1 * Q[mx, 2] + a2 * Q[mx, 1] + a3 * (b1 * QC[cmx + 1, 0] + b2 * QC[cmx, 0] + b3 * QC[cmx - 1, 0])
            Q[1, 0] = a1 * Q[1, 2] + a2 * Q[1, 1] + a3 * (b1 * QC[0, 0] + b2 * QC[1, 0] + b3 * QC[2, 0])
        else:
            ii = np.arange(2, mx + 1, 2)
            Q[ii, 0] = C23 * QC[ii // 2, 0] + C23 * Q[ii, 1] - C13 * Q[ii - 1, 1]
            ii = np.arange(1, mx, 2)
            Q[ii, 0] = C23 * QC[(ii + 1) // 2, 0] + C23 * Q[ii, 1] - C13 * Q[ii + 1, 1]
    if fint[3]:  # NORTH
        if qif:
            ii = np.arange(2, mx - 1, 2)
            Q[ii, my + 1] = a1 * Q[ii + 2, my - 1] + a2 * Q[ii + 1, my] + a3 * QC[ii // 2, cmy + 1]
            ii = np.arange(3, mx, 2)
            Q[ii, my + 1] = a1 * Q[ii - 2, my - 1] + a2 * Q[ii - 1, my] + a3 * QC[(ii + 1) // 2, cmy + 1]
            Q[mx, my + 1] = a1 * Q[mx, my - 1] + a2 * Q[mx, my] + a3 * (
                b1 * QC[cmx + 1, cmy + 1] + b2 * QC[cmx, cmy + 1] + b3 * QC[cmx - 1, cmy + 1])
            Q[1, my + 1] = a1 * Q[1, my - 1] + a2 * Q[1, my] + a3 * (
                b1 * QC[0, cmy + 1] + b2 * QC[1, cmy + 1] + b3 * QC[2, cmy + 1])
        else:
            ii = np.arange(2, mx + 1, 2)
            Q[ii, my + 1] = C23 * QC[ii // 2, cmy + 1] + C23 * Q[ii, my] - C13 * Q[ii - 1, my]
            ii = np.arange(1, mx, 2)
            Q[ii, my + 1] = C23 * QC[(ii + 1) // 2, cmy + 1] + C23 * Q[ii, my] - C13 * Q[ii + 1, my]
    # corners (ebInterpCf.m:123-153)
    if cCF[0]:
        Q[0, 0] = (a1 * Q[2, 2] + a2 * Q[1, 1] + a3 * QC[0, 0]) if qif else (C23 * QC[0, 0] + C13 * Q[1, 1])
    if cCF[1]:
        Q[mx + 1, 0] = (a1 * Q[mx - 1, 2] + a2 * Q[mx, 1] + a3 * QC[cmx + 1, 0]) if qif \
            else (C23 * QC[cmx + 1, 0] + C13 * Q[mx, 1])
    if cCF[2]:
        Q[0, my + 1] = (a1 * Q[2, my - 1] + a2 * Q[1, my] + a3 * QC[0, cmy + 1]) if qif \
            else (C23 * QC[0, cmy + 1] + C13 * Q[1, my])
    if cCF[3]:
        Q[mx + 1, my + 1] = (a1 * Q[mx - 1, my - 1] + a2 * Q[mx, my] + a3 * QC[cmx + 1, cmy + 1]) if qif \
            else (C23 * QC[cmx + 1, cmy + 1] + C13 * Q[mx, my])
    return Q


# --------------------------------------------------------------------------
def _rect_isect(a, b):
    r = [max(a[0], b[0]), min(a[1], b[1]), max(a[2], b[2]), min(a[3], b[3])]
    if r[0] > r[1] or r[2] > r[3]:
        return None
    return r


class _Lev:
    pass


class OracleHierarchy:
    """Self-contained copy of a host hierarchy with the reference's
    descriptors (built by ``build_ops``) and the V-cycle machinery."""

    def __init__(self, H, assemble=True):
        self.prob = H.prob
        self.opts = H.opts
        self.rootIdx = H.rootIdx
        self.lev = []
        for k in range(1, H.nlev + 1):
            src = H.L(k)
            L = _Lev()
            L.n = tuple(src.n)
            L.h = tuple(src.h)
            L.np = src.np
            L.box = np.array(src.box, dtype=np.int64)
            L.Q = [a.copy() for a in src.Q]
            L.FTrue = [a.copy() for a in src.FTrue]
            L.F = [a.copy() for a in src.FTrue]
            L.DeW = [a.copy() for a in src.DeW]
            L.DnS = [a.copy() for a in src.DnS]
            L.CC = [a.copy() for a in src.CC]
            L.InvDiagS = [a.copy() for a in src.InvDiagS]
            L.bc_code = [list(c) for c in src.bc_code]
            L.bc_g = [[g.copy() for g in gs] for gs in src.bc_g]
            L.par0 = list(src.par0)
            L.faceInternal = [list(f) for f in src.faceInternal]
            L.covered = [c.copy() for c in src.covered]
            L.cornerClass = [[1, 1, 1, 1] for _ in range(L.np)]
            L.land = [None] * L.np
            L.sameOps = [[] for _ in range(L.np)]
            L.ops = [None] * L.np
            if k > 1:
                L.QC = []
                for p in range(L.np):
                    mx, my = src.dims(p)
                    L.QC.append(np.zeros((mx // 2 + 2, my // 2 + 2)))
            self.lev.append(L)
        for k in range(2, self.nlev + 1):
            self.build_ops(k)
        self.gather_parent_coeffs()
        self.stats = {"stages": []}
        if assemble:
            self.coarsest_assemble()

    @property
    def nlev(self):
        return len(self.lev)

    def L(self, k):
        return self.lev[k - 1]

    # ---------------------------------------------------------------- tables
    def build_ops(self, k):
        """ebBuildOps.m:17-169 (descriptors of level k against level k-1)."""
        levF = self.L(k)
        levC = self.L(k - 1)
        nF = levF.n
        nC = levC.n
        OF = np.zeros(nF, dtype=np.int64)  # 1-based patch owner, 0 = none
        for p in range(levF.np):
            b = levF.box[p]
            OF[b[0] - 1:b[1], b[2] - 1:b[3]] = p + 1
        OC = np.zeros(nC, dtype=np.int64)
        for q in range(levC.np):
            b = levC.box[q]
            OC[b[0] - 1:b[1], b[2] - 1:b[3]] = q + 1
        for p in range(levF.np):
            b = levF.box[p]
            ilo, ihi, jlo, jhi = (int(v) for v in b)
            mx = ihi - ilo + 1
            my = jhi - jlo + 1
            if ilo % 2 != 1 or ihi % 2 != 0 or jlo % 2 != 1 or jhi % 2 != 0:
                raise ValueError("eb_build_ops:align: patch not aligned to parent cells.")
            Il, Ih, Jl, Jh = (ilo + 1) // 2, ihi // 2, (jlo + 1) // 2, jhi // 2
            cmx = Ih - Il + 1
            cmy = Jh - Jl + 1
            land = {"W": np.ones(my, np.int8), "E": np.ones(my, np.int8),
                    "S": np.ones(mx, np.int8), "N": np.ones(mx, np.int8)}
            same = []   # (sp, src_i, src_j, dst_i, dst_j) index arrays, 0-based padded

            def add_face(nb, gi, gj, dst_i, dst_j):
                idx = np.flatnonzero(nb > 0)
                if idx.size == 0:
                    return
                gi = np.broadcast_to(gi, nb.shape)
                gj = np.broadcast_to(gj, nb.shape)
                for s in np.unique(nb[idx]):
                    ii = idx[nb[idx] == s]
                    sb = levF.box[s - 1]
                    same.append((int(s) - 1, gi[ii] - sb[0] + 1, gj[ii] - sb[2] + 1,
                                 np.broadcast_to(dst_i, nb.shape)[ii], np.broadcast_to(dst_j, nb.shape)[ii]))

            gj = np.arange(jlo, jhi + 1)
            gi = np.arange(ilo, ihi + 1)
            if ilo > 1:
                nb = OF[ilo - 2, jlo - 1:jhi]
                land["W"][nb > 0] = 2
                land["W"][nb == 0] = 3
                add_face(nb, ilo - 1, gj, 0, np.arange(1, my + 1))
            if ihi < nF[0]:
                nb = OF[ihi, jlo - 1:jhi]
                land["E"][nb > 0] = 2
                land["E"][nb == 0] = 3
                add_face(nb, ihi + 1, gj, mx + 1, np.arange(1, my + 1))
            if jlo > 1:
                nb = OF[ilo - 1:ihi, jlo - 2]
                land["S"][nb > 0] = 2
                land["S"][nb == 0] = 3
                add_face(nb, gi, jlo - 1, np.arange(1, mx + 1), 0)
            if jhi < nF[1]:
                nb = OF[ilo - 1:ihi, jhi]
                land["N"][nb > 0] = 2
                land["N"][nb == 0] = 3
                add_face(nb, gi, jhi + 1, np.arange(1, mx + 1), my + 1)
            cdef = [(ilo - 1, jlo - 1, 0, 0), (ihi + 1, jlo - 1, mx + 1, 0),
                    (ilo - 1, jhi + 1, 0, my + 1), (ihi + 1, jhi + 1, mx + 1, my + 1)]
            cclass = [1, 1, 1, 1]
            for c, (cgi, cgj, di, dj) in enumerate(cdef):
                if cgi < 1 or cgi > nF[0] or cgj < 1 or cgj > nF[1]:
                    cclass[c] = 1
                else:
                    s = OF[cgi - 1, cgj - 1]
                    if s > 0:
                        cclass[c] = 2
                        sb = levF.box[s - 1]
                        same.append((int(s) - 1, np.array([cgi - sb[0] + 1]), np.array([cgj - sb[2] + 1]),
                                     np.array([di]), np.array([dj])))
                    else:
                        cclass[c] = 3
            levF.land[p] = land
            levF.cornerClass[p] = cclass
            levF.sameOps[p] = same

            ops = {}
            # ring: west column, east column (incl. corners), then south/north rows
            ring_gi = np.concatenate([np.full(cmy + 2, Il - 1), np.full(cmy + 2, Ih + 1),
                                      np.arange(Il, Ih + 1), np.arange(Il, Ih + 1)])
            ring_gj = np.concatenate([np.arange(Jl - 1, Jh + 2), np.arange(Jl - 1, Jh + 2),
                                      np.full(cmx, Jl - 1), np.full(cmx, Jh + 1)])
            ops["ring"] = self._resolve_cells(levC, OC, ring_gi, ring_gj, Il, Jl, cmx, cmy)
            ops["winrect"] = []
            ncov = sum(r[1].size for r in ops["ring"])
            for q in range(levC.np):
                qb = levC.box[q]
                r = _rect_isect([Il, Ih, Jl, Jh], qb)
                if r is None:
                    continue
                # (q, src rect 0-based padded [i0,i1,j0,j1], dst rect 0-based window)
                ops["winrect"].append((q, (r[0] - qb[0] + 1, r[1] - qb[0] + 1, r[2] - qb[2] + 1, r[3] - qb[2] + 1),
                                       (r[0] - Il + 1, r[1] - Il + 1, r[2] - Jl + 1, r[3] - Jl + 1)))
                ncov += (r[1] - r[0] + 1) * (r[3] - r[2] + 1)
            if ncov != (cmx + 2) * (cmy + 2):
                raise ValueError("eb_build_ops:nesting: parent window not fully covered")
            ops["usc"] = []
            ops["fsc"] = []
            for q in range(levC.np):
                qb = levC.box[q]
                r = _rect_isect([Il, Ih, Jl, Jh], qb)
                if r is None:
                    continue
                sr = (r[0] - Il, r[1] - Il, r[2] - Jl, r[3] - Jl)
                ops["usc"].append((q, sr, (r[0] - qb[0] + 1, r[1] - qb[0] + 1, r[2] - qb[2] + 1, r[3] - qb[2] + 1)))
                ops["fsc"].append((q, sr, (r[0] - qb[0], r[1] - qb[0], r[2] - qb[2], r[3] - qb[2])))
            ops["corr"] = [
                self._corr_face(levC, OC, land["W"][1::2] == 3, np.full(cmy, Il - 1), np.arange(Jl, Jh + 1)),
                self._corr_face(levC, OC, land["E"][1::2] == 3, np.full(cmy, Ih + 1), np.arange(Jl, Jh + 1)),
                self._corr_face(levC, OC, land["S"][1::2] == 3, np.arange(Il, Ih + 1), np.full(cmx, Jl - 1)),
                self._corr_face(levC, OC, land["N"][1::2] == 3, np.arange(Il, Ih + 1), np.full(cmx, Jh + 1)),
            ]
            levF.ops[p] = ops
        # covered masks of the parent level (ebBuildOps.m:174-189)
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

    @staticmethod
    def _resolve_cells(levC, OC, gis, gjs, Il, Jl, cmx, cmy):
        """ebBuildOps.m:221-252 -> list of (q, src_i, src_j, dst_i, dst_j) grouped by q."""
        nC = levC.n
        sp = np.zeros(gis.size, np.int64)
        si = np.zeros(gis.size, np.int64)
        sj = np.zeros(gis.size, np.int64)
        for t in range(gis.size):
            gi, gj = int(gis[t]), int(gjs[t])
            gic = min(max(gi, 1), nC[0])
            gjc = min(max(gj, 1), nC[1])
            q = OC[gic - 1, gjc - 1]
            if q == 0:
                raise ValueError(f"eb_build_ops:ring: parent cell ({gi},{gj}) not covered")
            qb = levC.box[q - 1]
            li = gi - qb[0] + 1
            lj = gj - qb[2] + 1
            mqx = qb[1] - qb[0] + 1
            mqy = qb[3] - qb[2] + 1
            if li < 0 or li > mqx + 1 or lj < 0 or lj > mqy + 1:
                raise ValueError("eb_build_ops:overhang")
            sp[t] = q - 1
            si[t] = li
            sj[t] = lj
        di = gis - Il + 1
        dj = gjs - Jl + 1
        out = []
        for q in np.unique(sp):
            ii = sp == q
            out.append((int(q), si[ii], sj[ii], di[ii], dj[ii]))
        return out

    @staticmethod
    def _corr_face(levC, OC, mask, gis, gjs):
        """ebBuildOps.m:255-281 -> list of (dp, src_j (0-based), dst_i, dst_j) grouped by dp."""
        idx = np.flatnonzero(mask)
        if idx.size == 0:
            return []
        dp = np.zeros(idx.size, np.int64)
        di = np.zeros(idx.size, np.int64)
        dj = np.zeros(idx.size, np.int64)
        for t, j in enumerate(idx):
            q = OC[gis[j] - 1, gjs[j] - 1]
            if q == 0:
                raise ValueError("eb_build_ops:corr: mass-correction target cell uncovered.")
            qb = levC.box[q - 1]
            dp[t] = q - 1
            di[t] = gis[j] - qb[0]
            dj[t] = gjs[j] - qb[2]
        out = []
        for q in np.unique(dp):
            ii = dp == q
            out.append((int(q), idx[ii], di[ii], dj[ii]))
        return out

    def gather_parent_coeffs(self):
        """ebGatherParentCoeffs.m:1-69 (rect descriptors + cfD vectors)."""
        for k in range(2, self.nlev + 1):
            levF = self.L(k)
            levC = self.L(k - 1)
            for p in range(levF.np):
                b = levF.box[p]
                Il, Ih, Jl, Jh = (b[0] + 1) // 2, b[1] // 2, (b[2] + 1) // 2, b[3] // 2
                cmx = Ih - Il + 1
                cmy = Jh - Jl + 1
                pcc, pdew, pdns = [], [], []
                for q in range(levC.np):
                    qb = levC.box[q]
                    r = _rect_isect([Il, Ih, Jl, Jh], qb)
                    if r is not None:
                        pcc.append((q, (r[0] - qb[0], r[1] - qb[0], r[2] - qb[2], r[3] - qb[2]),
                                    (r[0] - Il, r[1] - Il, r[2] - Jl, r[3] - Jl)))
                    rf = _rect_isect([Il, Ih + 1, Jl, Jh], [qb[0], qb[1] + 1, qb[2], qb[3]])
                    if rf is not None:
                        pdew.append((q, (rf[0] - qb[0], rf[1] - qb[0], rf[2] - qb[2], rf[3] - qb[2]),
                                     (rf[0] - Il, rf[1] - Il, rf[2] - Jl, rf[3] - Jl)))
                    rf = _rect_isect([Il, Ih, Jl, Jh + 1], [qb[0], qb[1], qb[2], qb[3] + 1])
                    if rf is not None:
                        pdns.append((q, (rf[0] - qb[0], rf[1] - qb[0], rf[2] - qb[2], rf[3] - qb[2]),
                                     (rf[0] - Il, rf[1] - Il, rf[2] - Jl, rf[3] - Jl)))
                ops = levF.ops[p]
                ops["pcc"], ops["pdew"], ops["pdns"] = pcc, pdew, pdns
                PDeW, PDnS, _ = self.parent_coeffs(k, p)
                if np.isnan(PDeW).any() or np.isnan(PDnS).any():
                    raise ValueError("eb_gather_parent_coeffs:cover")
                ops["cfD"] = (PDeW[0, :].copy(), PDeW[cmx, :].copy(), PDnS[:, 0].copy(), PDnS[:, cmy].copy())

    def parent_coeffs(self, k, p):
        """ebParentCoeffs.m:1-31."""
        b = self.L(k).box[p]
        cmx = (b[1] - b[0] + 1) // 2
        cmy = (b[3] - b[2] + 1) // 2
        ops = self.L(k).ops[p]
        levC = self.L(k - 1)
        PDeW = np.full((cmx + 1, cmy), np.nan)
        PDnS = np.full((cmx, cmy + 1), np.nan)
        PCC = np.full((cmx, cmy), np.nan)
        for q, s, d in ops["pdew"]:
            PDeW[d[0]:d[1] + 1, d[2]:d[3] + 1] = levC.DeW[q][s[0]:s[1] + 1, s[2]:s[3] + 1]
        for q, s, d in ops["pdns"]:
            PDnS[d[0]:d[1] + 1, d[2]:d[3] + 1] = levC.DnS[q][s[0]:s[1] + 1, s[2]:s[3] + 1]
        for q, s, d in ops["pcc"]:
            PCC[d[0]:d[1] + 1, d[2]:d[3] + 1] = levC.CC[q][s[0]:s[1] + 1, s[2]:s[3] + 1]
        return PDeW, PDnS, PCC

    # ------------------------------------------------------------ ghosts
    def fill_ghosts(self, k):
        """ebFillGhosts.m:1-95."""
        lev = self.L(k)
        qif = self.opts.GhostInterp == "qif"
        hx, hy = lev.h
        if k > 1:
            Qpar = self.L(k - 1).Q
            for p in range(lev.np):
                fint = lev.faceInternal[p]
                cCF = [c == 3 for c in lev.cornerClass[p]]
                if not any(fint) and not any(cCF):
                    continue
                QC = lev.QC[p]
                for q, si, sj, di, dj in lev.ops[p]["ring"]:
                    QC[di, dj] = Qpar[q][si, sj]
                interp_cf(lev.Q[p], QC, fint, cCF, qif)
        for p in range(lev.np):
            for sp, si, sj, di, dj in lev.sameOps[p]:
                lev.Q[p][di, dj] = lev.Q[sp][si, sj]
        for p in range(lev.np):
            codes = lev.bc_code[p]
            cc = lev.cornerClass[p]
            if all(c == 0 for c in codes) and all(c != 1 for c in cc):
                continue
            Q = lev.Q[p]
            g = lev.bc_g[p]
            mx = Q.shape[0] - 2
            my = Q.shape[1] - 2
            if codes[0] == 1:
                Q[0, 1:my + 1] = C83 * g[0] - 2 * Q[1, 1:my + 1] + C13 * Q[2, 1:my + 1]
            elif codes[0] == 2:
                Q[0, 1:my + 1] = Q[1, 1:my + 1] + hx * g[0]
            if codes[1] == 1:
                Q[mx + 1, 1:my + 1] = C83 * g[1] - 2 * Q[mx, 1:my + 1] + C13 * Q[mx - 1, 1:my + 1]
            elif codes[1] == 2:
                Q[mx + 1, 1:my + 1] = Q[mx, 1:my + 1] + hx * g[1]
            if codes[2] == 1:
                Q[1:mx + 1, 0] = C83 * g[2] - 2 * Q[1:mx + 1, 1] + C13 * Q[1:mx + 1, 2]
            elif codes[2] == 2:
                Q[1:mx + 1, 0] = Q[1:mx + 1, 1] + hy * g[2]
            if codes[3] == 1:
                Q[1:mx + 1, my + 1] = C83 * g[3] - 2 * Q[1:mx + 1, my] + C13 * Q[1:mx + 1, my - 1]
            elif codes[3] == 2:
                Q[1:mx + 1, my + 1] = Q[1:mx + 1, my] + hy * g[3]
            if cc[0] == 1:
                Q[0, 0] = Q[1, 0] + Q[0, 1] - Q[1, 1]
            if cc[1] == 1:
                Q[mx + 1, 0] = Q[mx, 0] + Q[mx + 1, 1] - Q[mx, 1]
            if cc[2] == 1:
                Q[0, my + 1] = Q[1, my + 1] + Q[0, my] - Q[1, my]
            if cc[3] == 1:
                Q[mx + 1, my + 1] = Q[mx, my + 1] + Q[mx + 1, my] - Q[mx, my]

    # ------------------------------------------------------------ smoother
    def relax_level(self, k, nsweeps):
        """ebRelaxLevel.m:28-38."""
        lev = self.L(k)
        om = self.opts.Omega
        hx, hy = lev.h
        for _ in range(nsweeps):
            for color in (1, 2):
                for p in range(lev.np):
                    relax2d(lev.Q[p], lev.F[p], lev.DeW[p], lev.DnS[p], lev.InvDiagS[p],
                            lev.bc_code[p], lev.bc_g[p], lev.par0[p], color, hx, hy, om)
                self.fill_ghosts(k)

    # ------------------------------------------------------------ transfers
    def gather_window(self, k, p):
        """ebGatherWindow.m:7-19."""
        ops = self.L(k).ops[p]
        b = self.L(k).box[p]
        W = np.zeros(((b[1] - b[0] + 1) // 2 + 2, (b[3] - b[2] + 1) // 2 + 2))
        Qpar = self.L(k - 1).Q
        for q, si, sj, di, dj in ops["ring"]:
            W[di, dj] = Qpar[q][si, sj]
        for q, s, d in ops["winrect"]:
            W[d[0]:d[1] + 1, d[2]:d[3] + 1] = Qpar[q][s[0]:s[1] + 1, s[2]:s[3] + 1]
        return W

    def transfer(self, k, what):
        """ebTransfer.m:10-79."""
        if what == "restrict":
            lev = self.L(k)
            Qpar = self.L(k - 1).Q
            for p in range(lev.np):
                R = restrict2(lev.Q[p][1:-1, 1:-1])
                lev.QC[p][1:-1, 1:-1] = R
                for q, s, d in lev.ops[p]["usc"]:
                    Qpar[q][d[0]:d[1] + 1, d[2]:d[3] + 1] = R[s[0]:s[1] + 1, s[2]:s[3] + 1]
        elif what == "ring":
            lev = self.L(k)
            Qpar = self.L(k - 1).Q
            for p in range(lev.np):
                QC = lev.QC[p]
                for q, si, sj, di, dj in lev.ops[p]["ring"]:
                    QC[di, dj] = Qpar[q][si, sj]
        elif what == "correct":
            lev = self.L(k)
            for p in range(lev.np):
                W = self.gather_window(k, p)
                E = W - lev.QC[p]
                lev.Q[p][1:-1, 1:-1] = lev.Q[p][1:-1, 1:-1] + prolong_bilin(E)
        elif what == "filldown":
            for kk in range(self.nlev, 1, -1):
                self.transfer(kk, "restrict")
                self.fill_ghosts(kk - 1)
        else:
            raise ValueError(f"eb_transfer:what: unknown transfer {what}")

    # ------------------------------------------------------------ FAS load
    def face_corrections(self, k, p):
        """ebFaceCorrections.m:29-55 -> [vals_W, vals_E, vals_S, vals_N]."""
        lev = self.L(k)
        Q = lev.Q[p]
        QC = lev.QC[p]
        DeW = lev.DeW[p]
        DnS = lev.DnS[p]
        cfW, cfE, cfS, cfN = lev.ops[p]["cfD"]
        mx = Q.shape[0] - 2
        my = Q.shape[1] - 2
        cmx = mx // 2
        cmy = my // 2
        hcx2 = self.L(k - 1).h[0] ** 2
        hcy2 = self.L(k - 1).h[1] ** 2
        j = np.arange(1, cmy + 1)
        vW = (DeW[0, 2 * j - 2] * (Q[1, 2 * j - 1] - Q[0, 2 * j - 1])
              + DeW[0, 2 * j - 1] * (Q[1, 2 * j] - Q[0, 2 * j])
              + cfW * (QC[0, j] - QC[1, j])) / hcx2
        vE = (DeW[mx, 2 * j - 2] * (Q[mx, 2 * j - 1] - Q[mx + 1, 2 * j - 1])
              + DeW[mx, 2 * j - 1] * (Q[mx, 2 * j] - Q[mx + 1, 2 * j])
              + cfE * (QC[cmx + 1, j] - QC[cmx, j])) / hcx2
        i = np.arange(1, cmx + 1)
        vS = (DnS[2 * i - 2, 0] * (Q[2 * i - 1, 1] - Q[2 * i - 1, 0])
              + DnS[2 * i - 1, 0] * (Q[2 * i, 1] - Q[2 * i, 0])
              + cfS * (QC[i, 0] - QC[i, 1])) / hcy2
        vN = (DnS[2 * i - 2, my] * (Q[2 * i - 1, my] - Q[2 * i - 1, my + 1])
              + DnS[2 * i - 1, my] * (Q[2 * i, my] - Q[2 * i, my + 1])
              + cfN * (QC[i, cmy + 1] - QC[i, cmy])) / hcy2
        return [vW, vE, vS, vN]

    def coarse_load(self, k):
        """ebCoarseLoad.m:17-46."""
        lev = self.L(k)
        hx, hy = lev.h
        hcx, hcy = self.L(k - 1).h
        Fpar = self.L(k - 1).F
        for p in range(lev.np):
            RF = lev.F[p] - op2d(lev.Q[p], lev.DeW[p], lev.DnS[p], lev.CC[p], hx, hy)
            PDeW, PDnS, PCC = self.parent_coeffs(k, p)
            loadC = restrict2(RF) + op2d(lev.QC[p], PDeW, PDnS, PCC, hcx, hcy)
            for q, s, d in lev.ops[p]["fsc"]:
                Fpar[q][d[0]:d[1] + 1, d[2]:d[3] + 1] = loadC[s[0]:s[1] + 1, s[2]:s[3] + 1]
            if not self.opts.MassCorrection:
                continue
            vals = self.face_corrections(k, p)
            for f in range(4):
                for q, src, di, dj in lev.ops[p]["corr"][f]:
                    Fpar[q][di, dj] = Fpar[q][di, dj] + vals[f][src]

    # ------------------------------------------------------------ coarsest
    def coarsest_assemble(self):
        """ebCoarsest.m:14-77 -> explicit inverse (LU with partial pivoting)."""
        lev = self.L(1)
        n1, n2 = lev.n
        N = n1 * n2
        hx, hy = lev.h
        hx2 = hx * hx
        hy2 = hy * hy
        DeW, DnS, CC = lev.DeW[0], lev.DnS[0], lev.CC[0]
        codes, gs = lev.bc_code[0], lev.bc_g[0]
        A, bcRhs = assemble_coarsest(n1, n2, hx, hy, DeW, DnS, CC, codes, gs)
        allneu = all(c != 1 for c in codes)
        cmax = 0.0
        for L in self.lev:
            for p in range(L.np):
                cmax = max(cmax, float(np.max(np.abs(L.CC[p]))))
        singular = allneu and cmax == 0
        if singular:
            M = np.zeros((N + 1, N + 1))
            M[:N, :N] = A
            M[:N, N] = 1.0
            M[N, :N] = 1.0
        else:
            M = A
        self.coarsest = {"Ainv": lu_inverse(M), "singular": singular, "bcRhs": bcRhs, "A": A}

    def coarsest_solve(self):
        """ebCoarsest.m:79-91 (dot products in ascending column order)."""
        lev = self.L(1)
        n1, n2 = lev.n
        N = n1 * n2
        rhs = lev.F[0].reshape(-1, order="F") + self.coarsest["bcRhs"]
        if self.coarsest["singular"]:
            rhs = np.concatenate([rhs, [0.0]])
        u = dense_apply(self.coarsest["Ainv"], rhs, N)
        lev.Q[0][1:-1, 1:-1] = u.reshape((n1, n2), order="F")
        self.fill_ghosts(1)

    # ------------------------------------------------------------ cycle
    def vcycle(self, k):
        """ebVcycle.m:15-35."""
        if k == 1:
            self.coarsest_solve()
            return
        self.relax_level(k, self.opts.NPre)
        self.transfer(k, "restrict")
        self.fill_ghosts(k - 1)
        self.transfer(k, "ring")
        self.coarse_load(k)
        self.vcycle(k - 1)
        self.transfer(k, "correct")
        self.fill_ghosts(k)
        self.relax_level(k, self.opts.NPost)

    def comp_residual(self):
        """ebCompResidual.m:23-84."""
        r0 = self.rootIdx
        chunks = {k: [[] for _ in range(self.L(k).np)] for k in range(r0, self.nlev + 1)}
        if self.opts.MassCorrection:
            for k in range(max(2, r0 + 1), self.nlev + 1):
                self.transfer(k, "ring")
                lev = self.L(k)
                for p in range(lev.np):
                    lev.QC[p][1:-1, 1:-1] = restrict2(lev.Q[p][1:-1, 1:-1])
            for k in range(r0 + 1, self.nlev + 1):
                lev = self.L(k)
                for p in range(lev.np):
                    vals = self.face_corrections(k, p)
                    for f in range(4):
                        for q, src, di, dj in lev.ops[p]["corr"][f]:
                            chunks[k - 1][q].append((di, dj, vals[f][src]))
        linf = 0.0
        l2sq = 0.0
        fnorm = 0.0
        ncell = 0
        for k in range(r0, self.nlev + 1):
            lev = self.L(k)
            area = lev.h[0] * lev.h[1]
            for p in range(lev.np):
                r = lev.FTrue[p] - op2d(lev.Q[p], lev.DeW[p], lev.DnS[p], lev.CC[p], lev.h[0], lev.h[1])
                for di, dj, v in chunks[k][p]:
                    r[di, dj] = r[di, dj] + v
                vis = ~lev.covered[p]
                e = r.T[vis.T]
                if e.size:
                    linf = max(linf, float(np.max(np.abs(e))))
                    l2sq = l2sq + float(np.sum(e ** 2)) * area
                    fnorm = max(fnorm, float(np.max(np.abs(lev.FTrue[p].T[vis.T]))))
                    ncell += e.size
        fnorm = max(fnorm, REALMIN)
        return {"linf": linf, "l2": np.sqrt(l2sq), "fnorm": fnorm, "rel": linf / fnorm, "ncell": ncell}

    def solve_cycles(self, max_cycles=None, verbose=0):
        """ebSolveCycles.m:8-79."""
        opts = self.opts
        maxv = opts.MaxVCycles if max_cycles is None else max_cycles
        for k in range(1, self.nlev + 1):
            self.fill_ghosts(k)
        res = self.comp_residual()
        hist = [res["linf"]]
        rel = res["rel"]
        it = 0
        while it < maxv and rel > opts.RelTol:
            it += 1
            for k in range(1, self.nlev + 1):
                lev = self.L(k)
                for p in range(lev.np):
                    lev.F[p] = lev.FTrue[p].copy()
            self.vcycle(self.nlev)
            self.transfer(0, "filldown")
            res = self.comp_residual()
            hist.append(res["linf"])
            rel = res["rel"]
            if verbose:
                print(f"   it {it:2d}   |r|_inf = {res['linf']:.6e}   rel = {rel:.3e}")
        ncell_tot = sum(self.L(k).FTrue[p].size for k in range(self.rootIdx, self.nlev + 1)
                        for p in range(self.L(k).np))
        stage = {"amrLevels": self.nlev - self.rootIdx + 1, "nlev": self.nlev, "vcycles": it,
                 "resHist": hist, "finalRel": rel, "visibleCells": res["ncell"], "totalCells": ncell_tot,
                 "avgFactor": (hist[-1] / max(hist[0], REALMIN)) ** (1.0 / (len(hist) - 1))
                 if len(hist) >= 2 else float("nan")}
        self.stats["stages"].append(stage)
        return stage

    def init_level_prolong(self, k):
        """ebAddLevel.m:25-32: fill parent ghosts, prolong the window, fill."""
        self.fill_ghosts(k - 1)
        lev = self.L(k)
        for p in range(lev.np):
            lev.Q[p][1:-1, 1:-1] = prolong_bilin(self.gather_window(k, p))
        self.fill_ghosts(k)

    # ------------------------------------------------------------ diagnostics
    def composite_mean(self):
        """ebCompositeMean.m:4-23 -> (mean, area, total)."""
        tot = 0.0
        area = 0.0
        for k in range(self.rootIdx, self.nlev + 1):
            lev = self.L(k)
            da = lev.h[0] * lev.h[1]
            for p in range(lev.np):
                vis = ~lev.covered[p]
                v = lev.Q[p][1:-1, 1:-1].T[vis.T]
                tot = tot + float(np.sum(v)) * da
                area = area + np.count_nonzero(vis) * da
        return tot / max(area, REALMIN), area, tot

    def mass_check(self):
        """ebMassCheck.m:14-66."""
        lhs = 0.0
        scale = 0.0
        for k in range(self.rootIdx, self.nlev + 1):
            lev = self.L(k)
            da = lev.h[0] * lev.h[1]
            for p in range(lev.np):
                vis = ~lev.covered[p]
                Qi = lev.Q[p][1:-1, 1:-1]
                v = lev.CC[p][vis] * Qi[vis] - lev.FTrue[p][vis]
                lhs = lhs + float(np.sum(v)) * da
                scale = scale + float(np.sum(np.abs(lev.FTrue[p][vis]))) * da
        rhs = 0.0
        for k in range(self.rootIdx, self.nlev + 1):
            lev = self.L(k)
            hx, hy = lev.h
            for p in range(lev.np):
                Q = lev.Q[p]
                mx = Q.shape[0] - 2
                my = Q.shape[1] - 2
                cov = lev.covered[p]
                codes = lev.bc_code[p]
                if codes[0] > 0:
                    m = ~cov[0, :]
                    fl = lev.DeW[p][0, :] * (Q[0, 1:my + 1] - Q[1, 1:my + 1]) / hx
                    rhs = rhs + float(np.sum(fl[m])) * hy
                if codes[1] > 0:
                    m = ~cov[mx - 1, :]
                    fl = lev.DeW[p][mx, :] * (Q[mx + 1, 1:my + 1] - Q[mx, 1:my + 1]) / hx
                    rhs = rhs + float(np.sum(fl[m])) * hy
                if codes[2] > 0:
                    m = ~cov[:, 0]
                    fl = lev.DnS[p][:, 0] * (Q[1:mx + 1, 0] - Q[1:mx + 1, 1]) / hy
                    rhs = rhs + float(np.sum(fl[m])) * hx
                if codes[3] > 0:
                    m = ~cov[:, my - 1]
                    fl = lev.DnS[p][:, my] * (Q[1:mx + 1, my + 1] - Q[1:mx + 1, my]) / hy
                    rhs = rhs + float(np.sum(fl[m])) * hx
        defect = lhs - rhs
        return {"lhs": lhs, "flux": rhs, "defect": defect,
                "rel": abs(defect) / max(scale, abs(lhs), abs(rhs), REALMIN)}

    def copy(self):
        return copy.deepcopy(self)


# --------------------------------------------------------------------------
def assemble_coarsest(n1, n2, hx, hy, DeW, DnS, CC, codes, gs):
    """Dense ghost-eliminated matrix of ebCoarsest.m:14-58 with duplicates
    summed in triplet order; bcRhs per ebCoarsest.m:95-109."""
    N = n1 * n2
    hx2 = hx * hx
    hy2 = hy * hy
    A = np.zeros((N, N))
    bcRhs = np.zeros(N)
    idx = np.arange(N).reshape((n1, n2), order="F")
    rows = [idx.reshape(-1, order="F")]
    cols = [idx.reshape(-1, order="F")]
    vals = [CC.reshape(-1, order="F")]
    c = (DeW[1:n1, :] / hx2).reshape(-1, order="F")
    iL = idx[0:n1 - 1, :].reshape(-1, order="F")
    iR = idx[1:n1, :].reshape(-1, order="F")
    rows += [iL, iR, iL, iR]
    cols += [iL, iR, iR, iL]
    vals += [c, c, -c, -c]
    c = (DnS[:, 1:n2] / hy2).reshape(-1, order="F")
    iL = idx[:, 0:n2 - 1].reshape(-1, order="F")
    iR = idx[:, 1:n2].reshape(-1, order="F")
    rows += [iL, iR, iL, iR]
    cols += [iL, iR, iR, iL]
    vals += [c, c, -c, -c]
    sides = [(DeW[0, :], idx[0, :], idx[1, :], gs[0], codes[0], hx, hx2),
             (DeW[n1, :], idx[n1 - 1, :], idx[n1 - 2, :], gs[1], codes[1], hx, hx2),
             (DnS[:, 0], idx[:, 0], idx[:, 1], gs[2], codes[2], hy, hy2),
             (DnS[:, n2], idx[:, n2 - 1], idx[:, n2 - 2], gs[3], codes[3], hy, hy2)]
    for d, i1, i2, g, code, h, h2 in sides:
        if code == 1:
            rows += [i1, i1]
            cols += [i1, i2]
            vals += [3 * d / h2, -d / (3 * h2)]
            bcRhs[i1] = bcRhs[i1] + C83 * d * g / h2
        elif code == 2:
            bcRhs[i1] = bcRhs[i1] + d * g / h
    for r, c, v in zip(rows, cols, vals):
        # sequential scatter-add preserves triplet order for repeated targets
        for t in range(r.size):
            A[r[t], c[t]] = A[r[t], c[t]] + v[t]
    return A, bcRhs


def lu_inverse(M):
    """Explicit inverse by Gaussian elimination with partial pivoting
    (first maximum wins) and column-oriented substitution.  Parity unpinned
    against MATLAB's ``decomposition`` (not reproducible); pinned by
    M @ inv(M) = I to round-off in tests."""
    A = np.array(M, dtype=float, copy=True)
    n = A.shape[0]
    perm = np.arange(n)
    for k in range(n):
        piv = k + int(np.argmax(np.abs(A[k:, k])))
        if A[piv, k] == 0:
            raise ValueError("coarsest matrix is singular")
        if piv != k:
            A[[k, piv], :] = A[[piv, k], :]
            perm[[k, piv]] = perm[[piv, k]]
        A[k + 1:, k] = A[k + 1:, k] / A[k, k]
        A[k + 1:, k + 1:] = A[k + 1:, k + 1:] - np.outer(A[k + 1:, k], A[k, k + 1:])
    X = np.zeros((n, n))
    X[np.arange(n), perm] = 1.0
    for k in range(n):
        X[k + 1:, :] = X[k + 1:, :] - np.outer(A[k + 1:, k], X[k, :])
    for k in range(n - 1, -1, -1):
        X[k, :] = X[k, :] / A[k, k]
        X[:k, :] = X[:k, :] - np.outer(A[:k, k], X[k, :])
    return X


def dense_apply(Ainv, rhs, nrows):
    """u_i = sum_j Ainv[i,j] rhs_j accumulated in ascending j."""
    acc = np.zeros(nrows)
    for j in range(rhs.size):
        acc = acc + Ainv[:nrows, j] * rhs[j]
    return acc


def eb_solve(prob, opts, host, verbose=0):
    """ebSolve.m:22-77 driven by the oracle (host layer supplies setup,
    tagging/clustering and level construction; the oracle supplies every
    solve-path operation)."""
    H = host.setup(prob, opts)
    O = OracleHierarchy(H)
    O.fill_all = lambda: [O.fill_ghosts(k) for k in range(1, O.nlev + 1)]
    O.fill_all()

    def sync_q():
        for k in range(1, O.nlev + 1):
            for p in range(O.L(k).np):
                H.L(k).Q[p] = O.L(k).Q[p].copy()

    def add(boxes):
        nonlocal O
        sync_q()
        old = O
        kNew = host.add_level(H, boxes)
        O = OracleHierarchy(H)
        # carry over state the reference keeps in H (QC buffers)
        for k in range(2, old.nlev + 1):
            O.L(k).QC = [a.copy() for a in old.L(k).QC]
        O.stats = old.stats
        O.init_level_prolong(kNew)

    if opts.Indicator.lower() == "fun":
        while H.nlev - H.rootIdx + 1 < opts.MaxLevels:
            boxes = host.tag_cluster(H)
            if not boxes:
                break
            add(boxes)
        O.solve_cycles(verbose=verbose)
    else:
        O.solve_cycles(verbose=verbose)
        while H.nlev - H.rootIdx + 1 < opts.MaxLevels:
            sync_q()
            boxes = host.tag_cluster(H)
            if not boxes:
                break
            add(boxes)
            O.solve_cycles(verbose=verbose)
    sync_q()
    return H, O


# --------------------------------------------------------------------------
# Parabolic stepping (ParabolicBSAM/src): every step is one ebSolveCycles of
# the oracle with C_eff = C + alpha/dt.  Host decisions (ladder, tagging,
# clustering, level construction, coefficient fill-down) come from the host
# module passed in, exactly as for eb_solve above.
def ebp_rhs(f, Uo, Uoo, dt, scheme):
    """ebpRhs.m:31-36 (be / bdf2 branches; cn needs the hierarchy, see
    ParabolicOracle.cn_rhs)."""
    if scheme == "be":
        return f + Uo / dt
    if scheme == "bdf2":
        return f + (2 * Uo - 0.5 * Uoo) / dt
    raise ValueError(f"ebp_rhs:scheme: unknown scheme {scheme}")


def conslin(W):
    """ebpRegrid.m:95-108: per coarse cell, value and centred slopes."""
    C0 = W[1:-1, 1:-1]
    sx = (W[2:, 1:-1] - W[:-2, 1:-1]) / 2
    sy = (W[1:-1, 2:] - W[1:-1, :-2]) / 2
    F = np.zeros((2 * C0.shape[0], 2 * C0.shape[1]))
    F[0::2, 0::2] = C0 - sx / 4 - sy / 4
    F[1::2, 0::2] = C0 + sx / 4 - sy / 4
    F[0::2, 1::2] = C0 - sx / 4 + sy / 4
    F[1::2, 1::2] = C0 + sx / 4 + sy / 4
    return F


def conslin_interior(P):
    """ebpRegrid.m:111-134: one-sided slopes on the edges."""
    cm, cn = P.shape
    sx = np.zeros((cm, cn))
    sy = np.zeros((cm, cn))
    if cm > 1:
        sx[1:-1, :] = (P[2:, :] - P[:-2, :]) / 2
        sx[0, :] = P[1, :] - P[0, :]
        sx[-1, :] = P[-1, :] - P[-2, :]
    if cn > 1:
        sy[:, 1:-1] = (P[:, 2:] - P[:, :-2]) / 2
        sy[:, 0] = P[:, 1] - P[:, 0]
        sy[:, -1] = P[:, -1] - P[:, -2]
    F = np.zeros((2 * cm, 2 * cn))
    F[0::2, 0::2] = P - sx / 4 - sy / 4
    F[1::2, 0::2] = P + sx / 4 - sy / 4
    F[0::2, 1::2] = P - sx / 4 + sy / 4
    F[1::2, 1::2] = P + sx / 4 + sy / 4
    return F


class ParabolicOracle:
    """ebpSolve.m / ebpInit.m / ebpSetTime.m / ebpRegrid.m on the oracle.

    ``host`` supplies setup/add_level/tag_cluster/coeff_filldown/ndgrid,
    ``par`` the problem freezing (eprob) of the host layer."""

    def __init__(self, pp, opts, host, par):
        self.pp, self.opts, self.host, self.par = pp, opts, host, par

    # ---- helpers on (H, O) pairs
    @staticmethod
    def _sync(H, O):
        for k in range(1, O.nlev + 1):
            for p in range(O.L(k).np):
                H.L(k).Q[p] = O.L(k).Q[p].copy()

    def _add(self, H, O, boxes):
        """ebAddLevel.m:6-33 (the new level's prolonged initial guess included)."""
        self._sync(H, O)
        kNew = self.host.add_level(H, boxes)
        On = OracleHierarchy(H)
        for k in range(2, O.nlev + 1):
            On.L(k).QC = [a.copy() for a in O.L(k).QC]
        On.stats = O.stats
        On.init_level_prolong(kNew)
        return kNew, On

    def _u0(self, O, k):
        lev = O.L(k)
        dom = self.pp.domain
        for p in range(lev.np):
            b = lev.box[p]
            xc = dom[0] + (np.arange(b[0], b[1] + 1, dtype=float) - 0.5) * lev.h[0]
            yc = dom[2] + (np.arange(b[2], b[3] + 1, dtype=float) - 0.5) * lev.h[1]
            X, Y = np.meshgrid(xc, yc, indexing="ij")
            v = self.pp.u0(X, Y)
            lev.Q[p][1:-1, 1:-1] = v if np.ndim(v) else float(v) * np.ones(X.shape)

    @staticmethod
    def interiors(O):
        return [[O.L(k).Q[p][1:-1, 1:-1].copy() for p in range(O.L(k).np)] for k in range(1, O.nlev + 1)]

    # ---- ebpInit.m:10-35
    def init(self, aDt1):
        host, opts = self.host, self.opts
        H = host.setup(self.par.eprob(self.pp, 0.0, aDt1), opts.eb)
        O = OracleHierarchy(H)
        for k in range(1, O.nlev + 1):
            self._u0(O, k)
        for k in range(1, O.nlev + 1):
            O.fill_ghosts(k)
        self._sync(H, O)
        while H.nlev - H.rootIdx + 1 < opts.eb.MaxLevels:
            boxes = host.tag_cluster(H)
            if not boxes:
                break
            kNew, O = self._add(H, O, boxes)
            self._u0(O, kNew)
            O.fill_ghosts(kNew)
            self._sync(H, O)
        O.transfer(0, "filldown")
        self._sync(H, O)
        self.H, self.O = H, O
        self.Uo, self.Uoo = self.interiors(O), None
        td = self.pp.tdD or self.pp.tdC or self.pp.tdBC
        self.stateKey = (0.0 * td, aDt1)

    # ---- ebpSetTime.m:14-55
    def set_time(self, t, aDt):
        pp = self.pp
        key = (t * (pp.tdD or pp.tdC or pp.tdBC), aDt)
        if key == self.stateKey:
            return
        H, O = self.H, self.O
        dom = pp.domain
        for k in range(1, H.nlev + 1):
            lev = H.L(k)
            for p in range(lev.np):
                b = lev.box[p]
                xc = dom[0] + (np.arange(b[0], b[1] + 1, dtype=float) - 0.5) * lev.h[0]
                yc = dom[2] + (np.arange(b[2], b[3] + 1, dtype=float) - 0.5) * lev.h[1]
                if pp.tdD:
                    xf = dom[0] + np.arange(b[0] - 1, b[1] + 1, dtype=float) * lev.h[0]
                    yf = dom[2] + np.arange(b[2] - 1, b[3] + 1, dtype=float) * lev.h[1]
                    XF, YC = np.meshgrid(xf, yc, indexing="ij")
                    lev.DeW[p] = np.broadcast_to(pp.D(XF, YC, t), XF.shape).astype(float)
                    XC, YF = np.meshgrid(xc, yf, indexing="ij")
                    lev.DnS[p] = np.broadcast_to(pp.D(XC, YF, t), XC.shape).astype(float)
                X, Y = np.meshgrid(xc, yc, indexing="ij")
                lev.CC[p] = np.broadcast_to(pp.C(X, Y, t), X.shape).astype(float) + aDt
                if pp.tdBC:
                    fint = lev.faceInternal[p]
                    if not fint[0]:
                        lev.bc_g[p][0] = np.ravel(np.broadcast_to(pp.bc[0].g(dom[0] * np.ones(yc.shape), yc, t), yc.shape)).astype(float)
                    if not fint[1]:
                        lev.bc_g[p][1] = np.ravel(np.broadcast_to(pp.bc[1].g(dom[1] * np.ones(yc.shape), yc, t), yc.shape)).astype(float)
                    if not fint[2]:
                        lev.bc_g[p][2] = np.ravel(np.broadcast_to(pp.bc[2].g(xc, dom[2] * np.ones(xc.shape), t), xc.shape)).astype(float)
                    if not fint[3]:
                        lev.bc_g[p][3] = np.ravel(np.broadcast_to(pp.bc[3].g(xc, dom[3] * np.ones(xc.shape), t), xc.shape)).astype(float)
        self.host.coeff_filldown(H)
        for k in range(1, H.nlev + 1):
            src, dst = H.L(k), O.L(k)
            dst.CC = [a.copy() for a in src.CC]
            dst.DeW = [a.copy() for a in src.DeW]
            dst.DnS = [a.copy() for a in src.DnS]
            dst.InvDiagS = [a.copy() for a in src.InvDiagS]
            dst.bc_g = [[g.copy() for g in gs] for gs in src.bc_g]
        O.gather_parent_coeffs()
        O.coarsest_assemble()
        self.stateKey = key

    # ---- ebpRhs.m:16-67, the cn branch: 2 f(t1 - dt/2) + (2/dt + aDt) Uo - L(u^n)
    # with the current (old-time) operator, plus the old state's coarse-fine
    # flux corrections added to the parent cells
    def cn_rhs(self, t1, dt):
        O = self.O
        dom = self.pp.domain
        O.transfer(0, "filldown")
        for k in range(1, O.nlev + 1):
            O.fill_ghosts(k)
        aDt = self.stateKey[1]
        for k in range(O.rootIdx, O.nlev + 1):
            lev = O.L(k)
            for p in range(lev.np):
                b = lev.box[p]
                X, Y = np.meshgrid(dom[0] + (np.arange(b[0], b[1] + 1, dtype=float) - 0.5) * lev.h[0],
                                   dom[2] + (np.arange(b[2], b[3] + 1, dtype=float) - 0.5) * lev.h[1],
                                   indexing="ij")
                f = np.broadcast_to(self.pp.f(X, Y, t1 - dt / 2), X.shape).astype(float)
                Le = op2d(lev.Q[p], lev.DeW[p], lev.DnS[p], lev.CC[p], lev.h[0], lev.h[1])
                lev.FTrue[p] = 2 * f + (2 / dt + aDt) * self.Uo[k - 1][p] - Le
        if O.opts.MassCorrection:
            for k in range(O.rootIdx + 1, O.nlev + 1):
                O.transfer(k, "ring")
                lev = O.L(k)
                FTpar = O.L(k - 1).FTrue
                for p in range(lev.np):
                    vals = O.face_corrections(k, p)
                    for fc in range(4):
                        for q, src, di, dj in lev.ops[p]["corr"][fc]:
                            FTpar[q][di, dj] = FTpar[q][di, dj] + vals[fc][src]

    # ---- one step of ebpSolve.m:40-58
    def step(self, t1, dt, scheme):
        O = self.O
        dom = self.pp.domain
        if scheme == "cn":
            self.cn_rhs(t1, dt)
            self.set_time(t1, ALPHA_OF[scheme] / dt)
            st = O.solve_cycles()
            self.Uoo = self.Uo
            self.Uo = self.interiors(O)
            m, area, _ = O.composite_mean()
            return st, m * area
        for k in range(O.rootIdx, O.nlev + 1):
            lev = O.L(k)
            for p in range(lev.np):
                b = lev.box[p]
                X, Y = np.meshgrid(dom[0] + (np.arange(b[0], b[1] + 1, dtype=float) - 0.5) * lev.h[0],
                                   dom[2] + (np.arange(b[2], b[3] + 1, dtype=float) - 0.5) * lev.h[1],
                                   indexing="ij")
                f = np.broadcast_to(self.pp.f(X, Y, t1), X.shape).astype(float)
                Uoo = self.Uoo[k - 1][p] if scheme == "bdf2" else None
                lev.FTrue[p] = ebp_rhs(f, self.Uo[k - 1][p], Uoo, dt, scheme)
        self.set_time(t1, ALPHA_OF[scheme] / dt)
        st = O.solve_cycles()
        self.Uoo = self.Uo
        self.Uo = self.interiors(O)
        m, area, _ = O.composite_mean()
        return st, m * area

    # ---- ebpRegrid.m:15-59
    def regrid(self, t, aDtNext):
        host, opts = self.host, self.opts
        Hold, Oold = self.H, self.O
        Hn = host.setup(self.par.eprob(self.pp, t, aDtNext), opts.eb)
        r = Hn.rootIdx
        On = OracleHierarchy(Hn)
        On.L(r).Q[0][1:-1, 1:-1] = Oold.L(r).Q[0][1:-1, 1:-1]
        UooN = [[] for _ in range(r - 1)] + [[self.Uoo[r - 1][0]]]
        for k in range(r - 1, 0, -1):
            On.L(k).Q[0][1:-1, 1:-1] = restrict2(On.L(k + 1).Q[0][1:-1, 1:-1])
        for k in range(1, r + 1):
            On.fill_ghosts(k)
        self._sync(Hn, On)
        while Hn.nlev - Hn.rootIdx + 1 < opts.eb.MaxLevels:
            boxes = host.tag_cluster(Hn)
            if not boxes:
                break
            kN, On = self._add(Hn, On, boxes)
            lev = On.L(kN)
            row = []
            for p in range(lev.np):
                Qf = conslin(On.gather_window(kN, p))
                b = lev.box[p]
                Pf = np.zeros(((b[1] - b[0] + 1) // 2, (b[3] - b[2] + 1) // 2))
                for q, s, d in lev.ops[p]["fsc"]:
                    Pf[s[0]:s[1] + 1, s[2]:s[3] + 1] = UooN[kN - 2][q][d[0]:d[1] + 1, d[2]:d[3] + 1]
                Uf = conslin_interior(Pf)
                if kN <= Oold.nlev:
                    for q in range(Oold.L(kN).np):
                        qb = Oold.L(kN).box[q]
                        rr = [max(b[0], qb[0]), min(b[1], qb[1]), max(b[2], qb[2]), min(b[3], qb[3])]
                        if rr[0] > rr[1] or rr[2] > rr[3]:
                            continue
                        di = slice(rr[0] - b[0], rr[1] - b[0] + 1)
                        dj = slice(rr[2] - b[2], rr[3] - b[2] + 1)
                        si = slice(rr[0] - qb[0], rr[1] - qb[0] + 1)
                        sj = slice(rr[2] - qb[2], rr[3] - qb[2] + 1)
                        Qf[di, dj] = Oold.L(kN).Q[q][1:-1, 1:-1][si, sj]
                        Uf[di, dj] = self.Uoo[kN - 1][q][si, sj]
                lev.Q[p][1:-1, 1:-1] = Qf
                row.append(Uf)
            UooN.append(row)
            On.fill_ghosts(kN)
            self._sync(Hn, On)
        On.transfer(0, "filldown")
        self._sync(Hn, On)
        On.stats = Oold.stats
        self.H, self.O = Hn, On
        self.Uo = self.interiors(On)
        self.Uoo = UooN
        td = self.pp.tdD or self.pp.tdC or self.pp.tdBC
        self.stateKey = (t * td, aDtNext)

    # ---- ebpSolve.m:23-78
    def solve(self, callback=None):
        opts = self.opts
        q = opts.TFinal / opts.Dt
        nsteps = max(1, int(np.floor(q + 0.5)))
        dt = opts.TFinal / nsteps
        scheme1 = "be" if opts.Scheme == "bdf2" else opts.Scheme
        self.init(ALPHA_OF[scheme1] / dt)
        m, area, _ = self.O.composite_mean()
        mass = [m * area]
        stages = []
        levels = []
        for n in range(1, nsteps + 1):
            t1 = n * dt
            scheme = scheme1 if n == 1 else opts.Scheme
            st, mn = self.step(t1, dt, scheme)
            stages.append(st)
            mass.append(mn)
            levels.append((self.O.nlev - self.O.rootIdx + 1, st["totalCells"]))
            if callback is not None:
                callback(self, n)
            if opts.RegridEvery > 0 and n % opts.RegridEvery == 0 and n < nsteps and opts.eb.MaxLevels > 1:
                self.regrid(t1, ALPHA_OF[opts.Scheme] / dt)
        return {"stages": stages, "mass": mass, "levels": levels, "dt": dt, "nsteps": nsteps}


ALPHA_OF = {"be": 1.0, "bdf2": 1.5, "cn": 2.0}
