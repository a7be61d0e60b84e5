"""Host-side diagnostics on a downloaded hierarchy.

* ``composite_mean`` <- EllipticBSAM/src/ebCompositeMean.m:1-24
* ``mass_check``     <- EllipticBSAM/src/ebMassCheck.m:1-67
* ``error_norms``    <- EllipticBSAM/src/ebErrorNorms.m:1-70
"""
from __future__ import annotations

import numpy as np

REALMIN = np.finfo(float).tiny


def composite_mean(H, E=None):
    tot = 0.0
    area = 0.0
    for k in range(H.rootIdx, H.nlev + 1):
        L = H.L(k)
        da = L.h[0] * L.h[1]
        for p in range(L.np):
            vis = ~L.covered[p]
            v = (E[k][p] if E is not None else L.Q[p][1:-1, 1:-1])[vis]
            tot += float(np.sum(v)) * da
            area += np.count_nonzero(vis) * da
    return tot / max(area, REALMIN), area


def mass_check(H):
    lhs = 0.0
    scale = 0.0
    for k in range(H.rootIdx, H.nlev + 1):
        L = H.L(k)
        da = L.h[0] * L.h[1]
        for p in range(L.np):
            vis = ~L.covered[p]
            Qi = L.Q[p][1:-1, 1:-1]
            lhs += float(np.sum(L.CC[p][vis] * Qi[vis] - L.FTrue[p][vis])) * da
            scale += float(np.sum(np.abs(L.FTrue[p][vis]))) * da
    rhs = 0.0
    for k in range(H.rootIdx, H.nlev + 1):
        L = H.L(k)
        hx, hy = L.h
        for p in range(L.np):
            Q = L.Q[p]
            mx, my = Q.shape[0] - 2, Q.shape[1] - 2
            cov = L.covered[p]
            c = L.bc_code[p]
            if c[0] > 0:
                rhs += float(np.sum((L.DeW[p][0, :] * (Q[0, 1:my + 1] - Q[1, 1:my + 1]) / hx)[~cov[0, :]])) * hy
            if c[1] > 0:
                rhs += float(np.sum((L.DeW[p][mx, :] * (Q[mx + 1, 1:my + 1] - Q[mx, 1:my + 1]) / hx)[~cov[mx - 1, :]])) * hy
            if c[2] > 0:
                rhs += float(np.sum((L.DnS[p][:, 0] * (Q[1:mx + 1, 0] - Q[1:mx + 1, 1]) / hy)[~cov[:, 0]])) * hx
            if c[3] > 0:
                rhs += float(np.sum((L.DnS[p][:, my] * (Q[1:mx + 1, my + 1] - Q[1:mx + 1, my]) / hy)[~cov[:, my - 1]])) * hx
    d = lhs - rhs
    return {"lhs": lhs, "flux": rhs, "defect": d, "rel": abs(d) / max(scale, abs(lhs), abs(rhs), REALMIN)}


def error_norms(H, ufun=None):
    ufun = ufun or H.prob.exact
    dom = H.prob.domain
    E = {}
    for k in range(H.rootIdx, H.nlev + 1):
        L = H.L(k)
        E[k] = []
        for p in range(L.np):
            XC, YC = L.cell_centers(dom, p)
            E[k].append(L.Q[p][1:-1, 1:-1] - np.asarray(ufun(XC, YC), float))
    m = 0.0
    if H.singular:
        m, _ = composite_mean(H, E)
        for k in E:
            E[k] = [e - m for e in E[k]]
    linf = 0.0
    l2sq = 0.0
    dof = 0
    for k in range(H.rootIdx, H.nlev + 1):
        L = H.L(k)
        da = L.h[0] * L.h[1]
        for p in range(L.np):
            e = E[k][p][~L.covered[p]]
            if e.size:
                linf = max(linf, float(np.max(np.abs(e))))
                l2sq += float(np.sum(e ** 2)) * da
                dof += e.size
    return {"linf": linf, "l2": float(np.sqrt(l2sq)), "dof": dof, "meanShift": m}
