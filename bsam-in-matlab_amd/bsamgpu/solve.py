"""Entry points of the host layer, mirroring EllipticBSAM/src/ebSolve.m:1-95.

Hierarchy construction (ladder, tagging, clustering, level build) stays on the
host; every solve stage (== ebSolveCycles) runs on the device library.
"""
from __future__ import annotations

import numpy as np

from . import diagnostics, hierarchy, tagcluster
from .device import GpuHierarchy
from .problem import Options, Problem


class Solution(dict):
    __getattr__ = dict.__getitem__


def _add_level(H, boxes, dev):
    """ebAddLevel.m:10-40: host level build, then the device prolongation of
    the parent window as the new level's initial guess.  The existing levels
    keep their device state (ebAddLevel leaves them unchanged): the new
    hierarchy takes their Q device to device from the old one."""
    kNew = hierarchy.add_level(H, boxes)
    new = GpuHierarchy(H)
    if dev is not None:
        new.inherit(dev)
        dev.close()
    dev = new
    dev.init_level_prolong(kNew)
    dev.pull_q(H)
    if H.opts.Verbose >= 1:
        L = H.L(kNew)
        ncell = sum(int(np.prod(L.dims(p))) for p in range(L.np))
        print(f"[ebsam] added level {kNew - H.rootIdx + 1} ({L.np} patch(es), {ncell} cells, h = {L.h[0]:.3e})")
    return dev


def _stage(H, dev, opts):
    st = dev.solve_cycles(opts)
    dev.pull_q(H)
    H.stats["stages"].append(st)
    if opts.Verbose >= 1:
        print(f"[ebsam] stage {len(H.stats['stages'])}: {st['amrLevels']} AMR level(s), {st['vcycles']} V-cycle(s), "
              f"|r|/|f| = {st['finalRel']:.2e}, avg factor {st['avgFactor']:.3f}, {st['visibleCells']} visible cells")
    return st


def eb_solve(prob: Problem, opts: Options | None = None, **kw) -> Solution:
    """Block-structured adaptive FAS multigrid solve of
    -div(D grad u) + C u = f  (ebSolve.m:1-95)."""
    opts = (opts or Options()).replace(**kw) if kw else (opts or Options())
    H = hierarchy.setup(prob, opts)
    dev = GpuHierarchy(H)
    if opts.Indicator.lower() == "fun":
        while H.nlev - H.rootIdx + 1 < opts.MaxLevels:
            boxes = tagcluster.tag_cluster(H)
            if not boxes:
                break
            dev = _add_level(H, boxes, dev)
        _stage(H, dev, opts)
    else:
        _stage(H, dev, opts)
        while H.nlev - H.rootIdx + 1 < opts.MaxLevels:
            boxes = tagcluster.tag_cluster(H, device=dev)
            if not boxes:
                if opts.Verbose >= 1:
                    print(f"[ebsam] no cells tagged; stopping at {H.nlev - H.rootIdx + 1} AMR level(s)")
                break
            dev = _add_level(H, boxes, dev)
            _stage(H, dev, opts)
    mz = opts.MeanZero
    if isinstance(mz, str) and mz.lower() == "auto":
        mz = H.singular
    if mz:
        m, _ = diagnostics.composite_mean(H)
        for k in range(1, H.nlev + 1):
            for p in range(H.L(k).np):
                H.L(k).Q[p] = H.L(k).Q[p] - m
    sol = Solution(H=H, stats=H.stats, device=dev)
    sol["err"] = diagnostics.error_norms(H) if prob.exact is not None else None
    sol["mass"] = diagnostics.mass_check(H)
    sol["levels"] = [{"n": H.L(k).n, "h": H.L(k).h, "np": H.L(k).np,
                      "cells": sum(int(np.prod(H.L(k).dims(p))) for p in range(H.L(k).np)),
                      "box": H.L(k).box.copy()} for k in range(H.rootIdx, H.nlev + 1)]
    return sol


def build_hierarchy(prob: Problem, opts: Options):
    """Host hierarchy of a geometry-tagged ('fun') problem with all levels
    added (ebSolve.m:30-37) and zero initial guess."""
    H = hierarchy.setup(prob, opts)
    while H.nlev - H.rootIdx + 1 < opts.MaxLevels:
        boxes = tagcluster.tag_cluster(H)
        if not boxes:
            break
        hierarchy.add_level(H, boxes)
    return H
