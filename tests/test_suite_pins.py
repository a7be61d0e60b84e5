"""The reference's verification suites reproduced on the CPU (no GPU): every
solve that EllipticBSAM/tests/runAllTests.m prints (T1-T5, T8), run through
the oracle with the host layer's hierarchy construction, against the numbers
published in verification/suite-output.txt:3-79; and the clustering outcomes
of the tree variant's suite (T8 with per-parent clustering, the T9 straddle,
EllipticBSAMTree/tests/runAllTreeTests.m:208-285, published at
verification/suite-output.txt:152-176).

The published rows pin more than the solver: the manufactured problems'
sampling, the host's tagging (fun and ulap), Berger-Rigoutsos clustering,
level construction and the composite error norms (ebErrorNorms.m) all enter
the printed error to 5 significant digits, the dof count and the patch
counts.  The same runs go through the device library in test_gpu_suite.py.
"""
import json
import os
import warnings

import numpy as np
import pytest

from bsamgpu import Options, diagnostics, hierarchy, tagcluster
from bsamgpu.problem import call_tagfun
from oracle import bsam_oracle as O
import suite_cases as S

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "published.json")))["suite"]
RUNS = S.suite_runs(GOLD)


class _Host:
    setup = staticmethod(hierarchy.setup)
    add_level = staticmethod(hierarchy.add_level)
    tag_cluster = staticmethod(tagcluster.tag_cluster)


def oracle_solve(prob, opts):
    """ebSolve.m:22-95 with every solve stage by the oracle; the mean pin and
    the diagnostics as the host layer's eb_solve applies them."""
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        H, Orc = O.eb_solve(prob, opts, _Host)
    mz = opts.MeanZero
    if isinstance(mz, str) and mz.lower() == "auto":
        mz = H.singular
    if mz:
        m, _ = diagnostics.composite_mean(H)
        for k in range(1, H.nlev + 1):
            for p in range(H.L(k).np):
                H.L(k).Q[p] = H.L(k).Q[p] - m
    return H, Orc.stats["stages"], diagnostics.error_norms(H), diagnostics.mass_check(H)


@pytest.mark.parametrize("run", RUNS, ids=[r[0] for r in RUNS])
def test_suite_solve_matches_published(run):
    label, prob, opts, row = run
    H, stages, err, mass = oracle_solve(prob, opts)
    S.check_published(label, H, stages[-1], err, mass, row)


def _grow(H, cluster):
    """The 'fun' hierarchy of ebSolve.m:30-37 / ebtSolve.m:32-40 with the
    given clustering of the current finest level; returns the number of
    dropped-sliver warnings."""
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        while H.nlev - H.rootIdx + 1 < H.opts.MaxLevels:
            boxes = cluster(H)
            if not boxes:
                break
            hierarchy.add_level(H, boxes)
    return sum("dropthin" in str(x.message) for x in w)


def _fun_tags(H):
    """'fun' tags of the finest level on the level-global map, dilated by
    TagBuffer and clipped to the erode-8 nesting region (ebTagCluster.m:27-90,
    ebtTagCluster.m:40-92 alike)."""
    opts, k = H.opts, H.nlev
    lev = H.L(k)
    T = np.zeros(lev.n, dtype=bool)
    cover = np.zeros(lev.n, dtype=bool)
    for p in range(lev.np):
        b = lev.box[p]
        XC, YC = lev.cell_centers(H.prob.domain, p)
        T[b[0] - 1:b[1], b[2] - 1:b[3]] |= call_tagfun(opts.TagFunction, XC, YC, k - H.rootIdx + 1)
        cover[b[0] - 1:b[1], b[2] - 1:b[3]] = True
    allowed = tagcluster.erode8(cover)
    return tagcluster.dilate(T, opts.TagBuffer) & allowed, allowed


def _tree_cluster(native):
    """ebtTagCluster.m:40-128 with 'fun' tags: tags on the level-global map,
    the global buffer and nesting clip, then Berger-Rigoutsos inside every
    parent patch.  Boxes lose their parent column for the level build."""

    def cluster(H):
        opts = H.opts
        T, allowed = _fun_tags(H)
        if not T.any():
            return []
        parents = [[int(v) for v in b] for b in H.L(H.nlev).box]
        f = tagcluster.cluster_per_parent_native if native else tagcluster.cluster_per_parent
        boxes = f(T, allowed, parents, opts.MinBoxWidth, opts.MinFillRatio)
        cluster.parents.append([b[0] for b in boxes])
        return [b[1:] for b in boxes]

    cluster.parents = []
    return cluster


def _patch_stats(H):
    bx = [b for k in range(H.rootIdx + 1, H.nlev + 1) for b in H.L(k).box]
    return len(bx), min(min(b[1] - b[0] + 1, b[3] - b[2] + 1) for b in bx)


def _t8_options():
    return Options(BaseCells=[64, 64], MaxLevels=3, Indicator="fun", TagFunction=S.t8_tagfn(), TagBuffer=1,
                   MinBoxWidth=4, Verbose=0, RelTol=1e-10, MaxVCycles=60)


@pytest.mark.parametrize("native", [False, True], ids=["restatement", "library"])
def test_t8_flat_clustering_published(native):
    """Flat T8: 49 patches, min width 4, one dropped sliver
    (suite-output.txt:70-79), with the host restatement and with the
    library's bsam_cluster_tags."""
    H = hierarchy.setup(S.mms_case("vardc"), _t8_options())

    def cluster(H):
        if not native:
            return tagcluster.tag_cluster(H)
        T, allowed = _fun_tags(H)
        if not T.any():
            return []
        return tagcluster.cluster_native(T, allowed, H.opts.MinBoxWidth, H.opts.MinFillRatio)

    nd = _grow(H, cluster)
    assert _patch_stats(H) == (GOLD["T8"]["patches"], GOLD["T8"]["min_width"])
    assert nd == GOLD["T8"]["dropthin_warnings"]


@pytest.mark.parametrize("native", [False, True], ids=["restatement", "library"])
def test_t8_tree_per_parent_clustering_published(native):
    """Tree T8: per-parent clustering gives 50 patches (one more than the
    flat variant), min width 4, two dropped slivers
    (suite-output.txt:152-168), with the restatement and with the library's
    bsam_cluster_tags_per_parent."""
    H = hierarchy.setup(S.mms_case("vardc"), _t8_options())
    nd = _grow(H, _tree_cluster(native))
    assert _patch_stats(H) == (GOLD["T8_tree"]["patches"], GOLD["T8_tree"]["min_width"])
    assert nd == GOLD["T8_tree"]["dropthin_warnings"]


@pytest.mark.parametrize("native", [False, True], ids=["restatement", "library"])
def test_t9_straddle_splits_per_parent(native):
    """T9 straddle (runAllTreeTests.m:258-283, suite-output.txt:173): a
    feature on the shared edge of two level-2 patches yields two level-3
    children, parents [1 2]."""
    opts = Options(BaseCells=[64, 64], MaxLevels=3, Indicator="fun", TagFunction=S.t9_tagfn, TagBuffer=0,
                   MinFillRatio=0.95, Verbose=0, RelTol=1e-11, MaxVCycles=60)
    H = hierarchy.setup(S.mms_case("gauss52"), opts)
    cl = _tree_cluster(native)
    _grow(H, cl)
    want = GOLD["T9_tree_straddle"]
    assert H.L(H.nlev - 1).np == want["level2_patches"]
    assert H.L(H.nlev).np == want["level3_patches"]
    assert cl.parents[-1] == want["parents"]
