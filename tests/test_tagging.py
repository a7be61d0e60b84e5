"""Tagging (EllipticBSAM/src/ebTagCluster.m:27-90, ebRte.m:14-42).

CPU: the host 'rte' indicator against closed forms --
  u = x^4, D = 1, C const: tau = L_2h(R u) - R(L_h u) = -6 h^2 on every
  coarse cell away from the x edges (second differences of x^4 and of the
  restricted x^4 + 1.5 h^2 x^2 + h^4/16);
  u linear along an edge, constant across it: the averaged ghost pair sits
  at the coarse ghost's tangential position, so tau = 0 on that edge (a
  shifted pair gives a nonzero tau there).
GPU: the device tag maps (bsam_tag_level) equal the host ones bit for bit on
every test hierarchy, both indicators and several buffers; the staged solve
with 'rte' tags on the device builds the oracle-driven hierarchy and Q bit
for bit and meets the reference's audit criterion (auditExtras.m:157-171:
three levels cut the max error to <= 0.35 of one level's).
"""
import numpy as np
import pytest

from bsamgpu import tagcluster


def _padded(fun, mx, my, h):
    xc = (np.arange(mx + 2) - 0.5) * h
    yc = (np.arange(my + 2) - 0.5) * h
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    return fun(X, Y)


def test_rte_quartic_closed_form():
    mx = my = 16
    h = 1.0 / 16
    Q = _padded(lambda x, y: x ** 4, mx, my, h)
    DeW = np.ones((mx + 1, my))
    DnS = np.ones((mx, my + 1))
    CC = np.full((mx, my), 0.75)
    ind = tagcluster.rte_indicator(Q, DeW, DnS, CC, h, h)
    assert ind.shape == (mx, my)
    inner = ind[2:mx - 2, :]  # coarse cells 1 .. cmx-2
    assert np.allclose(inner, 6 * h * h, rtol=0, atol=1e-10)
    # every 2x2 block carries its coarse value
    assert np.array_equal(ind[0::2, 0::2], ind[1::2, 1::2])


@pytest.mark.parametrize("axis", [0, 1])
def test_rte_edge_ghost_pairs(axis):
    mx, my = 12, 16
    h = 1.0 / 32
    lin = (lambda x, y: 0.3 + 2.0 * y) if axis == 0 else (lambda x, y: 0.3 + 2.0 * x)
    Q = _padded(lin, mx, my, h)
    D = np.ones  # constant D: L_h u = 0 for linear u with exact fine ghosts
    ind = tagcluster.rte_indicator(Q, D((mx + 1, my)), D((mx, my + 1)), np.zeros((mx, my)), h, h)
    tau = ind[0::2, 0::2]
    if axis == 0:  # u varies along the x edges (west/east ghost columns)
        edge = tau[[0, -1], 1:-1]
    else:  # u varies along the y edges (south/north ghost rows)
        edge = tau[1:-1, [0, -1]]
    assert np.max(edge) <= 1e-9
    # the same data with the edge ghosts shifted by one fine cell is caught
    Qs = Q.copy()
    if axis == 0:
        Qs[0, 1:-1] = Q[0, :-2]
    else:
        Qs[1:-1, 0] = Q[:-2, 0]
    bad = tagcluster.rte_indicator(Qs, D((mx + 1, my)), D((mx, my + 1)), np.zeros((mx, my)), h, h)[0::2, 0::2]
    assert np.max(bad[[0], 1:-1] if axis == 0 else bad[1:-1, [0]]) > 1.0


def test_rte_rejects_odd_widths():
    from bsamgpu.problem import BsamError

    with pytest.raises(BsamError):
        tagcluster.rte_indicator(np.zeros((7, 8)), np.ones((6, 6)), np.ones((5, 7)), np.zeros((5, 6)), 0.1, 0.1)


def _host_and_device_tags(H, G, indicator, thr, buffer):
    H.opts = H.opts.replace(Indicator=indicator, TagThreshold=thr, TagBuffer=buffer)
    boxes_h, info = tagcluster.tag_cluster(H, with_info=True)
    boxes_d, info_d = tagcluster.tag_cluster(H, with_info=True, device=G)
    return boxes_h, info, boxes_d, info_d


@pytest.mark.gpu
@pytest.mark.parametrize("indicator", ["ulap", "rte"])
def test_device_tags_equal_host_tags(indicator):
    from cases import CASES
    from bsamgpu.device import GpuHierarchy

    for name, make in CASES.items():
        H = make(seed=3)
        G = GpuHierarchy(H)
        for k in range(1, H.nlev + 1):
            G.fill_ghosts(k)
        G.pull_q(H)
        L = H.L(H.nlev)
        if indicator == "rte" and any(d % 2 for p in range(L.np) for d in L.dims(p)):
            G.close()
            continue
        s = np.concatenate([(tagcluster.ulap_indicator(L.Q[p]) if indicator == "ulap" else
                             tagcluster.rte_indicator(L.Q[p], L.DeW[p], L.DnS[p], L.CC[p], L.h[0], L.h[1])).ravel()
                            for p in range(L.np)])
        for q, buffer in ((0.5, 2), (0.9, 0), (0.97, 1), (1.0, 2)):
            thr = float(np.quantile(s, q))
            print(f"case {name} {indicator} threshold q{q} buffer {buffer}", flush=True)
            bh, ih, bd, idv = _host_and_device_tags(H, G, indicator, thr, buffer)
            assert np.array_equal(idv["Traw"], ih["Traw"]), name
            assert np.array_equal(idv["T"], ih["T"]), name
            assert bd == bh, name
        G.close()


def _threebumps():
    """EllipticBSAM/tests/mmsCases.m:41-53: three sharp Gaussians."""
    from bsamgpu import Problem

    s = 0.02
    xi, yi, Ai = (0.30, 0.66, 0.49), (0.34, 0.42, 0.72), (1.0, 0.8, -0.9)

    def g(x, y, a, b):
        return np.exp(-((x - a) ** 2 + (y - b) ** 2) / (2 * s ** 2))

    def lg(x, y, a, b):
        return g(x, y, a, b) * (((x - a) ** 2 + (y - b) ** 2) / s ** 4 - 2 / s ** 2)

    def uex(x, y):
        return Ai[0] * g(x, y, xi[0], yi[0]) + Ai[1] * g(x, y, xi[1], yi[1]) + Ai[2] * g(x, y, xi[2], yi[2])

    def lap(x, y):
        return Ai[0] * lg(x, y, xi[0], yi[0]) + Ai[1] * lg(x, y, xi[1], yi[1]) + Ai[2] * lg(x, y, xi[2], yi[2])

    return Problem([0, 1, 0, 1], 1, 1, lambda x, y: uex(x, y) - lap(x, y), ("dirichlet", uex), uex)


@pytest.mark.gpu
def test_rte_staged_solve_matches_oracle_and_audit():
    from bsamgpu import Options, hierarchy, solve
    from oracle import bsam_oracle as O

    class _Host:
        setup = staticmethod(hierarchy.setup)
        add_level = staticmethod(hierarchy.add_level)
        tag_cluster = staticmethod(tagcluster.tag_cluster)

    prob = _threebumps()
    errs = []
    for nlv in (1, 2, 3):
        opts = Options(BaseCells=64, MaxLevels=nlv, Indicator="rte", TagThreshold=lambda L: 2e-3 / 4 ** (L - 1),
                       TagBuffer=2, Verbose=0, RelTol=1e-9, MaxVCycles=40)
        sol = solve.eb_solve(prob, opts)
        sol["device"].close()
        Hg = sol["H"]
        errs.append(sol["err"]["linf"])
        print(f"rte MaxLevels={nlv}: levels {Hg.nlev}, patches {[Hg.L(k).np for k in range(1, Hg.nlev + 1)]}, "
              f"err_inf {errs[-1]:.4e}", flush=True)
        if nlv == 3:
            Ho, _ = O.eb_solve(prob, opts, _Host)
            assert Hg.nlev == Ho.nlev
            for k in range(1, Hg.nlev + 1):
                assert [list(b) for b in Hg.L(k).box] == [list(b) for b in Ho.L(k).box], k
                for p in range(Hg.L(k).np):
                    assert np.array_equal(Hg.L(k).Q[p], Ho.L(k).Q[p]), (k, p)
    assert errs[2] <= 0.35 * errs[0]


@pytest.mark.gpu
def test_tag_level_errors():
    import ctypes

    from cases import CASES
    from bsamgpu import _lib
    from bsamgpu.device import GpuHierarchy

    H = CASES["plain"](seed=1)
    G = GpuHierarchy(H)
    k = H.nlev
    n1, n2 = H.L(k).n
    assert G.lib.bsam_level_field_size(G.h, k, 4) == n1 * n2
    T = np.zeros(n1 * n2, dtype=np.uint8)
    assert G.lib.bsam_tag_level(G.h, k, 7, 0.1, 2, T.ctypes.data, None, None) == -1
    assert G.lib.bsam_tag_level(G.h, H.nlev + 1, 0, 0.1, 2, T.ctypes.data, None, None) == -1
    buf = np.zeros(n1 * n2)
    assert G.lib.bsam_get_level(G.h, k, 4, buf.ctypes.data_as(_lib.P_double)) == -1  # size-only field
    cnt = ctypes.c_longlong(-1)
    assert G.lib.bsam_tag_level(G.h, k, 0, np.inf, 2, None, None, ctypes.byref(cnt)) == 0 and cnt.value == 0
    G.close()
