"""bsam_tag_level_static: tagging without device allocation.

The entry point keeps its tag maps in a library-owned device array; its
output must equal bsam_tag_level's byte for byte (tags, raw tags, count), on
two calls in a row, the second on a smaller level map than the first, so a
stale byte left by the larger call would show.  Also: its size limits reject
a request with BSAM_E_ARG without touching the device.
"""
import ctypes

import numpy as np
import pytest


def _tag_both(G, k, code, thr, buffer):
    n1, n2 = (int(v) for v in G.H.L(k).n)
    out = []
    for name in ("bsam_tag_level", "bsam_tag_level_static"):
        T = np.full(n1 * n2, 7, dtype=np.uint8)
        R = np.full(n1 * n2, 7, dtype=np.uint8)
        cnt = ctypes.c_longlong(-1)
        rc = getattr(G.lib, name)(G.h, k, code, float(thr), int(buffer), T.ctypes.data, R.ctypes.data,
                                  ctypes.byref(cnt))
        assert rc == 0, (name, G.lib.bsam_last_error())
        cnt2 = ctypes.c_longlong(-1)
        rc = getattr(G.lib, name)(G.h, k, code, float(thr), int(buffer), None, None, ctypes.byref(cnt2))
        assert rc == 0 and cnt2.value == cnt.value, name
        out.append((T, R, cnt.value))
    return out


@pytest.mark.gpu
def test_static_tagging_equals_allocating_tagging():
    from cases import CASES
    from bsamgpu import _lib, tagcluster
    from bsamgpu.device import GpuHierarchy

    # the larger map first (README layout, 128^2 finest), then a smaller one (64^2)
    big = CASES["plain"](seed=3)
    small = CASES["neumann"](seed=5)
    assert np.prod(small.L(small.nlev).n) < np.prod(big.L(big.nlev).n)
    for H in (big, small):
        G = GpuHierarchy(H)
        for k in range(1, H.nlev + 1):
            G.fill_ghosts(k)
        G.pull_q(H)
        L = H.L(H.nlev)
        for code in (_lib.BSAM_TAG_ULAP, _lib.BSAM_TAG_RTE):
            s = np.concatenate([(tagcluster.ulap_indicator(L.Q[p]) if code == _lib.BSAM_TAG_ULAP else
                                 tagcluster.rte_indicator(L.Q[p], L.DeW[p], L.DnS[p], L.CC[p], L.h[0],
                                                          L.h[1])).ravel() for p in range(L.np)])
            for q, buffer in ((0.5, 2), (0.95, 1)):
                thr = float(np.quantile(s, q))
                print(f"levels {H.nlev} indicator {code} q{q} buffer {buffer}", flush=True)
                (Ta, Ra, na), (Tb, Rb, nb) = _tag_both(G, H.nlev, code, thr, buffer)
                assert na == nb and na > 0
                assert np.array_equal(Ta, Tb) and np.array_equal(Ra, Rb)
        G.close()


@pytest.mark.gpu
def test_static_tagging_limits_and_errors():
    from cases import CASES
    from bsamgpu.device import GpuHierarchy

    H = CASES["plain"](seed=1)
    G = GpuHierarchy(H)
    k = H.nlev
    n1, n2 = (int(v) for v in H.L(k).n)
    T = np.zeros(n1 * n2, dtype=np.uint8)
    assert G.lib.bsam_tag_level_static(G.h, k, 7, 0.1, 2, T.ctypes.data, None, None) == -1
    assert G.lib.bsam_tag_level_static(G.h, H.nlev + 1, 0, 0.1, 2, T.ctypes.data, None, None) == -1
    cnt = ctypes.c_longlong(-1)
    assert G.lib.bsam_tag_level_static(G.h, k, 0, np.inf, 2, None, None, ctypes.byref(cnt)) == 0 and cnt.value == 0
    G.close()
