"""The library's native hierarchy construction (bsam_build_*, host C++)
against the host layer's (hierarchy.setup / add_level, which restate
EllipticBSAM/src/ebSetup.m, ebMakeLevel.m, ebCoeffFilldown.m, ebAddLevel.m).

Stated tolerance: bit-identical descriptors (boxes, bc codes, Q, FTrue, DeW,
DnS, CC, InvDiagS, g) when both sample the same coefficient callables: after
setup and after every added level, on the README elliptic layout, variable D
and C with mixed Dirichlet/Neumann data, non-square and odd base sizes, an
initial guess, and served-workload requests.  Errors carry the reference's
identifiers.  The erode8 coverage map equals the host layer's.
"""
import numpy as np
import pytest

from bsamgpu import Options, Problem, configs, hierarchy, tagcluster
from bsamgpu.build import NativeBuild
from bsamgpu.device import pack_hierarchy
from bsamgpu.problem import BsamError
from cases import _vardc_problem


def _same(nb, H):
    d = nb.desc()
    ref = pack_hierarchy(H)
    assert d["nlev"] == H.nlev and d["root_idx"] == H.rootIdx
    for key in ("n", "np", "boxes", "bc"):
        assert np.array_equal(d[key], ref[key]), key
    for key in ("Q", "FTrue", "DeW", "DnS", "CC", "ID", "g"):
        assert d[key].shape == ref[key].shape, key
        assert np.array_equal(d[key], ref[key]), (key, np.max(np.abs(d[key] - ref[key])))
    nl, root, sing = nb.info()
    assert (nl, root, sing) == (H.nlev, H.rootIdx, H.singular)
    L = H.L(H.nlev)
    cover = np.zeros(L.n, dtype=bool)
    for b in L.box:
        cover[b[0] - 1:b[1], b[2] - 1:b[3]] = True
    assert np.array_equal(nb.allowed(), tagcluster.erode8(cover))


def _both(prob, opts, levels):
    H = hierarchy.setup(prob, opts)
    nb = NativeBuild(prob, opts)
    _same(nb, H)
    for boxes in levels:
        hierarchy.add_level(H, boxes)
        nb.add_level(boxes)
        _same(nb, H)
    nb.close()


def test_readme_layout():
    _both(configs.plain_problem(), configs.plain_options(),
          [[[1, 32, 1, 32]], [[7, 26, 7, 26], [7, 26, 39, 58], [39, 58, 7, 26], [39, 58, 39, 58]]])


def test_variable_coefficients_mixed_sides_and_initial_guess():
    bc = {"west": ("dirichlet", lambda x, y: np.sin(3 * y)), "east": ("neumann", 0.5),
          "south": ("neumann", lambda x, y: x), "north": ("dirichlet", 1.0)}
    opts = Options(BaseCells=32, MaxLevels=3, Verbose=0, InitialGuess=lambda x, y: x * y + 0.25)
    _both(_vardc_problem(bc), opts,
          [[[5, 12, 5, 8], [5, 8, 9, 12], [13, 16, 5, 6], [1, 4, 25, 32], [27, 32, 13, 20]],
           [[10, 15, 10, 22], [16, 23, 10, 15], [1, 6, 51, 62], [55, 62, 27, 38]]])


@pytest.mark.parametrize("base", [(48, 24), (40, 40), (12, 20)])
def test_ladders_of_other_base_sizes(base):
    prob = _vardc_problem(("dirichlet", lambda x, y: x - 2 * y))
    _both(prob, Options(BaseCells=base, MaxLevels=2, Verbose=0), [[[3, 8, 1, 4]]])


def test_served_requests():
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "config5_requests.json")) as f:
        reqs = json.load(f)["requests"]
    for par in reqs[:6]:
        prob, opts = configs.served_problem(par), configs.served_options(par)
        n = hierarchy.ladder_sizes(opts)[-1]
        _both(prob, opts, [[[3, int(n[0]) // 2, 5, int(n[1]) // 2 + 2]]])


def test_singular_neumann_projection():
    f = lambda x, y: np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.3  # noqa: E731
    prob = Problem([0, 1, 0, 1], 1, 0, f, ("neumann", 0))
    opts = Options(BaseCells=32, MaxLevels=2, Verbose=0)
    H = hierarchy.setup(prob, opts)
    nb = NativeBuild(prob, opts)
    assert nb.info()[2] and H.singular
    d = nb.desc()
    ref = pack_hierarchy(H)
    # the projection's sums run in another order than NumPy's pairwise sums
    assert np.allclose(d["FTrue"], ref["FTrue"], rtol=0, atol=1e-13)
    assert np.array_equal(d["DeW"], ref["DeW"]) and np.array_equal(d["ID"], ref["ID"])
    nb.close()


def test_errors_carry_reference_identifiers():
    prob = configs.plain_problem()
    nb = NativeBuild(prob, configs.plain_options())
    with pytest.raises(BsamError) as e:
        nb.add_level([[1, 40, 1, 8]])  # beyond the 32 x 32 root
    assert "eb_make_level:box" in str(e.value)
    nb.close()
    bad = Problem([0, 1, 0, 1], lambda x, y: x - 0.5, 1, 1, ("dirichlet", 0))
    with pytest.raises(BsamError) as e:
        NativeBuild(bad, Options(BaseCells=16, Verbose=0))
    assert "eb_make_level:D" in str(e.value)
    inf = Problem([0, 1, 0, 1], 1, lambda x, y: 1 / (x - x), 1, ("dirichlet", 0))
    with pytest.raises(BsamError) as e:
        NativeBuild(inf, Options(BaseCells=16, Verbose=0))
    assert "eb_make_level:eval" in str(e.value)
