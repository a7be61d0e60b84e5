"""Host-layer logic (no GPU): options/problem validation, ladder, tagging
morphology, Berger-Rigoutsos splitting, level construction errors, and the
C ABI library's exports and host-only table derivation."""
import copy
import ctypes
import os
import re

import numpy as np
import pytest

from bsamgpu import BsamError, Options, Problem, _lib, configs, hierarchy, tagcluster
from bsamgpu.device import GpuHierarchy
from cases import CASES
from oracle import bsam_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_options_defaults_and_validation():
    o = Options()
    assert (o.NPre, o.NPost, o.Omega, o.RelTol, o.MaxVCycles) == (3, 3, 1.0, 1e-9, 60)
    assert o.BaseCells == (64, 64) and o.GhostInterp == "qif" and o.MassCorrection
    assert Options(BaseCells=32).BaseCells == (32, 32)
    with pytest.raises(BsamError):
        Options(GhostInterp="cubic")


def test_problem_validation():
    with pytest.raises(BsamError):
        Problem([0, 1, 1, 0], 1, 1, 0, ("dirichlet", 0))
    with pytest.raises(BsamError):
        Problem([0, 1, 0, 1], 1, 1, 0, {"west": ("dirichlet", 0)})
    with pytest.raises(BsamError):
        Problem([0, 1, 0, 1], 1, 1, 0, ("robin", 0))


def test_ladder_sizes():
    """ebSetup.m:17-25: the coarsening ladder below each base size."""
    for base, ladder in ((32, [16, 32]), (64, [16, 32, 64]), (128, [16, 32, 64, 128]), (48, [12, 24, 48]),
                         (33, [33])):
        got = [int(s[0]) for s in hierarchy.ladder_sizes(Options(BaseCells=base))]
        assert got == ladder, (base, got)


def test_make_level_errors():
    prob = Problem([0, 1, 0, 1], lambda x, y: x - 0.5, 1, 0, ("dirichlet", 0))
    with pytest.raises(BsamError) as ei:
        hierarchy.setup(prob, Options(BaseCells=16))
    assert ei.value.ident == "eb_make_level:D"
    good = configs.plain_problem()
    with pytest.raises(BsamError) as ei:
        hierarchy.make_level(good, Options(), 1, (16, 16), [[0, 16, 1, 16]])
    assert ei.value.ident == "eb_make_level:box"


def test_dilate_erode():
    M = np.zeros((9, 9), bool)
    M[4, 4] = True
    D = tagcluster.dilate(M, 2)
    assert D.sum() == 25 and D[2:7, 2:7].all()
    C = np.ones((6, 6), bool)
    C[0, 3] = False
    A = tagcluster.erode8(C)
    assert not A[0:2, 2:5].any() and A[3:, :].all()


def test_pick_split():
    sig = np.array([3, 3, 3, 0, 0, 3, 3, 3], float)
    assert tagcluster.pick_split(sig, 2) == (4, 2)  # hole nearest the middle
    sig = np.array([1, 2, 3, 4, 4, 3, 6, 8, 8, 8], float)
    s, q = tagcluster.pick_split(sig, 3)
    assert q == 1 and 3 <= s <= 7


def test_br_cluster_two_blobs():
    T = np.zeros((40, 40), bool)
    T[4:10, 5:12] = True
    T[25:33, 20:30] = True
    boxes = tagcluster.br_cluster(T, 4, 0.7)
    cover = np.zeros_like(T)
    for b in boxes:
        cover[b[0] - 1:b[1], b[2] - 1:b[3]] = True
    assert (cover & T).sum() == T.sum()
    assert len(boxes) == 2


def _header_functions():
    src = open(os.path.join(ROOT, "include", "bsamgpu.h")).read()
    return sorted(set(re.findall(r"\b(bsam_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header_symbols():
    lib = _lib.load()
    names = _header_functions()
    assert names, "no declarations parsed"
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTS) == names
    assert b"gfx950" in lib.bsam_version()


def _host_only(H):
    return GpuHierarchy(H, device=-1)


@pytest.mark.parametrize("name", ["plain", "lshape", "neumann", "mixed_bc"])
def test_host_tables_match_oracle_descriptors(name):
    """Ghost classes (ebBuildOps.m:47-113 land/cornerClass) derived by the
    library's host code equal the oracle's restatement."""
    H = CASES[name](seed=None)
    Orc = O.OracleHierarchy(H)
    G = _host_only(H)
    for k in range(2, Orc.nlev + 1):
        L = Orc.L(k)
        ref = []
        for p in range(L.np):
            for s in "WESN":
                ref += [int(v) for v in L.land[p][s]]
            ref += [int(c) for c in L.cornerClass[p]]
        assert G.export(k, 0).tolist() == ref


@pytest.mark.parametrize("name", sorted(CASES))
def test_face_records_have_the_neighbour_assembled_layout(name):
    """Every level of the test layouts qualifies for the smoother passes that
    assemble coarse-fine ghosts from neighbours and read pushed sibling
    ghosts (the library checks each face record's stencil and push target)."""
    H = CASES[name](seed=None)
    G = _host_only(H)
    assert [int(G.export(k, 3)[0]) for k in range(1, H.nlev + 1)] == [1] * H.nlev


def test_starfish_face_records_have_the_neighbour_assembled_layout():
    from bsamgpu import solve

    prob, aux, g = configs.starfish_problem()
    H = solve.build_hierarchy(prob, configs.starfish_options(g))
    G = _host_only(H)
    assert [int(G.export(k, 3)[0]) for k in range(1, H.nlev + 1)] == [1] * H.nlev


def test_host_only_handle_rejects_device_calls():
    G = _host_only(CASES["plain"](seed=None))
    with pytest.raises(BsamError):
        G.fill_ghosts(1)


def test_abi_error_codes_match_reference_identifiers():
    H = CASES["plain"](seed=None)
    H.L(4).box[1] = [14, 53, 77, 116]
    with pytest.raises(BsamError) as ei:
        _host_only(H)
    assert ei.value.ident == "eb_build_ops:align"
    # same patch size (the packed arrays keep their length), moved so its
    # footprint leaves the parent level's union -> proper nesting violated
    H2 = CASES["lshape"](seed=None)
    b = H2.L(4).box[0]
    H2.L(4).box[0] = [3, 3 + b[1] - b[0], b[2], b[3]]
    with pytest.raises(BsamError) as ei:
        _host_only(H2)
    assert ei.value.ident in ("eb_build_ops:ring", "eb_build_ops:nesting")


def test_update_coeffs_validates_structure_and_values():
    """bsam_update_coeffs (the ebpSetTime refresh) accepts new coefficient
    values on the same grid and rejects structural changes and non-positive
    D with the reference's identifiers, leaving the hierarchy unchanged."""
    H = CASES["lshape"](seed=None)
    G = _host_only(H)
    for k in range(1, H.nlev + 1):
        for p in range(H.L(k).np):
            H.L(k).CC[p] = H.L(k).CC[p] + 2.0
    hierarchy.coeff_filldown(H)
    G.update_coeffs(H)
    bad = copy.deepcopy(H)
    bad.L(3).DnS[0][0, 0] = -1.0
    with pytest.raises(BsamError) as e:
        G.update_coeffs(bad)
    assert "eb_make_level:D" in str(e.value)
    bad = copy.deepcopy(H)
    bad.L(2).CC[0][1, 1] = float("nan")
    with pytest.raises(BsamError) as e:
        G.update_coeffs(bad)
    assert "eb_make_level:eval" in str(e.value)
    bad = copy.deepcopy(H)
    bad.L(2).bc_code[0] = [2, 2, 2, 2]
    with pytest.raises(BsamError) as e:
        G.update_coeffs(bad)
    assert "boundary type" in str(e.value)


@pytest.mark.parametrize("name", list(CASES))
def test_host_coarsest_inverse_equals_oracle(name):
    """The library's coarsest inverse (bsam_lu.cpp, host code) equals the
    NumPy oracle's lu_inverse bit for bit: regular and KKT (all-Neumann)
    systems of order 256 / 257."""
    H = CASES[name](seed=None)
    Hh = GpuHierarchy(H, device=-1)
    got = Hh.export(1, 6)
    Hh.close()
    ref = np.asfortranarray(O.OracleHierarchy(H).coarsest["Ainv"]).ravel(order="F").view(np.int64)
    assert got.size == ref.size and np.array_equal(got, ref)
