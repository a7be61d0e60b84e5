"""Pins of the CPU oracle (no GPU): the reference's published numbers, its
unit-kernel identities, closed forms and brute force.

Each check is chosen so that a plausible slip in the restatement (a dropped
flux-correction term, a swapped ghost index, a wrong sign in a ghost formula,
a transposed operand) fails it.
"""
import json
import os

import numpy as np
import pytest

from bsamgpu import Options, Problem, configs, hierarchy, solve, tagcluster
from oracle import bsam_oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "published.json")))


class _Host:
    setup = staticmethod(hierarchy.setup)
    add_level = staticmethod(hierarchy.add_level)
    tag_cluster = staticmethod(tagcluster.tag_cluster)


def test_readme_elliptic_stage_log():
    """verification/readme-examples.txt:2-6 -- three stages, 6 V-cycles each,
    published rel / factor / visible cells, and the added patch layouts."""
    opts = configs.plain_options(RelTol=1e-9)
    H, Orc = O.eb_solve(configs.plain_problem(), opts, _Host)
    ref = GOLD["readme_elliptic"]
    assert len(Orc.stats["stages"]) == 3
    for st, want in zip(Orc.stats["stages"], ref["stages"]):
        assert st["amrLevels"] == want["amrLevels"]
        assert st["vcycles"] == want["vcycles"]
        assert f"{st['finalRel']:.2e}" == want["finalRel"]
        assert f"{st['avgFactor']:.3f}" == want["avgFactor"]
        assert st["visibleCells"] == want["visibleCells"]
    for k, want in zip(range(H.rootIdx + 1, H.nlev + 1), ref["added_levels"]):
        L = H.L(k)
        assert L.np == want["np"]
        assert sum(np.prod(L.dims(p)) for p in range(L.np)) == want["cells"]
    # T7-style discrete conservation at the solved fixed point
    assert Orc.mass_check()["rel"] < GOLD["t7_mass"]["with_corrections_rel_max"]


def test_mass_defect_without_corrections_is_large():
    """suite-output.txt:64-68: without flux-balance corrections the composite
    conservation identity fails at O(h); with them it holds to round-off."""
    from cases import lshape

    H = lshape(MassCorrection=False)
    Orc = O.OracleHierarchy(H)
    Orc.solve_cycles()
    assert Orc.mass_check()["rel"] > 1e-8
    H.opts = H.opts.replace(MassCorrection=True)
    Orc = O.OracleHierarchy(H)
    Orc.solve_cycles()
    assert Orc.mass_check()["rel"] < 1e-12


@pytest.fixture(scope="module")
def starfish():
    prob, aux, g = configs.starfish_problem()
    opts = configs.starfish_options(g)
    return solve.build_hierarchy(prob, opts), g


def test_starfish_hierarchy_counts(starfish):
    """verification/starfish-full.json:13-37 -- patches and cells per level."""
    H, _ = starfish
    ref = GOLD["starfish_full"]
    assert [H.L(k).np for k in range(H.rootIdx, H.nlev + 1)] == ref["patchesPerLevel"]
    cells = [sum(int(np.prod(H.L(k).dims(p))) for p in range(H.L(k).np)) for k in range(H.rootIdx, H.nlev + 1)]
    assert cells == ref["cellsPerLevel"]
    assert sum(cells) == ref["storedAmrCells"]
    vis = sum(int((~c).sum()) for k in range(H.rootIdx, H.nlev + 1) for c in H.L(k).covered)
    assert vis == ref["visibleCells"]


@pytest.mark.slow
def test_starfish_residual_history(starfish):
    """starfish-full-residual.csv:2-12: the oracle's FP64 cycle sequence
    agrees with MATLAB's to round-off-limited precision (early cycles to
    1e-9 relative; the last cycles sit a few ulps-of-L*u above the floor)."""
    H, _ = starfish
    Orc = O.OracleHierarchy(H)
    st = Orc.solve_cycles()
    ref = GOLD["starfish_full"]
    assert st["vcycles"] == ref["vcycles"]
    hist = np.array(st["resHist"])
    want = np.array(ref["residualHistory"])
    rel = np.abs(hist - want) / want
    assert np.all(rel[:4] < 1e-8), rel
    assert np.all(rel[4:8] < 1e-3), rel
    assert np.all(rel < 5e-2), rel
    assert abs(st["finalRel"] - ref["relativeResidual"]) / ref["relativeResidual"] < 5e-2
    assert f"{st['avgFactor']:.3f}" == ref["avgFactor"]
    assert Orc.mass_check()["rel"] < 1e-13


# ------------------------------------------------------------- kernel KATs
@pytest.mark.parametrize("method", ["qif", "lif"])
def test_interp_reproduces_polynomials(method):
    """unitKernels.m:12-35: QIF reproduces quadratics, LIF linears (<=1e-12)."""
    if method == "qif":
        u = lambda x, y: 1.5 + 0.3 * x - 0.7 * y + 0.25 * x ** 2 - 0.4 * x * y + 0.6 * y ** 2  # noqa: E731
    else:
        u = lambda x, y: 1.5 + 0.3 * x - 0.7 * y  # noqa: E731
    mx, my, h, x0, y0 = 8, 12, 0.05, 0.3, -0.2
    XF, YF = np.meshgrid(x0 + (np.arange(mx + 2) - 0.5) * h, y0 + (np.arange(my + 2) - 0.5) * h, indexing="ij")
    Qex = u(XF, YF)
    Q = Qex.copy()
    Q[0, :] = Q[-1, :] = Q[:, 0] = Q[:, -1] = 99
    cmx, cmy = mx // 2, my // 2
    XC, YC = np.meshgrid(x0 + (np.arange(cmx + 2) - 1) * 2 * h + h, y0 + (np.arange(cmy + 2) - 1) * 2 * h + h,
                         indexing="ij")
    QC = u(XC, YC)
    O.interp_cf(Q, QC, [True] * 4, [True] * 4, method == "qif")
    ghost = np.zeros(Q.shape, bool)
    ghost[[0, -1], :] = True
    ghost[:, [0, -1]] = True
    assert np.max(np.abs(Q[ghost] - Qex[ghost])) <= GOLD["unit_kernels"]["qif_ghost_tol"]
    if method == "qif":  # a quadratic must defeat LIF (the test has teeth)
        Q2 = Qex.copy()
        O.interp_cf(Q2, QC, [True] * 4, [True] * 4, False)
        assert np.max(np.abs(Q2[ghost] - Qex[ghost])) > 1e-6


def test_operator_second_order():
    """unitKernels.m:37-66: variable-D/C operator truncation rate >= 1.8."""
    uex = lambda x, y: np.sin(np.pi * x) * np.cos(np.pi * y)  # noqa: E731
    Df = lambda x, y: 1 + 0.5 * np.sin(2 * np.pi * x) * np.cos(np.pi * y)  # noqa: E731
    Cf = lambda x, y: 1 + x * y  # noqa: E731
    Dx = lambda x, y: np.pi * np.cos(2 * np.pi * x) * np.cos(np.pi * y)  # noqa: E731
    Dy = lambda x, y: -0.5 * np.pi * np.sin(2 * np.pi * x) * np.sin(np.pi * y)  # noqa: E731
    ux = lambda x, y: np.pi * np.cos(np.pi * x) * np.cos(np.pi * y)  # noqa: E731
    uy = lambda x, y: -np.pi * np.sin(np.pi * x) * np.sin(np.pi * y)  # noqa: E731
    ff = lambda x, y: -(Dx(x, y) * ux(x, y) + Dy(x, y) * uy(x, y) + Df(x, y) * (-2 * np.pi ** 2 * uex(x, y))) \
        + Cf(x, y) * uex(x, y)  # noqa: E731
    errs = []
    for n in (32, 64, 128):
        h = 1 / n
        XF, YC = np.meshgrid(np.arange(n + 1) * h, (np.arange(n) + 0.5) * h, indexing="ij")
        XC, YF = np.meshgrid((np.arange(n) + 0.5) * h, np.arange(n + 1) * h, indexing="ij")
        X, Y = np.meshgrid((np.arange(n + 2) - 0.5) * h, (np.arange(n + 2) - 0.5) * h, indexing="ij")
        L = O.op2d(uex(X, Y), Df(XF, YC), Df(XC, YF), Cf(X[1:-1, 1:-1], Y[1:-1, 1:-1]), h, h)
        errs.append(np.max(np.abs(L - ff(X[1:-1, 1:-1], Y[1:-1, 1:-1]))))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates >= GOLD["unit_kernels"]["op_rate_min"]), rates


def test_restrict_prolong_affine():
    """unitKernels.m:68-78: restrict(prolong(affine)) == affine, and
    prolongation is exact on affine data."""
    II, JJ = np.meshgrid(np.arange(6), np.arange(8), indexing="ij")
    W = 0.3 + 0.2 * II - 0.45 * JJ
    Fp = O.prolong_bilin(W)
    assert np.max(np.abs(O.restrict2(Fp) - W[1:-1, 1:-1])) <= GOLD["unit_kernels"]["restrict_prolong_tol"]
    fi, fj = np.meshgrid(np.arange(8) * 0.5 + 0.75, np.arange(12) * 0.5 + 0.75, indexing="ij")
    assert np.max(np.abs(Fp - (0.3 + 0.2 * fi - 0.45 * fj))) < 1e-13


def test_uniform_vcycle_matches_direct():
    """unitKernels.m:80-89: uniform 64^2 V-cycles reach rel <= 1e-11 and the
    discretisation error of the manufactured solution."""
    uex = lambda x, y: np.sin(np.pi * x) * np.cos(np.pi * y)  # noqa: E731
    prob = Problem([0, 1, 0, 1], 1, 1, lambda x, y: (1 + 2 * np.pi ** 2) * uex(x, y), ("dirichlet", uex), uex)
    opts = Options(BaseCells=64, MaxLevels=1, Verbose=0, RelTol=1e-11)
    H = hierarchy.setup(prob, opts)
    Orc = O.OracleHierarchy(H)
    st = Orc.solve_cycles()
    assert st["finalRel"] <= 1e-11
    XC, YC = H.L(H.nlev).cell_centers(prob.domain, 0)
    assert np.max(np.abs(Orc.L(Orc.nlev).Q[0][1:-1, 1:-1] - uex(XC, YC))) < 5e-3


def _bc_problem():
    D = lambda x, y: 1 + 0.3 * np.cos(3 * x + y)  # noqa: E731
    bc = {"west": ("dirichlet", lambda x, y: 1 + y), "east": ("neumann", lambda x, y: np.sin(y)),
          "south": ("dirichlet", 0.5), "north": ("neumann", -0.25)}
    return Problem([0, 1, 0, 2], D, lambda x, y: 1 + x, lambda x, y: x * y, bc)


def test_coarsest_matrix_is_ghost_eliminated_operator():
    """ebCoarsest.m:14-109 vs brute force: A u - bcRhs equals op2d(u) with the
    physical ghosts of ebFillGhosts.m:53-83, for random u; the explicit
    inverse satisfies A A^-1 = I; the folded diagonal (ebBcFold.m) equals
    1/diag(A)."""
    H = hierarchy.setup(_bc_problem(), Options(BaseCells=(16, 32), Verbose=0))
    Orc = O.OracleHierarchy(H)
    L1 = Orc.L(1)
    A = Orc.coarsest["A"]
    bcRhs = Orc.coarsest["bcRhs"]
    rng = np.random.default_rng(0)
    for _ in range(3):
        u = rng.standard_normal(A.shape[0])
        L1.Q[0][1:-1, 1:-1] = u.reshape(L1.n, order="F")
        Orc.fill_ghosts(1)
        Lu = O.op2d(L1.Q[0], L1.DeW[0], L1.DnS[0], L1.CC[0], *L1.h).reshape(-1, order="F")
        assert np.max(np.abs(A @ u - bcRhs - Lu)) < 1e-10 * np.max(np.abs(Lu))
    Ainv = Orc.coarsest["Ainv"]
    assert np.max(np.abs(A @ Ainv - np.eye(A.shape[0]))) < 1e-11
    fold = 1.0 / L1.InvDiagS[0].reshape(-1, order="F")
    assert np.max(np.abs(fold - np.diag(A)) / np.abs(np.diag(A))) < 1e-14


def test_singular_kkt_pins_mean():
    """ebCoarsest.m:59-90: all-Neumann, C = 0 uses the KKT system; the solve
    returns a zero-mean exact solution of the compatible system."""
    f = lambda x, y: np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)  # noqa: E731
    H = hierarchy.setup(Problem([0, 1, 0, 1], 1, 0, f, ("neumann", 0)), Options(BaseCells=16, Verbose=0))
    Orc = O.OracleHierarchy(H)
    assert Orc.coarsest["singular"]
    L1 = Orc.L(1)
    Orc.coarsest_solve()
    u = L1.Q[0][1:-1, 1:-1]
    assert abs(u.mean()) < 1e-13
    Lu = O.op2d(L1.Q[0], L1.DeW[0], L1.DnS[0], L1.CC[0], *L1.h)
    assert np.max(np.abs(Lu - L1.F[0])) < 1e-10 * np.max(np.abs(L1.F[0]))


def test_relax_solves_own_equation():
    """relax2d (ebKernels.m:37-95) brute force: after a colour pass every
    updated cell satisfies its own ghost-eliminated equation exactly (the
    boundary fold of ebBcFold.m makes physical-boundary updates exact)."""
    H = hierarchy.setup(_bc_problem(), Options(BaseCells=(16, 32), Verbose=0))
    Orc = O.OracleHierarchy(H)
    L1 = Orc.L(1)
    rng = np.random.default_rng(4)
    L1.Q[0][1:-1, 1:-1] = rng.standard_normal(L1.n)
    Orc.fill_ghosts(1)
    for color in (1, 2):
        Q = L1.Q[0]
        O.relax2d(Q, L1.F[0], L1.DeW[0], L1.DnS[0], L1.InvDiagS[0], L1.bc_code[0], L1.bc_g[0], L1.par0[0],
                  color, L1.h[0], L1.h[1], 1.0)
        Orc.fill_ghosts(1)
        r = O.op2d(L1.Q[0], L1.DeW[0], L1.DnS[0], L1.CC[0], *L1.h) - L1.F[0]
        mask = O.colour_mask(*L1.n, L1.par0[0], color)
        assert np.max(np.abs(r[mask])) < 1e-10 * np.max(np.abs(L1.F[0]) + 1)
        assert np.max(np.abs(r[~mask])) > 1e-6


def test_starfish_distance_brute_force(starfish):
    """tests/testStarfish.m:5-16: tree distance == brute force (5e-14), sign
    == point-in-polygon, zero on the vertices."""
    _, g = starfish
    rng = np.random.default_rng(1127)
    X = 4 * rng.random(300) - 2
    Y = 4 * rng.random(300) - 2
    v = g.vertices
    brute = np.full(X.shape, np.inf)
    for k in range(v.shape[0] - 1):
        a = v[k]
        e = v[k + 1] - a
        t = np.clip(((X - a[0]) * e[0] + (Y - a[1]) * e[1]) / np.sum(e ** 2), 0, 1)
        brute = np.minimum(brute, np.hypot(X - a[0] - t * e[0], Y - a[1] - t * e[1]))
    d = g.signed_distance(X, Y)
    assert np.max(np.abs(np.abs(d) - brute)) < 5e-14
    inside = np.zeros(X.shape, bool)  # even-odd ray casting
    for k in range(v.shape[0] - 1):
        (x1, y1), (x2, y2) = v[k], v[k + 1]
        cross = ((y1 > Y) != (y2 > Y)) & (X < (x2 - x1) * (Y - y1) / (y2 - y1 + 1e-300) + x1)
        inside ^= cross
    assert np.all((d >= 0) == inside)
    assert g.signed_distance(0.0, 0.0) > 0 and g.signed_distance(2.0, 2.0) < 0
    assert np.max(np.abs(g.signed_distance(v[:, 0], v[:, 1]))) < 1e-13
