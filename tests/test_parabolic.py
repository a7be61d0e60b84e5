"""Parabolic stepping: the oracle against the published heat log, the host
options layer, and (on the GPU) the device step loop against the oracle.

The device path (bsam_parabolic_step: BDF2/BE right-hand side from the
resident history, ebSolveCycles, history rotation, composite mass) must
reproduce the oracle's per-step residual histories and final solution bit
for bit; masses are compared to 1e-14 relative (parallel sum).
"""
import json
import os

import numpy as np
import pytest

from bsamgpu import _lib
from bsamgpu import parabolic as par
from bsamgpu.problem import BsamError
from oracle import bsam_oracle as O
from parabolic_cases import CASES, Host, diffusion_t_cn, fixed_patch_cn, heat_config, heat_readme, reaction_t_cn

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "published.json")))


def _oracle(case_fn, **kw):
    pp, opts = case_fn(**kw)
    po = O.ParabolicOracle(pp, opts, Host, par)
    snaps = []
    R = po.solve(callback=lambda s, n: snaps.append([[a.copy() for a in s.O.L(k).Q]
                                                     for k in range(1, s.O.nlev + 1)]))
    return po, R, snaps


# ------------------------------------------------------------------ CPU
def test_oracle_reproduces_published_heat_log():
    """README heat example: V-cycles, |r|/|f|, AMR levels, cells and mass at
    steps 10..50 (verification/readme-examples.txt:23-27)."""
    g = GOLD["heat_readme"]
    _, R, _ = _oracle(heat_readme)
    for i, n in enumerate(g["steps"]):
        st = R["stages"][n - 1]
        lev, cells = R["levels"][n - 1]
        assert st["vcycles"] == g["vcycles"][i], (n, st["vcycles"])
        assert float(f"{st['finalRel']:.1e}") == g["rel"][i], (n, st["finalRel"])
        assert lev == g["levels"][i] and cells == g["cells"][i], (n, lev, cells)
        assert f"{R['mass'][n]:.6e}" == f"{g['mass'][i]:.6e}"


def test_oracle_heat_mass_drift_at_tight_tolerance():
    """Zero source, zero reaction, homogeneous Neumann: BE bootstrap + BDF2
    keep the composite mass constant up to the algebraic residual
    (ParabolicBSAM/README.md:118-123); at RelTol 1e-12 per step the drift over
    50 steps and 9 regrids stays below 1e-12 relative."""
    _, R, _ = _oracle(heat_config)
    m = np.array(R["mass"])
    assert np.max(np.abs(m - m[0])) / abs(m[0]) <= 1e-12


def test_oracle_cn_published_regressions():
    """Crank-Nicolson (ebpRhs.m cn branch) against the published regression
    log (verification/suite-output.txt:178-181, runParabolicRegressions.m):
    errors of the time-dependent reaction at 6 digits, Cauchy differences
    with time-dependent D and boundary data at 6 digits, and the composite
    mass of the fixed-patch heat problem (drift below the reference's 2e-10,
    largest per-step residual 1.157e-12 at 4 digits, without and with regrids)."""
    errs = []
    for dt in (0.1, 0.05, 0.025, 0.0125):
        po, _, _ = _oracle(reaction_t_cn, Dt=dt)
        q = po.O.L(po.O.nlev).Q[0][1:-1, 1:-1]
        errs.append(float(np.max(np.abs(q - np.exp(-0.4)))))
    assert [float(f"{e:.6g}") for e in errs] == [0.000359769, 8.98676e-05, 2.24623e-05, 5.61527e-06], errs
    U = []
    for dt in (0.1, 0.05, 0.025, 0.0125):
        po, _, _ = _oracle(diffusion_t_cn, Dt=dt)
        U.append(po.O.L(po.O.nlev).Q[0][1:-1, 1:-1].copy())
    d = [float(np.max(np.abs(U[j] - U[j + 1]))) for j in range(3)]
    assert [float(f"{x:.6g}") for x in d] == [7.6541e-05, 1.80441e-05, 4.53075e-06], d
    for regrid in (0, 2):
        po, R, _ = _oracle(fixed_patch_cn, RegridEvery=regrid)
        assert po.O.nlev - po.O.rootIdx + 1 == 2
        m = np.array(R["mass"])
        drift = float(np.max(np.abs(m - m[0])))
        rel = max(st["finalRel"] for st in R["stages"])
        print(f"cn fixed patch regrid={regrid}: drift {drift:.3e}, max rel {rel:.3e}", flush=True)
        assert drift < 2e-10
        assert f"{rel:.3e}" == "1.157e-12"


def test_oracle_reaction_t_exact_in_space():
    """C = 1 + t, f = t e^-t, u0 = 1: the solution stays spatially constant
    and BDF2 tracks exp(-t) to O(dt^2) (runParabolicRegressions.m:6-20)."""
    errs = []
    for dt in (0.1, 0.05, 0.025):
        po, R, _ = _oracle(reaction_t_case, Dt=dt)
        q = po.O.L(po.O.nlev).Q[0][1:-1, 1:-1]
        assert np.ptp(q) < 1e-10
        errs.append(float(np.max(np.abs(q - np.exp(-0.4)))))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > 1.8, (errs, rates)


def reaction_t_case(**kw):
    return CASES["reaction_t"](**kw)


def test_conservative_prolongations_pinned():
    """conslin / conslin_interior (ebpRegrid.m:95-134) of the oracle: children
    average to the parent exactly in exact arithmetic and a linear field is
    reproduced.  (The device transfer is compared with the oracle's regrid
    by the parabolic GPU tests.)"""
    rng = np.random.default_rng(3)
    W = rng.standard_normal((7, 6))
    for f in (O.conslin,):
        F = f(W)
        avg = 0.25 * (F[0::2, 0::2] + F[1::2, 0::2] + F[0::2, 1::2] + F[1::2, 1::2])
        assert np.allclose(avg, W[1:-1, 1:-1], rtol=0, atol=1e-14)
    x = np.arange(7)[:, None] * 0.5 + np.arange(6)[None, :] * 0.25
    F = O.conslin(x)
    fx = (np.arange(10)[:, None] - 0.5) * 0.25 + 0.5 + ((np.arange(8)[None, :] - 0.5) * 0.125 + 0.25)
    assert np.allclose(F, fx + 0 * F, atol=1e-14)
    P = rng.standard_normal((5, 4))
    F = O.conslin_interior(P)
    avg = 0.25 * (F[0::2, 0::2] + F[1::2, 0::2] + F[0::2, 1::2] + F[1::2, 1::2])
    assert np.allclose(avg, P, rtol=0, atol=1e-14)
    # one-sided slopes at the edges reproduce a linear field too
    lin = np.arange(5)[:, None] * 0.5 + np.arange(4)[None, :] * 0.25
    fl = (np.arange(10)[:, None] - 0.5) * 0.25 + (np.arange(8)[None, :] - 0.5) * 0.125
    assert np.allclose(O.conslin_interior(lin), fl, atol=1e-14)


def test_options_and_problem_errors():
    for bad in (0, -0.1, float("nan"), float("inf")):
        with pytest.raises(BsamError):
            par.parabolic_options(Dt=bad, TFinal=1)
    with pytest.raises(BsamError) as e:
        par.parabolic_options(Dt=0.1)
    assert e.value.ident == "ebp_options:time"
    with pytest.raises(BsamError) as e:
        par.parabolic_options(Dt=0.1, TFinal=1, Scheme="rk4")
    assert e.value.ident == "ebp_options:scheme"
    with pytest.raises(BsamError) as e:
        par.ParabolicProblem([0, 1, 0, 1], 1, 0, 0, ("robin", 0), lambda x, y: x)
    assert e.value.ident == "ebp_problem:bc"
    with pytest.raises(BsamError) as e:
        par.ParabolicProblem([0, 1, 0, 1], lambda x: x, 0, 0, ("neumann", 0), lambda x, y: x)
    assert e.value.ident == "ebp_problem:arity"
    with pytest.raises(BsamError) as e:
        par.ParabolicProblem([0, 1, 0, 1], 1, 0, 0, ("neumann", 0), 3.0)
    assert e.value.ident == "ebp_problem:u0"
    o = par.parabolic_options(Dt=0.01, TFinal=0.1, BaseCells=32, MaxVCyclesStep=7)
    assert o.eb.MaxVCycles == 7 and o.eb.Verbose == 0 and o.eb.BaseCells == (32, 32)
    pp = par.ParabolicProblem([0, 1, 0, 1], 1, lambda x, y, t: t, 0, ("neumann", lambda x, y: 0 * x), lambda x, y: x)
    assert pp.tdC and not pp.tdD and not pp.tdF and not pp.tdBC


# ------------------------------------------------------------------ GPU
def _gpu(case_fn, **kw):
    pp, opts = case_fn(**kw)
    snaps = []
    opts.Callback = lambda S, t, n: snaps.append([S.G.get_field(k, _lib.BSAM_Q) for k in range(1, S.H.nlev + 1)])
    sol = par.ebp_solve(pp, opts)
    return sol, snaps


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_device_steps_match_oracle(name):
    print(f"parabolic case {name}", flush=True)
    po, R, osnaps = _oracle(CASES[name])
    sol, gsnaps = _gpu(CASES[name])
    assert len(sol["stages"]) == len(R["stages"])
    for n, (sg, so) in enumerate(zip(sol["stages"], R["stages"]), 1):
        assert sg["vcycles"] == so["vcycles"], (n, sg["vcycles"], so["vcycles"])
        assert sg["resHist"] == so["resHist"], (n, sg["resHist"], so["resHist"])
        assert sg["totalCells"] == so["totalCells"], n
        for k, (qa, qb) in enumerate(zip(gsnaps[n - 1], osnaps[n - 1]), 1):
            for p, (a, b) in enumerate(zip(qa, qb)):
                if not np.array_equal(a, b):
                    d = np.abs(a - b)
                    raise AssertionError(f"step {n} level {k} patch {p}: max diff {d.max():.3e}")
    mg = np.array(sol["mass"])
    mo = np.array(R["mass"])
    assert np.allclose(mg, mo, rtol=1e-14, atol=0), np.max(np.abs(mg - mo) / np.abs(mo))


@pytest.mark.gpu
def test_device_heat_config_mass_drift():
    """Heat at BaseCells 64 / MaxLevels 4, RelTol 1e-12 per step: composite
    mass drift <= 1e-12 relative over 50 steps with 9 regrids."""
    sol, _ = _gpu(heat_config)
    m = np.array(sol["mass"])
    drift = np.max(np.abs(m - m[0])) / abs(m[0])
    print(f"heat config: {sum(sol['cycles'])} V-cycles, {sol['nregrids']} regrids, drift {drift:.3e}", flush=True)
    assert drift <= 1e-12
    assert sol["nregrids"] == 9


@pytest.mark.gpu
def test_device_heat_readme_published_log():
    g = GOLD["heat_readme"]
    sol, _ = _gpu(heat_readme)
    for i, n in enumerate(g["steps"]):
        st = sol["stages"][n - 1]
        assert st["vcycles"] == g["vcycles"][i]
        assert float(f"{st['finalRel']:.1e}") == g["rel"][i]
        assert sol["levels_hist"][n - 1] == (g["levels"][i], g["cells"][i])


@pytest.mark.gpu
def test_device_temporal_orders():
    """BE ~ first and BDF2 ~ second order in time on a fixed uniform mesh
    (Dt-Cauchy differences, runParabolicTests.m:35-62 at 64^2)."""
    rates = {}
    for scheme in ("be", "bdf2", "cn"):
        U = []
        for dt in (0.04, 0.02, 0.01, 0.005):
            pp, opts = CASES["mms_bump"](Scheme=scheme, Dt=dt, RegridEvery=0, BaseCells=64, MaxLevels=1,
                                         RelTol=1e-10)
            sol = par.ebp_solve(pp, opts)
            H = sol["H"]
            U.append(H.L(H.nlev).Q[0][1:-1, 1:-1].copy())
        d = [float(np.max(np.abs(U[j] - U[j + 1]))) for j in range(len(U) - 1)]
        rates[scheme] = np.log2(np.array(d[:-1]) / np.array(d[1:]))
        print(scheme, d, rates[scheme], flush=True)
    assert rates["be"][-1] > 0.9
    assert rates["bdf2"][-1] > 1.8
    assert rates["cn"][-1] > 1.8


@pytest.mark.gpu
def test_device_cn_published_regressions():
    """Crank-Nicolson on the device against the published regression log
    (verification/suite-output.txt:178-181): time-dependent reaction errors at
    6 digits and the fixed-patch composite mass with regrids (drift below the
    reference's 2e-10, largest per-step residual 1.157e-12 at 4 digits)."""
    errs = []
    for dt in (0.1, 0.05, 0.025, 0.0125):
        sol, _ = _gpu(reaction_t_cn, Dt=dt)
        H = sol["H"]
        errs.append(float(np.max(np.abs(H.L(H.nlev).Q[0][1:-1, 1:-1] - np.exp(-0.4)))))
    print("cn reaction errors", errs, flush=True)
    assert [float(f"{e:.6g}") for e in errs] == [0.000359769, 8.98676e-05, 2.24623e-05, 5.61527e-06], errs
    sol, _ = _gpu(fixed_patch_cn, RegridEvery=2)
    m = np.array(sol["mass"])
    assert float(np.max(np.abs(m - m[0]))) < 2e-10
    assert f"{max(sol['relres']):.3e}" == "1.157e-12"


@pytest.mark.gpu
def test_device_cn_mixed_precision():
    """Crank-Nicolson with FP32 inner cycles (the right side stays FP64): the
    fixed-patch problem with regrids ends within 1e-10 relative L2 of the FP64
    run, with the same V-cycle counts and composite mass drift below 2e-10."""
    out = {}
    for prec in ("fp64", "mixed"):
        sol, _ = _gpu(fixed_patch_cn, RegridEvery=2, Precision=prec)
        H = sol["H"]
        out[prec] = (sol, np.concatenate([H.L(k).Q[p][1:-1, 1:-1].ravel() for k in range(H.rootIdx, H.nlev + 1)
                                          for p in range(H.L(k).np)]))
    (s64, q64), (smx, qmx) = out["fp64"], out["mixed"]
    err = float(np.linalg.norm(qmx - q64) / np.linalg.norm(q64))
    m = np.array(smx["mass"])
    print(f"cn mixed: rel L2 vs fp64 {err:.2e}, cycles {smx['cycles']} vs {s64['cycles']}", flush=True)
    assert err < 1e-10
    assert smx["cycles"] == s64["cycles"]
    assert float(np.max(np.abs(m - m[0]))) < 2e-10


@pytest.mark.gpu
def test_device_history_roundtrip_and_errors():
    pp, opts = heat_readme(TFinal=0.004)
    S = par.ebp_init(pp, opts, 1.0 / 0.002)
    G = S.G
    U = G.get_history(0)
    for k in range(1, S.H.nlev + 1):
        Q = G.get_field(k, _lib.BSAM_Q)
        for p in range(S.H.L(k).np):
            assert np.array_equal(U[k - 1][p], Q[p][1:-1, 1:-1])
    with pytest.raises(BsamError) as e:
        G.parabolic_rhs(0.002, "bdf2")     # no u^{n-1} before the first step
    assert "ebp_rhs:history" in str(e.value)
    with pytest.raises(BsamError):
        G.parabolic_rhs(-1.0, "be")
    with pytest.raises(BsamError) as e:
        G.parabolic_rhs(0.002, "cn")       # the operator's shift was never given
    assert "ebp_rhs:shift" in str(e.value)
    with pytest.raises(BsamError):
        G.parabolic_set_shift(float("nan"))
    V = [[u * 2.0 for u in lev] for lev in U]
    G.set_history(Uo=V, Uoo=U)
    assert all(np.array_equal(a, b) for la, lb in zip(G.get_history(0), V) for a, b in zip(la, lb))
    assert all(np.array_equal(a, b) for la, lb in zip(G.get_history(1), U) for a, b in zip(la, lb))
