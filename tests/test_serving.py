"""Served workload: seeded request fixture, and (GPU) the continuous-batching
loop against the CPU oracle and against one-at-a-time solves.

Stated tolerances:
- FP64: bit-identical.  All 512 requests of the served workload, served
  together, give the oracle's boxes, V-cycle counts, residual histories and
  Q (per-level SHA-256 digests of the padded arrays, tests/golden/
  config5_oracle.json, written by tests/golden/make_config5_oracle.py); a
  request's result is the same whether it is served alone or in the mix.
- FP32-inner: every request converges to its RelTol 1e-8; the visible-cell
  relative L2 difference from the FP64 (= oracle) solution is reported and
  must be below 1e-6 wherever both forms built the same hierarchy; alone and
  in the mix the result is bit-identical.
"""
import json
import os
import time

import numpy as np
import pytest

import config5
from bsamgpu import configs, serving

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _fixture():
    with open(os.path.join(GOLD, "config5_requests.json")) as f:
        return json.load(f)


def test_request_fixture_matches_generator():
    d = _fixture()
    assert d["requests"] == configs.served_params(512, seed=5)
    assert np.allclose(d["arrivals_at_rate_1"], configs.served_arrivals(512, 1.0, seed=55), rtol=0, atol=0)
    for r in d["requests"]:
        assert r["base"] in (32, 64, 128) and 2 <= r["levels"] <= 5
        assert sum(abs(a) for a in r["a"]) <= 0.5 and 0 <= r["C"] <= 2
        assert all(b in ("dirichlet", "neumann") for b in r["bc"])


def test_config2_fixture_matches_generator():
    with open(os.path.join(GOLD, "config2_params.json")) as f:
        d = json.load(f)
    ref = configs.varcoef_params(2)
    assert all(d[k] == ref[k] for k in ref)


def test_request_stopping_rule():
    """ebSolveCycles.m:24-41: stop at rel <= RelTol, at the cycle cap, or on
    a non-finite residual."""
    o = configs.served_options({"base": 32, "levels": 2}, MaxVCycles=3)
    r = serving.Request(0, None, o)
    assert not r.take({"linf": 1.0, "rel": 1.0, "ncell": 10})
    assert not r.take({"linf": 0.5, "rel": 0.5, "ncell": 10})
    assert not r.take({"linf": 0.1, "rel": 0.1, "ncell": 10})
    assert r.take({"linf": 0.01, "rel": 0.01, "ncell": 10})  # third cycle: cap
    assert r.it == 3 and r.hist == [1.0, 0.5, 0.1, 0.01]
    r = serving.Request(1, None, o)
    assert r.take({"linf": 1e-12, "rel": 1e-12, "ncell": 10})  # converged before any cycle
    r = serving.Request(2, None, o)
    r.take({"linf": 1.0, "rel": 1.0, "ncell": 10})
    assert r.take({"linf": float("nan"), "rel": float("nan"), "ncell": 10})


def _requests(ids, rate=None):
    d = _fixture()
    arr = d["arrivals_at_rate_1"]
    out = []
    for i in ids:
        par = d["requests"][i]
        t = arr[i] / rate if rate else 0.0
        out.append(serving.Request(i, configs.served_problem(par), configs.served_options(par), t))
    return out


# small and mid-size requests of every kind: Dirichlet/Neumann mixes, 1-3 stages
PICK = [0, 2, 3, 6, 8, 9, 11, 12, 17, 20, 25, 31]


def _same(a, b):
    assert len(a.stages) == len(b.stages), (a.id, len(a.stages), len(b.stages))
    for sa, sb in zip(a.stages, b.stages):
        assert sa["vcycles"] == sb["vcycles"] and sa["resHist"] == sb["resHist"], a.id
    for qa, qb in zip(a.Q, b.Q):
        for x, y in zip(qa, qb):
            assert np.array_equal(x, y), a.id


@pytest.mark.gpu
def test_batched_fp64_equals_plain_solves():
    reqs = _requests(PICK)
    srv = serving.Server(slots=64)
    done = {r.id: r for r in srv.run(reqs, realtime=False)}
    srv.close()
    for r in reqs:
        assert done[r.id].error is None, done[r.id].error
        stages, Q = serving.serial_solve(r)
        print(f"request {r.id}: {len(stages)} stage(s), cycles {[s['vcycles'] for s in stages]}", flush=True)
        assert [s["vcycles"] for s in stages] == [s["vcycles"] for s in done[r.id].stages]
        assert [s["resHist"] for s in stages] == [s["resHist"] for s in done[r.id].stages]
        for qa, qb in zip(Q, done[r.id].Q):
            for x, y in zip(qa, qb):
                assert np.array_equal(x, y), r.id


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_result_is_the_same_alone_and_in_the_mix(precision):
    alone = {}
    for r in _requests(PICK[:6]):
        srv = serving.Server(slots=4, precision=precision)
        alone[r.id] = srv.run([r], realtime=False)[0]
        srv.close()
    srv = serving.Server(slots=64, precision=precision)
    mix = srv.run(_requests(PICK), realtime=False)
    srv.close()
    for r in mix:
        assert r.error is None, r.error
        if r.id in alone:
            _same(r, alone[r.id])
        assert r.stages[-1]["finalRel"] <= r.opts.RelTol


@pytest.mark.gpu
def test_mixed_precision_serving_accuracy():
    """FP32-inner served solves against the FP64 ones: same converged
    solution to the requests' tolerance (1e-8 relative residual)."""
    reqs64 = _requests(PICK[:8])
    srv = serving.Server(slots=16, precision="fp64")
    r64 = {r.id: r for r in srv.run(reqs64, realtime=False)}
    srv.close()
    srv = serving.Server(slots=16, precision="mixed")
    rmx = {r.id: r for r in srv.run(_requests(PICK[:8]), realtime=False)}
    srv.close()
    for i, a in r64.items():
        b = rmx[i]
        if len(a.Q) != len(b.Q):
            continue  # a different hierarchy (tagging saw a different stage solution)
        num = sum(float(np.sum((x - y)[1:-1, 1:-1] ** 2)) for qa, qb in zip(a.Q, b.Q) for x, y in zip(qa, qb))
        den = sum(float(np.sum(x[1:-1, 1:-1] ** 2)) for qa in a.Q for x in qa)
        print(f"request {i}: rel L2 mixed vs fp64 {np.sqrt(num / den):.2e}", flush=True)
        assert np.sqrt(num / den) < 1e-6


@pytest.mark.gpu
def test_realtime_poisson_arrivals():
    reqs = _requests(range(24), rate=200.0)
    srv = serving.Server(slots=32, mem_budget=1 << 30)
    done = srv.run(reqs, realtime=True)
    srv.close()
    assert len(done) == 24 and all(r.error is None for r in done)
    assert all(r.t_done >= r.arrival and r.t_admit >= r.arrival - 1e-9 for r in done)
    st = serving.latency_stats(done)
    print(st, f"ticks {srv.ticks}, mean batch {srv.tick_items / max(srv.ticks, 1):.1f}", flush=True)
    assert st["p99_ms"] >= st["p50_ms"] > 0


@pytest.mark.gpu
def test_parabolic_requests_in_the_mix_match_ebp_solve():
    """Parabolic requests (one stage per time step, BDF2 history resident;
    Crank-Nicolson with time-dependent coefficients and with regrids) served
    together with elliptic ones give exactly what ebp_solve gives."""
    from bsamgpu import parabolic as par
    from parabolic_cases import diffusion_t_cn, fixed_patch_cn, heat_readme, mms_bump

    cases = {"heat": heat_readme(TFinal=0.02), "bump": mms_bump(TFinal=0.2), "diff_cn": diffusion_t_cn(),
             "patch_cn": fixed_patch_cn()}
    ref = {}
    for name, (pp, po) in cases.items():
        ref[name] = par.ebp_solve(pp, po)
        ref[name]["device"].close()
    reqs = [serving.ParabolicRequest(name, pp, po) for name, (pp, po) in cases.items()] + _requests(PICK[:4])
    srv = serving.Server(slots=16)
    done = {r.id: r for r in srv.run(reqs, realtime=False)}
    srv.close()
    for name in cases:
        r = done[name]
        assert r.error is None, r.error
        sol = ref[name]
        assert [s["vcycles"] for s in r.stages] == sol["cycles"]
        assert [s["resHist"] for s in r.stages] == [s["resHist"] for s in sol["stages"]]
        assert r.mass == sol["mass"], name
        assert r.nregrids == sol["nregrids"]
        H = sol["H"]
        for k in range(1, H.nlev + 1):
            for p in range(H.L(k).np):
                assert np.array_equal(r.Q[k - 1][p], H.L(k).Q[p]), (name, k, p)


def test_oracle_fixture_matches_oracle():
    """A sample of the committed oracle records, recomputed by the oracle:
    one request of every (BaseCells, MaxLevels, boundary mix) kind."""
    d = config5.fixture()["requests"]
    gold = config5.oracle_fixture()["requests"]
    assert len(gold) == len(d) == 512
    seen = set()
    for par in d:
        kind = (par["base"], par["levels"], tuple(par["bc"]).count("dirichlet") in (0, 4))
        if kind in seen or len(seen) >= 12:
            continue
        seen.add(kind)
        assert config5.oracle_record(par) == gold[par["id"]], par["id"]
    # the workload refines: some requests run 2 and 3 stages
    assert {len(g["vcycles"]) for g in gold} == {1, 2, 3}
    assert max(g["finalRel"][-1] for g in gold) <= 1e-8


def _served_record(r):
    return config5.record([[b for b in r.H.L(k).box] for k in range(1, r.H.nlev + 1)], r.stages, r.Q)


def _serve_all(precision, slots=128):
    reqs = _requests(range(512))
    srv = serving.Server(slots=slots, precision=precision)
    t0 = time.perf_counter()
    done = {r.id: r for r in srv.run(reqs, realtime=False)}
    wall = time.perf_counter() - t0
    print(f"{precision}: 512 requests in {wall:.2f} s, {srv.ticks} ticks, "
          f"mean batch {srv.tick_items / max(srv.ticks, 1):.1f}", flush=True)
    srv.close()
    assert len(done) == 512
    for r in done.values():
        assert r.error is None, (r.id, r.error)
    return done


@pytest.mark.gpu
def test_served_workload_fp64_equals_oracle():
    """All 512 requests served together in FP64 equal the oracle's ebSolve
    bit for bit (EllipticBSAM/src/ebSolve.m:22-77 per request)."""
    gold = config5.oracle_fixture()["requests"]
    done = _serve_all("fp64")
    bad = [i for i, r in done.items() if _served_record(r) != gold[i]]
    assert not bad, bad[:10]
    assert all(r.stages[-1]["finalRel"] <= 1e-8 for r in done.values())


@pytest.mark.gpu
def test_served_workload_mixed_converges_with_stated_error():
    gold = config5.oracle_fixture()["requests"]
    done64 = _serve_all("fp64")
    done = _serve_all("mixed")
    errs = []
    for i, r in done.items():
        assert r.stages[-1]["finalRel"] <= 1e-8, (i, r.stages[-1]["finalRel"])
        a = done64[i]
        if [list(map(list, a.H.L(k).box)) for k in range(1, a.H.nlev + 1)] != \
                [list(map(list, r.H.L(k).box)) for k in range(1, r.H.nlev + 1)]:
            continue  # tagging on the mixed stage solution chose another hierarchy
        num = den = 0.0
        for k in range(a.H.rootIdx, a.H.nlev + 1):
            L = a.H.L(k)
            for p in range(L.np):
                vis = ~L.covered[p]
                x = r.Q[k - 1][p][1:-1, 1:-1][vis]
                y = a.Q[k - 1][p][1:-1, 1:-1][vis]
                num += float(((x - y) ** 2).sum())
                den += float((y ** 2).sum())
        errs.append(np.sqrt(num / den))
    errs = np.array(errs)
    print(f"mixed vs fp64 (= oracle): {errs.size} requests on the same hierarchy, rel L2 max {errs.max():.2e}, "
          f"median {np.median(errs):.2e}", flush=True)
    assert errs.size >= 400 and errs.max() < 1e-6
    assert sum(_served_record(done64[i]) == gold[i] for i in done64) == 512


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_served_alone_equals_in_the_full_mix(precision):
    """A request of every stage count and size, served alone, gives exactly
    what it gives inside the full 512-request mix."""
    gold = config5.oracle_fixture()["requests"]
    pick = []
    for nst in (1, 2, 3):
        for base in (32, 64, 128):
            ids = [i for i, g in enumerate(gold) if len(g["vcycles"]) == nst and
                   config5.fixture()["requests"][i]["base"] == base]
            pick += ids[:2]
    # the workload's (stage count, BaseCells) kinds: 1 x 64, 1 x 128, 2 x 32, 2 x 64, 3 x 32
    assert len(pick) == 10
    mix = _serve_all(precision)
    for r in _requests(pick):
        srv = serving.Server(slots=4, precision=precision)
        alone = srv.run([r], realtime=False)[0]
        srv.close()
        print(f"request {r.id}: stages {[s['vcycles'] for s in alone.stages]}", flush=True)
        _same(alone, mix[r.id])
