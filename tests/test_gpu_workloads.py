"""Device path on the benchmark workloads at their real sizes.

  README elliptic (base 32, 3 levels) through eb_solve: the published stage log.
  Variable-D 256^2: bit-exact against the oracle, and the finding
  that ulap at 1e-3 tags nothing there (one stage, one AMR level).
  Starfish 64 -> 1024 (544,664 cells): the whole FP64 solve equals
  the oracle's bit for bit (residual history, every Q), and the mixed form
  reaches the same solution within 1e-10 relative L2.
  Variable-D refined to 2048^2 spacing (the 256^2 problem with the ulap threshold at
  1e-6, 1.75M cells): the staged solve (tagging after each stage) builds the
  same hierarchy as the oracle-driven one and ends on the same Q bit for bit;
  the mixed form of the final stage is within 1e-10 of the FP64 one.
"""
import json
import os

import numpy as np
import pytest

from bsamgpu import _lib, configs, solve
from bsamgpu.device import GpuHierarchy
from oracle import bsam_oracle as O

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "published.json")))


def test_readme_elliptic_published_log_on_device():
    sol = solve.eb_solve(configs.plain_problem(), configs.plain_options(RelTol=1e-9))
    ref = GOLD["readme_elliptic"]
    stages = sol["stats"]["stages"]
    assert len(stages) == 3
    for st, want in zip(stages, ref["stages"]):
        assert st["amrLevels"] == want["amrLevels"]
        assert st["vcycles"] == want["vcycles"]
        assert f"{st['finalRel']:.2e}" == want["finalRel"]
        assert f"{st['avgFactor']:.3f}" == want["avgFactor"]
        assert st["visibleCells"] == want["visibleCells"]
    sol["device"].close()


def test_varcoef_256_bit_exact_and_untagged():
    with open(os.path.join(os.path.dirname(__file__), "golden", "config2_params.json")) as f:
        par = json.load(f)
    prob = configs.varcoef_problem(par)
    opts = configs.varcoef_options()
    sol = solve.eb_solve(prob, opts)
    H = sol["H"]
    assert len(sol["stats"]["stages"]) == 1 and H.nlev == H.rootIdx  # nothing tagged at 1e-3
    sol["device"].close()
    from bsamgpu import hierarchy

    H0 = hierarchy.setup(prob, opts)
    Orc = O.OracleHierarchy(H0)
    so = Orc.solve_cycles()
    G = GpuHierarchy(H0)
    sg = G.solve_cycles()
    assert sg["resHist"] == so["resHist"] and sg["vcycles"] == so["vcycles"]
    for k in range(1, H0.nlev + 1):
        for p, q in enumerate(G.get_field(k, _lib.BSAM_Q)):
            assert np.array_equal(q, Orc.L(k).Q[p]), (k, p)
    G.close()


@pytest.fixture(scope="module")
def starfish():
    prob, aux, g = configs.starfish_problem()
    return solve.build_hierarchy(prob, configs.starfish_options(g))


def test_starfish_full_size_bit_exact(starfish):
    H = starfish
    print("starfish: oracle solve (CPU)", flush=True)
    Orc = O.OracleHierarchy(H)
    so = Orc.solve_cycles()
    print(f"oracle: {so['vcycles']} cycles, rel {so['finalRel']:.3e}; device solve", flush=True)
    G = GpuHierarchy(H)
    sg = G.solve_cycles()
    assert sg["vcycles"] == so["vcycles"] == GOLD["starfish_full"]["vcycles"]
    assert sg["resHist"] == so["resHist"]
    assert sg["finalRel"] == so["finalRel"]
    for k in range(1, H.nlev + 1):
        got = G.get_field(k, _lib.BSAM_Q)
        for p in range(H.L(k).np):
            assert np.array_equal(got[p], Orc.L(k).Q[p]), (k, p)
    # fused red+black sweep launches: the same sequence bit for bit
    G.reset()
    ss = G.solve_cycles(smoother="sweep")
    print(f"sweep smoother: {ss['vcycles']} cycles", flush=True)
    assert ss["resHist"] == so["resHist"]
    for k in range(1, H.nlev + 1):
        got = G.get_field(k, _lib.BSAM_Q)
        for p in range(H.L(k).np):
            assert np.array_equal(got[p], Orc.L(k).Q[p]), ("sweep", k, p)
    # FP32-inner form: same fixed point (1e-10 rel L2 on visible cells, the FP64 bound)
    G.reset()
    sm = G.solve_cycles(precision="mixed")
    num = den = 0.0
    for k in range(H.rootIdx, H.nlev + 1):
        got = G.get_field(k, _lib.BSAM_Q)
        for p in range(H.L(k).np):
            vis = ~H.L(k).covered[p]
            a = got[p][1:-1, 1:-1][vis]
            b = Orc.L(k).Q[p][1:-1, 1:-1][vis]
            num += float(((a - b) ** 2).sum())
            den += float((b ** 2).sum())
    err = (num / den) ** 0.5
    print(f"mixed: {sm['vcycles']} cycles, rel {sm['finalRel']:.3e}, rel L2 vs FP64 oracle {err:.2e}", flush=True)
    assert sm["finalRel"] <= 1e-10 and err < 1e-10
    # the cycles stop their FP64 update at the root; after the last one the
    # sub-root levels hold the restriction of the root, as after FP64 cycles
    assert H.rootIdx > 1
    for k in range(1, H.rootIdx):
        assert H.L(k).np == 1 and H.L(k + 1).np == 1
        child = G.get_field(k + 1, _lib.BSAM_Q)[0][1:-1, 1:-1]
        assert np.array_equal(G.get_field(k, _lib.BSAM_Q)[0][1:-1, 1:-1], O.restrict2(child)), k
    G.close()


class _Host:
    @staticmethod
    def setup(prob, opts):
        from bsamgpu import hierarchy

        return hierarchy.setup(prob, opts)

    @staticmethod
    def tag_cluster(H):
        from bsamgpu import tagcluster

        return tagcluster.tag_cluster(H)

    @staticmethod
    def add_level(H, boxes):
        from bsamgpu import hierarchy

        return hierarchy.add_level(H, boxes)


def test_varcoef_refined_to_2048_bit_exact_and_mixed():
    with open(os.path.join(os.path.dirname(__file__), "golden", "config2_params.json")) as f:
        par = json.load(f)
    prob = configs.varcoef_problem(par)
    opts = configs.varcoef_refined_options()
    print("varcoef refined: device stages", flush=True)
    sol = solve.eb_solve(prob, opts)
    Hg = sol["H"]
    sol["device"].close()
    print(f"device: {Hg.nlev} levels, patches {[Hg.L(k).np for k in range(1, Hg.nlev + 1)]}; oracle stages",
          flush=True)
    Ho, Orc = O.eb_solve(prob, opts, _Host)
    assert Hg.nlev == Ho.nlev == 8
    assert [Hg.L(k).np for k in range(1, 9)] == [1, 1, 1, 1, 1, 1, 1, 20]
    for k in range(1, Hg.nlev + 1):
        assert [list(b) for b in Hg.L(k).box] == [list(b) for b in Ho.L(k).box], k
        for p in range(Hg.L(k).np):
            assert np.array_equal(Hg.L(k).Q[p], Ho.L(k).Q[p]), (k, p)
    # the final stage from zero: FP64 against the FP32-inner form
    for k in range(1, Hg.nlev + 1):
        for p in range(Hg.L(k).np):
            Hg.L(k).Q[p][:] = 0.0
    G = GpuHierarchy(Hg)
    s64 = G.solve_cycles()
    Q64 = {k: G.get_field(k, _lib.BSAM_Q) for k in range(Hg.rootIdx, Hg.nlev + 1)}
    G.reset()
    ss = G.solve_cycles(smoother="sweep")
    assert ss["resHist"] == s64["resHist"]
    for k in range(Hg.rootIdx, Hg.nlev + 1):
        for a, b in zip(G.get_field(k, _lib.BSAM_Q), Q64[k]):
            assert np.array_equal(a, b), ("sweep", k)
    G.reset()
    smx = G.solve_cycles(precision="mixed")
    num = den = 0.0
    for k in range(Hg.rootIdx, Hg.nlev + 1):
        got = G.get_field(k, _lib.BSAM_Q)
        for p in range(Hg.L(k).np):
            vis = ~Hg.L(k).covered[p]
            a = got[p][1:-1, 1:-1][vis]
            b = Q64[k][p][1:-1, 1:-1][vis]
            num += float(((a - b) ** 2).sum())
            den += float((b ** 2).sum())
    err = (num / den) ** 0.5
    print(f"final stage: fp64 {s64['vcycles']} cycles rel {s64['finalRel']:.2e}; mixed {smx['vcycles']} cycles "
          f"rel {smx['finalRel']:.2e}, rel L2 vs fp64 {err:.2e}", flush=True)
    assert s64["finalRel"] <= 1e-10 and smx["finalRel"] <= 1e-10 and err < 1e-10
    G.close()
