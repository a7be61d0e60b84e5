"""Per-hierarchy option search: the candidates it may pick, and (GPU) that the
chosen options give the same answer as the defaults.

Stated tolerance: the exact-smoother FP64 candidates are bit-identical to the
default solve (same residual history, same Q); any candidate it returns meets
RelTol.
"""
import numpy as np
import pytest

from bsamgpu import autotune, configs, hierarchy, _lib


def test_exact_candidates_are_fp64_exact_only():
    for c in autotune.EXACT_CANDIDATES:
        assert c["smoother"] in ("exact", "sweep") and c["precision"] == "fp64"
    keys = {tuple(sorted(c.items())) for c in autotune.ALL_CANDIDATES}
    assert len(keys) == len(autotune.ALL_CANDIDATES)
    assert all(c in autotune.ALL_CANDIDATES for c in autotune.EXACT_CANDIDATES)


@pytest.mark.gpu
def test_autotuned_options_reproduce_default_solve():
    from bsamgpu.device import GpuHierarchy

    H = hierarchy.setup(configs.plain_problem(), configs.plain_options(RelTol=1e-9))
    G = GpuHierarchy(H)
    G.reset()
    ref = G.solve_cycles()
    Qref = {k: G.get_field(k, _lib.BSAM_Q) for k in range(1, H.nlev + 1)}
    best, table = autotune.autotune(G, repeats=2)
    print(best, [(r["coarse_cells"], round(r["seconds"] * 1e3, 3)) for r in table], flush=True)
    assert len(table) == len(autotune.EXACT_CANDIDATES)
    for r in table:
        assert r["vcycles"] == ref["vcycles"] and r["finalRel"] == ref["finalRel"]
    G.reset()
    st = G.solve_cycles(**best)
    assert st["resHist"] == ref["resHist"]
    for k, qs in Qref.items():
        for a, b in zip(G.get_field(k, _lib.BSAM_Q), qs):
            assert np.array_equal(a, b), k
    best_all, table_all = autotune.autotune(G, bit_exact_only=False, repeats=1)
    assert any(r["finalRel"] <= G.opts.RelTol and r is not None for r in table_all)
    G.reset()
    st = G.solve_cycles(**best_all)
    assert st["finalRel"] <= G.opts.RelTol
    G.close()
