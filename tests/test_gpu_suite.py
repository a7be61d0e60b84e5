"""The reference's elliptic verification suite through the device library:
every solve that EllipticBSAM/tests/runAllTests.m prints (T1-T5, T8; 49
solves, uniform 32^2 to 256^2 roots, box-in-box QIF and LIF hierarchies,
L-shaped and scattered patches, ulap-adaptive staging) run by the host
layer's eb_solve with each stage on the GPU, against the numbers published
in verification/suite-output.txt:3-79, and bit for bit against the same run
through the oracle (the CPU twin is test_suite_pins.py).
"""
import json
import os
import warnings

import numpy as np
import pytest

from bsamgpu import solve
import suite_cases as S
from test_suite_pins import oracle_solve

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "published.json")))["suite"]
RUNS = S.suite_runs(GOLD)


@pytest.mark.parametrize("run", RUNS, ids=[r[0] for r in RUNS])
def test_suite_solve_on_device(run):
    label, prob, opts, row = run
    print(f"suite case {label}", flush=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sol = solve.eb_solve(prob, opts)
    sol["device"].close()
    H = sol["H"]
    st = sol["stats"]["stages"][-1]
    S.check_published(label, H, st, sol["err"], sol["mass"], row)
    Ho, stages, _, _ = oracle_solve(prob, opts)
    assert [s["resHist"] for s in sol["stats"]["stages"]] == [s["resHist"] for s in stages], label
    assert Ho.nlev == H.nlev
    for k in range(1, H.nlev + 1):
        assert [b.tolist() for b in Ho.L(k).box] == [b.tolist() for b in H.L(k).box], (label, k)
        for p in range(H.L(k).np):
            assert np.array_equal(H.L(k).Q[p], Ho.L(k).Q[p]), (label, k, p)
