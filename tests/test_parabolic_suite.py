"""The reference's parabolic verification suite on the CPU oracle
(ParabolicOracle: ebpInit / ebpSetTime / ebpRhs / ebpRegrid / ebpSolve with
every solve by the oracle), against the numbers published in
verification/suite-output.txt:184-213; the checks live in parabolic_suite.py
and the device path runs them in test_gpu_parabolic_suite.py."""
import pytest

import parabolic_suite as PS
from bsamgpu import diagnostics
from bsamgpu import parabolic as par
from oracle import bsam_oracle as O
from parabolic_cases import Host


def oracle_solve(pp, opts):
    """ebpSolve.m:17-98 on the oracle, returned with the fields the suite
    prints (the regrid count follows ebpSolve.m:73-77)."""
    po = O.ParabolicOracle(pp, opts, Host, par)
    cb = opts.Callback
    R = po.solve(callback=None if cb is None else (lambda s, n: cb(s.H, n * s_dt(s), n)))
    po._sync(po.H, po.O)
    H = po.H
    n = R["nsteps"]
    nregrids = sum(1 for i in range(1, n + 1)
                   if opts.RegridEvery > 0 and i % opts.RegridEvery == 0 and i < n and opts.eb.MaxLevels > 1)
    T = opts.TFinal
    err = diagnostics.error_norms(H, lambda x, y: pp.exact(x, y, T)) if pp.exact is not None else None
    return {"cycles": [s["vcycles"] for s in R["stages"]], "relres": [s["finalRel"] for s in R["stages"]],
            "mass": R["mass"], "nregrids": nregrids, "H": H, "err": err,
            "levels": [{"cells": sum(int(H.L(k).dims(p)[0] * H.L(k).dims(p)[1]) for p in range(H.L(k).np))}
                       for k in range(H.rootIdx, H.nlev + 1)]}


def s_dt(s):
    return s.opts.TFinal / max(1, int(s.opts.TFinal / s.opts.Dt + 0.5))


def test_p1a_temporal_cauchy_differences():
    PS.p1a(oracle_solve)


@pytest.mark.parametrize("j", [0, 1, 2])
def test_p1b_joint_refinement_with_regrids(j):
    PS.p1b(oracle_solve, j)


def test_p2_heat_mass_across_regrids_and_p4():
    PS.p2_and_p4_mean(oracle_solve)


def test_p3_rotating_source_tracking():
    PS.p3(oracle_solve)


def test_p4_large_dt_backward_euler():
    PS.p4_large_dt(oracle_solve)
