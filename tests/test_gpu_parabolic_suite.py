"""The reference's parabolic verification suite through the device library:
every time step's solve, the BDF2 / BE / CN right sides and the conservative
regrids run on the GPU (bsam_parabolic_step, bsam_regrid_level) under the
host layer's ebp_solve, against the numbers published in
verification/suite-output.txt:184-213 (checks in parabolic_suite.py; the
oracle runs the same checks in test_parabolic_suite.py)."""
import pytest

import parabolic_suite as PS
from bsamgpu import parabolic as par

pytestmark = pytest.mark.gpu


def device_solve(pp, opts):
    cb = opts.Callback
    if cb is not None:
        opts.Callback = lambda S, t1, n: cb(S.H, t1, n)
    sol = par.ebp_solve(pp, opts)
    sol["device"].close()
    return sol


def test_p1a_temporal_cauchy_differences():
    PS.p1a(device_solve)


@pytest.mark.parametrize("j", [0, 1, 2])
def test_p1b_joint_refinement_with_regrids(j):
    PS.p1b(device_solve, j)


def test_p2_heat_mass_across_regrids_and_p4():
    PS.p2_and_p4_mean(device_solve)


def test_p3_rotating_source_tracking():
    PS.p3(device_solve)


def test_p4_large_dt_backward_euler():
    PS.p4_large_dt(device_solve)
