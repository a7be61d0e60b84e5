"""The C++ restatement of the solve stage (oracle/bsam_cpu.cpp, the CPU
baseline the benchmark times) against the NumPy oracle.

Stated tolerance: bit-identical residual histories, V-cycle counts and final
Q, F and QC of every patch, on every parity case (README layout, L-shaped
union with QIF and LIF ghosts, singular Neumann, mixed sides), at one thread
and at four (the OpenMP split over patches changes no sum order).
"""
import numpy as np
import pytest

from cases import CASES
from oracle import bsam_cpu
from oracle import bsam_oracle as O


@pytest.fixture(scope="module", autouse=True)
def _built():
    bsam_cpu.build()


def _same_state(A, B):
    for k in range(1, A.nlev + 1):
        for p in range(A.L(k).np):
            assert np.array_equal(A.L(k).Q[p], B.L(k).Q[p]), ("Q", k, p)
            assert np.array_equal(A.L(k).F[p], B.L(k).F[p]), ("F", k, p)
            if k > 1:
                assert np.array_equal(A.L(k).QC[p], B.L(k).QC[p]), ("QC", k, p)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("threads", [1, 4])
def test_cycles_bit_identical(name, threads):
    H = CASES[name](seed=5)
    A = O.OracleHierarchy(H)
    B = O.OracleHierarchy(H)
    st = A.solve_cycles(max_cycles=5)
    c = bsam_cpu.CpuSolve(B)
    hist, it, rel = c.solve(max_cycles=5, threads=threads)
    assert it == st["vcycles"]
    assert np.array_equal(np.asarray(st["resHist"]), hist)
    assert rel == st["finalRel"]
    c.load_into(B)
    _same_state(A, B)


def test_solve_to_tolerance_and_reset():
    H = CASES["plain"]()
    A = O.OracleHierarchy(H)
    st = A.solve_cycles()
    c = bsam_cpu.CpuSolve(O.OracleHierarchy(H))
    for _ in range(2):  # reset restores the uploaded state
        c.reset()
        hist, it, rel = c.solve(threads=2)
        assert it == st["vcycles"] and np.array_equal(np.asarray(st["resHist"]), hist)
    assert rel <= A.opts.RelTol


def test_port_catches_a_planted_change():
    """A one-ulp change in one coefficient of one patch must show in the
    history: the comparison above is not vacuous."""
    H = CASES["lshape"](seed=3)
    A = O.OracleHierarchy(H)
    st = A.solve_cycles(max_cycles=3)
    B = O.OracleHierarchy(H)
    L = B.L(B.nlev)
    L.DeW[0][3, 2] = np.nextafter(L.DeW[0][3, 2], np.inf)
    hist, _, _ = bsam_cpu.CpuSolve(B).solve(max_cycles=3, threads=1)
    assert not np.array_equal(np.asarray(st["resHist"]), hist)
