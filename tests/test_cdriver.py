"""The plain C driver (cdriver/bsam_run.c) over the same C ABI.

CPU: the descriptor file written by the host layer parses and the library
derives its tables from it (host-only hierarchy, --device -1).
GPU: the driver's solve equals the Python-driven solve bit for bit.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from bsamgpu import _lib, descfile
from bsamgpu.device import GpuHierarchy
from cases import CASES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "bsam-in-matlab_amd", "cdriver", "bsam_run")


def _driver():
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (run __graft_entry__.build())")
    return DRIVER


@pytest.mark.parametrize("name", ["plain", "lshape", "neumann"])
def test_descriptor_round_trip_host_only(name, tmp_path):
    H = CASES[name](seed=None)
    path = tmp_path / f"{name}.desc"
    descfile.write_desc(H, path)
    r = subprocess.run([_driver(), str(path), "--device", "-1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["created"] and out["nlev"] == H.nlev
    assert out["patches"] == sum(H.L(k).np for k in range(1, H.nlev + 1))


def test_driver_reports_library_errors(tmp_path):
    H = CASES["plain"](seed=None)
    H.L(4).box[0] = [14, 53, 13, 52]  # misaligned (ebBuildOps.m:38-42)
    path = tmp_path / "bad.desc"
    descfile.write_desc(H, path)
    r = subprocess.run([_driver(), str(path), "--device", "-1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "eb_build_ops:align" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["lshape", "neumann", "mixed_bc"])
def test_driver_solve_matches_library(name, tmp_path):
    H = CASES[name](seed=None)
    desc = tmp_path / "h.desc"
    qout = tmp_path / "q.bin"
    descfile.write_desc(H, desc)
    print(f"driver case {name}", flush=True)
    r = subprocess.run([_driver(), str(desc), "--out", str(qout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rec = json.loads(r.stdout)
    G = GpuHierarchy(H)
    st = G.solve_cycles()
    assert rec["vcycles"] == st["vcycles"]
    assert rec["res_hist"] == st["resHist"]
    Qd = descfile.read_q(H, qout)
    for k in range(1, H.nlev + 1):
        for p, q in enumerate(G.get_field(k, _lib.BSAM_Q)):
            assert np.array_equal(q, Qd[k - 1][p]), (k, p)


SOLVER = os.path.join(ROOT, "bsam-in-matlab_amd", "cdriver", "bsam_solve")


def _libm_plain_problem():
    """The README elliptic problem with the C library's sin (as the C driver
    evaluates it), so both drivers sample the same coefficients."""
    import math

    from bsamgpu import Problem

    vsin = np.vectorize(math.sin, otypes=[float])
    c = 1.0 + 8.0 * math.pow(math.pi, 2.0)
    return Problem([0, 1, 0, 1], 1, 1, lambda x, y: c * (vsin((2.0 * math.pi) * x) * vsin((2.0 * math.pi) * y)),
                   ("dirichlet", 0))


@pytest.mark.gpu
def test_c_driver_runs_the_staged_readme_solve(tmp_path):
    """bsam_solve builds, tags, clusters, refines and solves the README
    elliptic example with no Python (native hierarchy construction): the
    published stage log (verification/readme-examples.txt:2-6) and boxes,
    and the Python-driven eb_solve of the same problem bit for bit."""
    from bsamgpu import configs, solve

    if not os.path.exists(SOLVER):
        pytest.fail(f"{SOLVER} is not built (run __graft_entry__.build())")
    qout = tmp_path / "q.bin"
    r = subprocess.run([SOLVER, "--out", str(qout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(x) for x in r.stdout.splitlines() if x.strip()]
    stages, final = lines[:-1], lines[-1]
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "published.json")))["readme_elliptic"]
    assert len(stages) == 3
    for st, want in zip(stages, gold["stages"]):
        print(f"stage {st['stage']}: {st['vcycles']} cycles, rel {st['final_rel']:.3e}", flush=True)
        assert st["amr_levels"] == want["amrLevels"] and st["vcycles"] == want["vcycles"]
        assert f"{st['final_rel']:.2e}" == want["finalRel"] and st["visible_cells"] == want["visibleCells"]
        assert f"{st['avg_factor']:.3f}" == want["avgFactor"]
    added = final["boxes"][final["root_idx"]:]
    for boxes, want in zip(added, gold["added_levels"]):
        assert len(boxes) == want["np"]
        assert sum((b[1] - b[0] + 1) * (b[3] - b[2] + 1) for b in boxes) == want["cells"]
    # the same staged solve driven from Python (host layer + device library)
    sol = solve.eb_solve(_libm_plain_problem(), configs.plain_options(RelTol=1e-9))
    H = sol["H"]
    sol["device"].close()
    assert [[list(map(int, b)) for b in H.L(k).box] for k in range(1, H.nlev + 1)] == final["boxes"]
    for st, sp in zip(stages, H.stats["stages"]):
        assert st["vcycles"] == sp["vcycles"] and st["res_hist"] == sp["resHist"]
    q = np.fromfile(qout, dtype=np.float64)
    off = 0
    for k in range(1, H.nlev + 1):
        for p in range(H.L(k).np):
            a = H.L(k).Q[p]
            b = q[off:off + a.size].reshape(a.shape, order="F")
            off += a.size
            assert np.array_equal(a, b), (k, p)
    assert off == q.size
