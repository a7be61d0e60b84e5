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
