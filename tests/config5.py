"""The served workload's requests (seeded fixture) and the digest of a solved
hierarchy used to compare served results with the oracle's.

Digest of a hierarchy: per level, SHA-256 of the level's patches' padded Q
arrays (FP64, column-major, patch after patch) -- equal digests mean equal Q
bit for bit.  Stage records keep V-cycle counts and the residual histories as
exact doubles.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture():
    with open(os.path.join(GOLD, "config5_requests.json")) as f:
        return json.load(f)


def oracle_fixture():
    with open(os.path.join(GOLD, "config5_oracle.json")) as f:
        return json.load(f)


def q_digests(Q):
    """Q: per level (1..nlev) a list of padded patch arrays."""
    out = []
    for qs in Q:
        h = hashlib.sha256()
        for q in qs:
            h.update(np.asarray(q, dtype=np.float64).tobytes(order="F"))
        out.append(h.hexdigest())
    return out


def record(boxes, stages, Q):
    """The comparable record of one solved request."""
    return {"boxes": [[list(map(int, b)) for b in lv] for lv in boxes],
            "vcycles": [int(s["vcycles"]) for s in stages],
            "resHist": [[float(x) for x in s["resHist"]] for s in stages],
            "finalRel": [float(s["finalRel"]) for s in stages],
            "q": q_digests(Q)}


def oracle_record(par):
    """ebSolve of one request driven by the CPU oracle."""
    from bsamgpu import configs, hierarchy, tagcluster
    from oracle import bsam_oracle as O

    class _Host:
        setup = staticmethod(hierarchy.setup)
        add_level = staticmethod(hierarchy.add_level)
        tag_cluster = staticmethod(tagcluster.tag_cluster)

    H, Orc = O.eb_solve(configs.served_problem(par), configs.served_options(par), _Host)
    boxes = [[b for b in H.L(k).box] for k in range(1, H.nlev + 1)]
    Q = [[Orc.L(k).Q[p] for p in range(H.L(k).np)] for k in range(1, H.nlev + 1)]
    return record(boxes, Orc.stats["stages"], Q)
