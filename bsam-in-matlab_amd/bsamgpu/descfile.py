"""Descriptor files for the plain C driver (cdriver/bsam_run.c).

``write_desc`` stores a host hierarchy's bsam_hier_desc arrays and the
solver options in the little-endian layout documented in bsam_run.c;
``read_q`` reads the driver's solution file back into per-level, per-patch
padded arrays.
"""
from __future__ import annotations

import numpy as np

from .device import pack_hierarchy

MAGIC = b"BSAMHD01"


def write_desc(H, path, opts=None):
    opts = opts or H.opts
    pk = pack_hierarchy(H)
    with open(path, "wb") as f:
        f.write(MAGIC)
        np.array([H.nlev, H.rootIdx], dtype="<i4").tofile(f)
        np.array(H.prob.domain, dtype="<f8").tofile(f)
        for key in ("n", "np", "boxes", "bc"):
            pk[key].astype("<i4").tofile(f)
        np.array([pk["Q"].size, pk["FTrue"].size, pk["DeW"].size, pk["DnS"].size, pk["g"].size],
                 dtype="<i8").tofile(f)
        for key in ("Q", "FTrue", "DeW", "DnS", "CC", "ID", "g"):
            pk[key].astype("<f8").tofile(f)
        np.array([opts.NPre, opts.NPost, opts.MaxVCycles, 0 if opts.GhostInterp == "qif" else 1,
                  1 if opts.MassCorrection else 0], dtype="<i4").tofile(f)
        np.array([opts.Omega, opts.RelTol], dtype="<f8").tofile(f)


def read_q(H, path):
    buf = np.fromfile(path, dtype="<f8")
    out = []
    off = 0
    for k in range(1, H.nlev + 1):
        lev = []
        for p in range(H.L(k).np):
            mx, my = H.L(k).dims(p)
            n = (mx + 2) * (my + 2)
            lev.append(buf[off:off + n].reshape((mx + 2, my + 2), order="F"))
            off += n
        out.append(lev)
    if off != buf.size:
        raise ValueError(f"{path}: {buf.size} values, expected {off}")
    return out
