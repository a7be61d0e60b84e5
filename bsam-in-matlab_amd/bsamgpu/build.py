"""The library's native hierarchy construction (bsam_build_*) from Python.

The host layer (``hierarchy.setup`` / ``hierarchy.add_level``) builds the
hierarchy value in NumPy; the library builds the same value in C++ from
coefficient callbacks, which is what the plain C driver and callers without
Python use (EllipticBSAM/src/ebSetup.m:10-68, ebMakeLevel.m:1-105,
ebCoeffFilldown.m:1-62, ebAddLevel.m:10-24).  ``NativeBuild`` drives it with
a ``Problem``'s callables, so both constructions can be compared on the same
coefficient values.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check
from .problem import Options, Problem


def _field(fn):
    """A bsam_field_fn calling fn(x, y) on the points it is given."""
    if fn is None:
        return _lib.FIELD_FN()

    def cb(user, x, y, n, out):
        X = np.ctypeslib.as_array(x, (n,))
        Y = np.ctypeslib.as_array(y, (n,))
        v = np.asarray(fn(X, Y), dtype=float)
        o = np.ctypeslib.as_array(out, (n,))
        o[:] = np.broadcast_to(v.reshape(-1) if v.size == n else v, (n,))

    return _lib.FIELD_FN(cb)


class NativeBuild:
    """A ``bsam_build`` handle: ladder and root at creation, refined levels
    added with ``add_level`` (parent-space boxes)."""

    def __init__(self, prob: Problem, opts: Options):
        self.lib = _lib.load()
        p = _lib.Problem()
        for i in range(4):
            p.domain[i] = float(prob.domain[i])
        self._keep = [_field(prob.D), _field(prob.C), _field(prob.f)]
        p.D, p.C, p.f = self._keep
        for s in range(4):
            p.bc_type[s] = 1 if prob.bc[s].type == "dirichlet" else 2
            cb = _field(prob.bc[s].g)
            self._keep.append(cb)
            p.g[s] = cb
        cb = _field(opts.InitialGuess)
        self._keep.append(cb)
        p.u0 = cb
        o = _lib.BuildOpts()
        o.base_cells[0], o.base_cells[1] = int(opts.BaseCells[0]), int(opts.BaseCells[1])
        o.coarsest_cells = int(opts.CoarsestCells)
        o.source_quad = int(opts.SourceQuad)
        self._p, self._o = p, o
        self.b = ctypes.c_void_p()
        check(self.lib.bsam_build_create(ctypes.byref(p), ctypes.byref(o), ctypes.byref(self.b)), "bsam_build_create")

    def add_level(self, boxes):
        b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 4))
        check(self.lib.bsam_build_add_level(self.b, b.ctypes.data_as(_lib.P_int), b.shape[0]), "bsam_build_add_level")

    def info(self):
        n, r, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(self.lib.bsam_build_info(self.b, ctypes.byref(n), ctypes.byref(r), ctypes.byref(s)), "bsam_build_info")
        return n.value, r.value, bool(s.value)

    def desc(self):
        """The bsam_hier_desc and its arrays as NumPy copies (the layout of
        device.pack_hierarchy)."""
        d = _lib.HierDesc()
        check(self.lib.bsam_build_desc(self.b, ctypes.byref(d)), "bsam_build_desc")
        nlev = d.nlev
        n = np.ctypeslib.as_array(d.n, (2 * nlev,)).copy()
        np_ = np.ctypeslib.as_array(d.np, (nlev,)).copy()
        P = int(np_.sum())
        boxes = np.ctypeslib.as_array(d.boxes, (4 * P,)).copy()
        bc = np.ctypeslib.as_array(d.bc_code, (4 * P,)).copy()
        nq = ncell = nx = ny = ng = 0
        for b in boxes.reshape(-1, 4):
            mx, my = int(b[1] - b[0] + 1), int(b[3] - b[2] + 1)
            nq += (mx + 2) * (my + 2)
            ncell += mx * my
            nx += (mx + 1) * my
            ny += mx * (my + 1)
            ng += 2 * (mx + my)
        out = {"nlev": nlev, "root_idx": d.root_idx, "n": n, "np": np_, "boxes": boxes, "bc": bc}
        for key, field, size in (("Q", "Q", nq), ("FTrue", "FTrue", ncell), ("DeW", "DeW", nx), ("DnS", "DnS", ny),
                                 ("CC", "CC", ncell), ("ID", "InvDiagS", ncell), ("g", "g", ng)):
            out[key] = np.ctypeslib.as_array(getattr(d, field), (size,)).copy() if size else np.zeros(0)
        return out

    def allowed(self):
        d = self.desc()
        n1, n2 = int(d["n"][-2]), int(d["n"][-1])
        a = np.zeros(n1 * n2, dtype=np.uint8)
        check(self.lib.bsam_build_allowed(self.b, a.ctypes.data), "bsam_build_allowed")
        return a.reshape((n1, n2), order="F").astype(bool)

    def close(self):
        if getattr(self, "b", None):
            self.lib.bsam_build_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
