"""The C++ restatement of the solve stage (oracle/bsam_cpu.cpp) driven from an
``OracleHierarchy`` -- TEST AND BASELINE INFRASTRUCTURE ONLY.

Only ``tests/`` and the ``cpu_baseline`` leg of ``bench.py`` use it: the
tests check that it reproduces the NumPy oracle's residual history and
solution bit for bit, and the benchmark times it as the CPU baseline at one
thread and at every host core.  The library never loads it.

``CpuSolve(O)`` flattens the oracle's per-patch arrays and descriptors
(``build_ops`` / ``gather_parent_coeffs``: EllipticBSAM/src/ebBuildOps.m,
ebGatherParentCoeffs.m) into one double buffer and integer tables, and
``solve(max_cycles, rel_tol, threads)`` runs ebSolveCycles.m:8-79 on it.
``load_into(O)`` copies the state back into the oracle's arrays.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libbsamcpu.so")


class _Desc(ctypes.Structure):
    _fields_ = [("nlev", ctypes.c_int), ("root", ctypes.c_int), ("npre", ctypes.c_int), ("npost", ctypes.c_int),
                ("mass", ctypes.c_int), ("qif", ctypes.c_int), ("N", ctypes.c_int), ("M", ctypes.c_int),
                ("singular", ctypes.c_int), ("omega", ctypes.c_double),
                ("lev_n", ctypes.c_void_p), ("lev_h", ctypes.c_void_p), ("lev_np", ctypes.c_void_p),
                ("pint", ctypes.c_void_p), ("poff", ctypes.c_void_p), ("pcnt", ctypes.c_void_p),
                ("pdesc", ctypes.c_void_p), ("desc", ctypes.c_void_p), ("data", ctypes.c_void_p),
                ("cov", ctypes.c_void_p), ("Ainv", ctypes.c_void_p), ("bcRhs", ctypes.c_void_p)]


_lib = None


def build():
    """Compiles oracle/bsam_cpu.cpp with oracle/Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "_build/libbsamcpu.so"], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        lib = ctypes.CDLL(LIB_PATH)
        lib.bsam_cpu_solve.argtypes = [ctypes.POINTER(_Desc), ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_double)]
        lib.bsam_cpu_solve.restype = ctypes.c_int
        _lib = lib
    return _lib


class CpuSolve:
    """A flattened copy of an OracleHierarchy's state and tables."""

    def __init__(self, O):
        self.O = O
        data = []          # column-major arrays
        pos = [0]

        def put(a):
            a = np.asarray(a, dtype=float)
            off = pos[0]
            flat = a.ravel(order="F")
            data.append(flat)
            pos[0] += flat.size
            return off

        cov = []
        cpos = [0]

        def put_cov(c):
            off = cpos[0]
            flat = np.asarray(c, dtype=np.uint8).ravel(order="F")
            cov.append(flat)
            cpos[0] += flat.size
            return off

        zero = put(np.zeros(8))
        pint, poff, pcnt, pdesc, desc = [], [], [], [], []
        self._q_off = []
        self._qc_off = []
        for k in range(1, O.nlev + 1):
            L = O.L(k)
            for p in range(L.np):
                Q = L.Q[p]
                mx, my = Q.shape[0] - 2, Q.shape[1] - 2
                ccl = L.cornerClass[p]
                pint += [mx, my, int(L.par0[p])] + [int(c) for c in L.bc_code[p]] + \
                    [1 if f else 0 for f in L.faceInternal[p]] + [int(c) for c in ccl]
                offs = [put(Q), put(L.F[p]), put(L.FTrue[p]), put(L.DeW[p]), put(L.DnS[p]), put(L.CC[p]),
                        put(L.InvDiagS[p])]
                self._q_off.append((k, p, offs[0], offs[1]))
                if k > 1:
                    qc = put(L.QC[p])
                    self._qc_off.append((k, p, qc))
                    offs.append(qc)
                else:
                    offs.append(zero)
                for s in range(4):
                    g = np.asarray(L.bc_g[p][s], dtype=float).ravel()
                    offs.append(put(g) if g.size else zero)
                if k > 1:
                    for v in L.ops[p]["cfD"]:
                        offs.append(put(v))
                else:
                    offs += [zero] * 4
                poff += offs + [put_cov(L.covered[p])]
                pdesc.append(len(desc))
                cnt = []
                same = [(sp, int(a), int(b), int(c), int(d)) for sp, si, sj, di, dj in L.sameOps[p]
                        for a, b, c, d in zip(*np.broadcast_arrays(si, sj, di, dj))]
                cnt.append(len(same))
                for e in same:
                    desc += list(e)
                if k > 1:
                    ops = L.ops[p]
                    ring = [(q, int(a), int(b), int(c), int(d)) for q, si, sj, di, dj in ops["ring"]
                            for a, b, c, d in zip(si, sj, di, dj)]
                    cnt.append(len(ring))
                    for e in ring:
                        desc += list(e)
                    # rectangles (q, source rect, destination rect): winrect parent
                    # padded -> window; usc R -> parent padded; fsc loadC -> parent F;
                    # pcc/pdew/pdns parent -> footprint
                    for key in ("winrect", "usc", "fsc", "pcc", "pdew", "pdns"):
                        cnt.append(len(ops[key]))
                        for q, s, d in ops[key]:
                            desc += [q] + [int(v) for v in s] + [int(v) for v in d]
                    for f in range(4):
                        ent = [(q, int(a), int(b), int(c)) for q, src, di, dj in ops["corr"][f]
                               for a, b, c in zip(src, di, dj)]
                        cnt.append(len(ent))
                        for e in ent:
                            desc += list(e)
                else:
                    cnt += [0] * 11
                pcnt += cnt
        self.data = np.concatenate(data)
        self._data0 = self.data.copy()
        self.cov = np.concatenate(cov) if cov else np.zeros(1, np.uint8)
        self.pint = np.asarray(pint, np.int32)
        self.poff = np.asarray(poff, np.int64)
        self.pcnt = np.asarray(pcnt, np.int32)
        self.pdesc = np.asarray(pdesc, np.int64)
        self.desc = np.asarray(desc if desc else [0], np.int32)
        self.lev_n = np.asarray([v for k in range(1, O.nlev + 1) for v in O.L(k).n], np.int32)
        self.lev_h = np.asarray([v for k in range(1, O.nlev + 1) for v in O.L(k).h], float)
        self.lev_np = np.asarray([O.L(k).np for k in range(1, O.nlev + 1)], np.int32)
        c = O.coarsest
        self.Ainv = np.asfortranarray(c["Ainv"], dtype=float)
        self.bcRhs = np.ascontiguousarray(c["bcRhs"], dtype=float)
        n1, n2 = O.L(1).n
        d = _Desc()
        d.nlev, d.root = O.nlev, O.rootIdx
        d.npre, d.npost = int(O.opts.NPre), int(O.opts.NPost)
        d.mass, d.qif = int(bool(O.opts.MassCorrection)), int(O.opts.GhostInterp == "qif")
        d.N, d.M, d.singular = n1 * n2, self.Ainv.shape[0], int(bool(c["singular"]))
        d.omega = float(O.opts.Omega)
        for name in ("lev_n", "lev_h", "lev_np", "pint", "poff", "pcnt", "pdesc", "desc", "data", "cov", "Ainv",
                     "bcRhs"):
            setattr(d, name, getattr(self, name).ctypes.data)
        self.d = d

    def reset(self):
        """Restores the state the flattened copy was made from."""
        self.data[:] = self._data0

    def solve(self, max_cycles=None, rel_tol=None, threads=0):
        """ebSolveCycles on the flattened state; returns (residual history,
        cycles, final relative residual)."""
        lib = load()
        maxv = int(self.O.opts.MaxVCycles if max_cycles is None else max_cycles)
        tol = float(self.O.opts.RelTol if rel_tol is None else rel_tol)
        hist = np.zeros(maxv + 1)
        it = ctypes.c_int()
        rel = ctypes.c_double()
        rc = lib.bsam_cpu_solve(ctypes.byref(self.d), maxv, tol, int(threads), hist.ctypes.data, hist.size,
                                ctypes.byref(it), ctypes.byref(rel))
        if rc:
            raise RuntimeError(f"bsam_cpu_solve returned {rc}")
        return hist[:it.value + 1].copy(), it.value, rel.value

    def q(self, k, p):
        """Q of patch p on level k from the flattened state."""
        for kk, pp, off, _ in self._q_off:
            if kk == k and pp == p:
                shape = self.O.L(k).Q[p].shape
                return self.data[off:off + shape[0] * shape[1]].reshape(shape, order="F").copy()
        raise KeyError((k, p))

    def load_into(self, O=None):
        """Copies Q, F and QC back into the oracle's arrays."""
        O = self.O if O is None else O
        for k, p, qo, fo in self._q_off:
            L = O.L(k)
            sq, sf = L.Q[p].shape, L.F[p].shape
            L.Q[p][...] = self.data[qo:qo + sq[0] * sq[1]].reshape(sq, order="F")
            L.F[p] = self.data[fo:fo + sf[0] * sf[1]].reshape(sf, order="F").copy()
        for k, p, off in self._qc_off:
            s = O.L(k).QC[p].shape
            O.L(k).QC[p][...] = self.data[off:off + s[0] * s[1]].reshape(s, order="F")
