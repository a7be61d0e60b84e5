"""Device-resident hierarchy handle: packs a host ``Hierarchy`` into the C ABI
descriptor, owns the ``bsam_hier`` handle and exposes the solve-path calls
under the reference's names (ebSolveCycles, ebVcycle, ebFillGhosts, ...)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr

_CTX = {}


def context(device=0):
    """One library context per HIP device (created on first use)."""
    if device not in _CTX:
        lib = _lib.load()
        h = ctypes.c_void_p()
        check(lib.bsam_ctx_create(device, ctypes.byref(h)), "bsam_ctx_create")
        _CTX[device] = h
    return _CTX[device]


def solve_opts(opts, precision=None, use_graph=True, inner_cycles=1, max_vcycles=None, coarse_cells=0,
               smoother=None, stream_rows=0):
    o = _lib.default_opts()
    o.npre = int(opts.NPre)
    o.npost = int(opts.NPost)
    o.omega = float(opts.Omega)
    o.rel_tol = float(opts.RelTol)
    o.max_vcycles = int(opts.MaxVCycles if max_vcycles is None else max_vcycles)
    o.ghost_interp = 0 if opts.GhostInterp == "qif" else 1
    o.mass_correction = 1 if opts.MassCorrection else 0
    prec = precision or getattr(opts, "Precision", "fp64")
    o.precision = _lib.BSAM_MIXED if prec == "mixed" else _lib.BSAM_FP64
    o.inner_cycles = int(inner_cycles)
    o.use_graph = 1 if use_graph else 0
    o.coarse_cells = int(coarse_cells)
    sm = smoother if smoother is not None else getattr(opts, "Smoother", "exact")
    o.smoother = {"exact": 0, "tiled": 1, "sweep": 2, "passes": 3, "classic": 4}[sm] if isinstance(sm, str) else int(sm)
    o.stream_rows = int(stream_rows)
    return o


def _packF(arrs):
    if not arrs:
        return np.zeros(0)
    return np.concatenate([np.asarray(a, float).reshape(-1, order="F") for a in arrs])


def pack_hierarchy(H):
    """Descriptor arrays in the C ABI layout (levels 1..nlev, patches in order,
    column-major per patch)."""
    n, np_, boxes, codes = [], [], [], []
    Q, FT, DeW, DnS, CC, ID, G = [], [], [], [], [], [], []
    for k in range(1, H.nlev + 1):
        L = H.L(k)
        n += [L.n[0], L.n[1]]
        np_.append(L.np)
        for p in range(L.np):
            boxes += [int(v) for v in L.box[p]]
            codes += [int(c) for c in L.bc_code[p]]
            Q.append(L.Q[p])
            FT.append(L.FTrue[p])
            DeW.append(L.DeW[p])
            DnS.append(L.DnS[p])
            CC.append(L.CC[p])
            ID.append(L.InvDiagS[p])
            G += list(L.bc_g[p])
    return {
        "n": np.ascontiguousarray(n, dtype=np.int32), "np": np.ascontiguousarray(np_, dtype=np.int32),
        "boxes": np.ascontiguousarray(boxes, dtype=np.int32), "bc": np.ascontiguousarray(codes, dtype=np.int32),
        "Q": _packF(Q), "FTrue": _packF(FT), "DeW": _packF(DeW), "DnS": _packF(DnS), "CC": _packF(CC),
        "ID": _packF(ID), "g": np.concatenate([np.asarray(g, float).ravel() for g in G]),
    }


class GpuHierarchy:
    """A ``bsam_hier`` handle built from a host hierarchy."""

    def __init__(self, H, opts=None, device=0):
        self.lib = _lib.load()
        self.H = H
        self.opts = opts or H.opts
        self.nlev = H.nlev
        self.rootIdx = H.rootIdx
        ctx = context(device)
        d, self._pk = self._desc(H)
        so = solve_opts(self.opts)
        self.h = ctypes.c_void_p()
        check(self.lib.bsam_hier_create(ctx, ctypes.byref(d), ctypes.byref(so), ctypes.byref(self.h)),
              "bsam_hier_create")

    @staticmethod
    def _desc(H):
        pk = pack_hierarchy(H)
        d = _lib.HierDesc()
        d.nlev = H.nlev
        d.root_idx = H.rootIdx
        for i in range(4):
            d.domain[i] = float(H.prob.domain[i])
        d.n, d.np, d.boxes, d.bc_code = iptr(pk["n"]), iptr(pk["np"]), iptr(pk["boxes"]), iptr(pk["bc"])
        d.Q, d.FTrue, d.DeW, d.DnS = dptr(pk["Q"]), dptr(pk["FTrue"]), dptr(pk["DeW"]), dptr(pk["DnS"])
        d.CC, d.InvDiagS, d.g = dptr(pk["CC"]), dptr(pk["ID"]), dptr(pk["g"])
        return d, pk  # pk keeps the arrays behind the pointers alive

    def close(self):
        if getattr(self, "h", None):
            self.lib.bsam_hier_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ solve path
    def solve_cycles(self, opts=None, **kw):
        """== ebSolveCycles (ebSolveCycles.m:1-80); returns the stage record."""
        o = solve_opts(opts or self.opts, **kw)
        cap = int(o.max_vcycles) + 1
        hist = np.zeros(cap)
        st = _lib.StageStats()
        st.res_hist = dptr(hist)
        st.res_hist_cap = cap
        check(self.lib.bsam_solve_cycles(self.h, ctypes.byref(o), ctypes.byref(st)), "bsam_solve_cycles")
        return _stage(st, hist)

    def comp_residual(self):
        r = _lib.Residual()
        check(self.lib.bsam_comp_residual(self.h, ctypes.byref(r)), "bsam_comp_residual")
        return {"linf": r.linf, "l2": r.l2, "fnorm": r.fnorm, "rel": r.rel, "ncell": r.ncell}

    def vcycle(self):
        check(self.lib.bsam_vcycle(self.h), "bsam_vcycle")

    def fill_ghosts(self, k):
        check(self.lib.bsam_fill_ghosts(self.h, k), "bsam_fill_ghosts")

    def relax_level(self, k, nsweeps):
        check(self.lib.bsam_relax_level(self.h, k, nsweeps), "bsam_relax_level")

    def transfer(self, k, what):
        code = {"restrict": 0, "ring": 1, "correct": 2, "filldown": 3}[what]
        check(self.lib.bsam_transfer(self.h, k, code), "bsam_transfer")

    def coarse_load(self, k):
        check(self.lib.bsam_coarse_load(self.h, k), "bsam_coarse_load")

    def coarsest_solve(self):
        check(self.lib.bsam_coarsest_solve(self.h), "bsam_coarsest_solve")

    def reset_force(self):
        check(self.lib.bsam_reset_force(self.h), "bsam_reset_force")

    def init_level_prolong(self, k):
        check(self.lib.bsam_init_level_prolong(self.h, k), "bsam_init_level_prolong")

    def composite_mean(self):
        m = ctypes.c_double()
        a = ctypes.c_double()
        check(self.lib.bsam_composite_mean(self.h, ctypes.byref(m), ctypes.byref(a)), "bsam_composite_mean")
        return m.value, a.value

    def inherit(self, src):
        """Copy Q and the history of every level whose boxes equal src's
        (bsam_hier_inherit)."""
        check(self.lib.bsam_hier_inherit(self.h, src.h), "bsam_hier_inherit")

    def regrid_level(self, k, old):
        """transferLevel of ebpRegrid.m for level k from the old hierarchy."""
        check(self.lib.bsam_regrid_level(self.h, k, old.h), "bsam_regrid_level")

    def parabolic_uo_from_q(self):
        check(self.lib.bsam_parabolic_uo_from_q(self.h), "bsam_parabolic_uo_from_q")

    def tag_level(self, k, indicator, threshold, buffer):
        """Device tagging of level k (ebTagCluster.m:27-90): returns (T, Traw),
        boolean arrays over the level's index space -- the tags after buffer
        and nesting clip, and the indicator tags before them."""
        code = {"ulap": _lib.BSAM_TAG_ULAP, "rte": _lib.BSAM_TAG_RTE}[indicator.lower()]
        n1, n2 = (int(v) for v in self.H.L(k).n)
        T = np.zeros((n1, n2), dtype=np.uint8, order="F")
        R = np.zeros((n1, n2), dtype=np.uint8, order="F")
        cnt = _lib.c_ll()
        check(self.lib.bsam_tag_level(self.h, k, code, float(threshold), int(buffer), T.ctypes.data, R.ctypes.data,
                                      ctypes.byref(cnt)), "bsam_tag_level")
        return T.astype(bool), R.astype(bool)

    # ------------------------------------------------------------ data
    def _shapes(self, k, field):
        L = self.H.L(k)
        out = []
        for p in range(L.np):
            mx, my = L.dims(p)
            if field == _lib.BSAM_Q:
                out.append((mx + 2, my + 2))
            elif field == _lib.BSAM_QC:
                out.append((mx // 2 + 2, my // 2 + 2))
            else:
                out.append((mx, my))
        return out

    def get_field(self, k, field):
        n = self.lib.bsam_level_field_size(self.h, k, field)
        buf = np.zeros(max(n, 0))
        if n > 0:
            check(self.lib.bsam_get_level(self.h, k, field, dptr(buf)), "bsam_get_level")
        out = []
        off = 0
        for shp in self._shapes(k, field):
            sz = shp[0] * shp[1]
            out.append(buf[off:off + sz].reshape(shp, order="F").copy())
            off += sz
        return out

    def set_field(self, k, field, arrs):
        buf = _packF(arrs)
        check(self.lib.bsam_set_level(self.h, k, field, dptr(buf)), "bsam_set_level")

    def pull_q(self, H=None):
        H = H or self.H
        for k in range(1, self.nlev + 1):
            for p, a in enumerate(self.get_field(k, _lib.BSAM_Q)):
                H.L(k).Q[p] = a

    def export(self, k, what):
        n = self.lib.bsam_export_size(self.h, k, what)
        out = np.zeros(max(n, 0), dtype=np.int64)
        if n > 0:
            check(self.lib.bsam_export(self.h, k, what, out.ctypes.data_as(_lib.P_ll)), "bsam_export")
        return out

    def reset(self):
        check(self.lib.bsam_hier_reset(self.h), "bsam_hier_reset")

    def time_smoother(self, k, nlaunch, mixed=False):
        ms = ctypes.c_double()
        check(self.lib.bsam_time_smoother(self.h, k, 1 if mixed else 0, nlaunch, ctypes.byref(ms)),
              "bsam_time_smoother")
        return ms.value

    def streaming_levels(self):
        """Levels whose sweeps run as streaming sweep launches."""
        return [k for k in range(1, self.H.nlev + 1) if self.lib.bsam_level_streams(self.h, k) == 1]

    def smoother_bytes(self, k, elem_bytes=8):
        return self.lib.bsam_smoother_bytes(self.h, k, elem_bytes)

    # ------------------------------------------------------------ parabolic
    def update_coeffs(self, H=None):
        """ebpSetTime's refresh (ebpSetTime.m:19-51) from the host hierarchy's
        current CC/DeW/DnS/InvDiagS/bc data (same boxes)."""
        d, pk = self._desc(H or self.H)
        check(self.lib.bsam_update_coeffs(self.h, ctypes.byref(d)), "bsam_update_coeffs")

    def _pack_cells(self, U):
        return _packF([U[k - 1][p] for k in range(1, self.nlev + 1) for p in range(self.H.L(k).np)])

    def _unpack_cells(self, buf):
        out, off = [], 0
        for k in range(1, self.nlev + 1):
            lev = []
            for p in range(self.H.L(k).np):
                mx, my = self.H.L(k).dims(p)
                lev.append(buf[off:off + mx * my].reshape((mx, my), order="F").copy())
                off += mx * my
            out.append(lev)
        return out

    def _ncell_all(self):
        return sum(int(np.prod(self.H.L(k).dims(p))) for k in range(1, self.nlev + 1) for p in range(self.H.L(k).np))

    def parabolic_init(self):
        """Uo = interior(Q), no u^{n-1} yet (ebpInit.m:32-35)."""
        check(self.lib.bsam_parabolic_init(self.h), "bsam_parabolic_init")

    def set_history(self, Uo=None, Uoo=None):
        """Upload u^n and/or u^{n-1}: per level, per patch (mx, my) arrays."""
        a = self._pack_cells(Uo) if Uo is not None else None
        b = self._pack_cells(Uoo) if Uoo is not None else None
        check(self.lib.bsam_parabolic_set_history(self.h, dptr(a) if a is not None else None,
                                                  dptr(b) if b is not None else None),
              "bsam_parabolic_set_history")

    def get_history(self, which=0):
        buf = np.zeros(self._ncell_all())
        check(self.lib.bsam_parabolic_get_history(self.h, int(which), dptr(buf)), "bsam_parabolic_get_history")
        return self._unpack_cells(buf)

    def set_source(self, F=None):
        """Resident source f(t1) (per level, per patch); None means f = 0."""
        a = self._pack_cells(F) if F is not None else None
        check(self.lib.bsam_parabolic_set_source(self.h, dptr(a) if a is not None else None),
              "bsam_parabolic_set_source")

    def parabolic_set_shift(self, aDt):
        """The shift alpha/Dt that the resident CC carries (H.par.aDt)."""
        check(self.lib.bsam_parabolic_set_shift(self.h, float(aDt)), "bsam_parabolic_set_shift")

    def parabolic_rhs(self, dt, scheme):
        check(self.lib.bsam_parabolic_rhs(self.h, float(dt), _scheme_code(scheme)), "bsam_parabolic_rhs")

    def parabolic_advance(self):
        check(self.lib.bsam_parabolic_advance(self.h), "bsam_parabolic_advance")

    def parabolic_step(self, dt, scheme, opts=None, **kw):
        """One accepted step (ebpSolve.m:40-58): rhs, ebSolveCycles, history,
        compMass.  Returns (stage record, mass)."""
        o = solve_opts(opts or self.opts, **kw)
        cap = int(o.max_vcycles) + 1
        hist = np.zeros(cap)
        st = _lib.StageStats()
        st.res_hist = dptr(hist)
        st.res_hist_cap = cap
        m = ctypes.c_double()
        check(self.lib.bsam_parabolic_step(self.h, float(dt), _scheme_code(scheme), ctypes.byref(o),
                                           ctypes.byref(st), ctypes.byref(m)), "bsam_parabolic_step")
        return _stage(st, hist), m.value


def _scheme_code(scheme):
    codes = {"be": _lib.BSAM_BE, "bdf2": _lib.BSAM_BDF2, "cn": _lib.BSAM_CN}
    if scheme not in codes:
        raise _lib.BsamError("ebp_rhs:scheme", f"unknown scheme {scheme}")
    return codes[scheme]


def _stage(st, hist):
    return {"amrLevels": st.amr_levels, "nlev": st.nlev, "vcycles": st.vcycles, "finalRel": st.final_rel,
            "visibleCells": st.visible_cells, "totalCells": st.total_cells, "avgFactor": st.avg_factor,
            "resHist": hist[:st.res_hist_len].tolist(), "gpu_ms": st.gpu_ms}
