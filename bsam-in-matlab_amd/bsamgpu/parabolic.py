"""Parabolic stepping (ParabolicBSAM) on the device library.

    du/dt - div(D grad u) + C u = f

Every implicit step is one warm-started AFAS solve with C_eff = C + alpha/dt.
The step itself -- right-hand side from the device-resident history,
ebSolveCycles, history rotation and the composite mass -- is one call into
the library (``bsam_parabolic_step``); the hierarchy and the BDF2 history
stay in HBM between steps.  What stays on the host is what the reference
evaluates from user functions or decides structurally:

  * ``ParabolicProblem``   <- ParabolicBSAM/src/ebpProblem.m:1-78
  * ``parabolic_options``  <- ParabolicBSAM/src/ebpOptions.m:1-57
  * ``eprob``              <- ParabolicBSAM/src/ebpEprob.m:1-19
  * ``ebp_init``           <- ParabolicBSAM/src/ebpInit.m:1-61
  * ``set_time``           <- ParabolicBSAM/src/ebpSetTime.m:1-59
  * ``ebp_regrid``         <- ParabolicBSAM/src/ebpRegrid.m:1-162
  * ``ebp_solve``          <- ParabolicBSAM/src/ebpSolve.m:1-107

Schemes: 'be', 'bdf2' (BE bootstrap on the first step) and 'cn'.  The
Crank-Nicolson right-hand side applies the old-time composite operator, flux
corrections included, on the device before the coefficients move to t1
(``cn_rhs``; ParabolicBSAM/src/ebpRhs.m:15-67).
"""
from __future__ import annotations

import dataclasses
import inspect
import math
from typing import Any, Callable, Optional

import numpy as np

from . import _lib, hierarchy, tagcluster
from .device import GpuHierarchy
from .problem import SIDES, BcSide, BsamError, Options, Problem

ALPHA = {"be": 1.0, "bdf2": 1.5, "cn": 2.0}


# ---------------------------------------------------------------- problem
def _arity(fn):
    try:
        ps = inspect.signature(fn).parameters.values()
    except (TypeError, ValueError):
        return -1
    if any(p.kind == p.VAR_POSITIONAL for p in ps):
        return -1
    return sum(1 for p in ps if p.default is p.empty and p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD))


def _wrap_t(v, label):
    """Normalise to fn(X, Y, T); the flag says whether it depends on T
    (ebpProblem.m:61-78: arity 3 or varargs -> time-dependent)."""
    if callable(v):
        na = _arity(v)
        if na == 3 or na < 0:
            return v, True
        if na == 2:
            return (lambda X, Y, T, _f=v: _f(X, Y)), False
        raise BsamError("ebp_problem:arity", f"{label} must take (x,y) or (x,y,t).")
    if np.isscalar(v) and np.isreal(v) and not isinstance(v, (str, bytes)):
        vv = float(v)
        return (lambda X, Y, T, _v=vv: _v * np.ones(np.shape(X))), False
    raise BsamError("ebp_problem:badinput", f"{label} must be a handle or scalar.")


class ParabolicProblem:
    """du/dt - div(D grad u) + C u = f on domain, u(.,0) = u0 (ebpProblem.m:1-59)."""

    def __init__(self, domain, D, C, f, bc, u0, exact=None):
        domain = [float(v) for v in np.ravel(domain)]
        if len(domain) != 4 or domain[1] <= domain[0] or domain[3] <= domain[2]:
            raise BsamError("ebp_problem:domain", "domain must be [a b c d].")
        self.domain = domain
        self.D, self.tdD = _wrap_t(D, "D")
        self.C, self.tdC = _wrap_t(C, "C")
        self.f, self.tdF = _wrap_t(f, "f")
        if isinstance(bc, (tuple, list)) and len(bc) == 2 and isinstance(bc[0], str):
            bc = {s: bc for s in SIDES}
        self.bc = []
        self.tdBC = False
        for s in SIDES:
            t, g = bc[s]
            t = t.lower()
            if t not in ("dirichlet", "neumann"):
                raise BsamError("ebp_problem:bc", "bc type must be dirichlet or neumann")
            gf, td = _wrap_t(g, "g_" + s)
            self.bc.append(BcSide(t, gf))
            self.tdBC = self.tdBC or td
        if not callable(u0):
            raise BsamError("ebp_problem:u0", "u0 must be @(X,Y).")
        self.u0 = u0
        self.exact = exact


@dataclasses.dataclass
class ParabolicOptions:
    """ebpOptions.m:14-57; ``eb`` holds the forwarded elliptic options."""

    Scheme: str = "bdf2"
    Dt: Optional[float] = None
    TFinal: Optional[float] = None
    RegridEvery: int = 8
    StepVerbose: int = 10
    Callback: Optional[Callable] = None
    MaxVCyclesStep: int = 20
    eb: Any = None


_TIME_FIELDS = {f.name.lower(): f.name for f in dataclasses.fields(ParabolicOptions) if f.name != "eb"}


def parabolic_options(**kw) -> ParabolicOptions:
    """Split name-value pairs into time-stepping fields and elliptic options
    (ebpOptions.m:22-57)."""
    t = {}
    rest = {}
    for k, v in kw.items():
        if k.lower() in _TIME_FIELDS:
            t[_TIME_FIELDS[k.lower()]] = v
        else:
            rest[k] = v
    o = ParabolicOptions(**t)
    o.Scheme = str(o.Scheme).lower()
    if o.Scheme not in ALPHA:
        raise BsamError("ebp_options:scheme", "Scheme must be be, bdf2 or cn.")
    if o.Dt is None or o.TFinal is None:
        raise BsamError("ebp_options:time", "Dt and TFinal are required.")
    for name in ("Dt", "TFinal"):
        v = getattr(o, name)
        if not (np.isscalar(v) and np.isfinite(v) and v > 0):
            raise BsamError("ebp_options:" + name, f"{name} must be a finite positive scalar.")
    for name in ("RegridEvery", "StepVerbose"):
        v = getattr(o, name)
        if not (float(v) == int(v) and v >= 0):
            raise BsamError("ebp_options:" + name, f"{name} must be a nonnegative integer.")
    if not (float(o.MaxVCyclesStep) == int(o.MaxVCyclesStep) and o.MaxVCyclesStep > 0):
        raise BsamError("ebp_options:MaxVCyclesStep", "MaxVCyclesStep must be a positive integer.")
    if o.Callback is not None and not callable(o.Callback):
        raise BsamError("ebp_options:callback", "Callback must be empty or a function handle.")
    o.eb = Options(**rest).replace(Verbose=0, MaxVCycles=int(o.MaxVCyclesStep))
    return o


def eprob(pp: ParabolicProblem, t, aDt) -> Problem:
    """The elliptic problem frozen at time t with C_eff = C + aDt (ebpEprob.m:7-19)."""
    bc = {}
    for s, side in zip(SIDES, pp.bc):
        bc[s] = (side.type, (lambda x, y, _g=side.g: _g(x, y, t)))
    return Problem(pp.domain, lambda x, y: pp.D(x, y, t), lambda x, y: pp.C(x, y, t) + aDt,
                   lambda x, y: np.zeros(np.shape(x)), bc)


def _eval(fn, X, Y, T):
    v = fn(X, Y, T)
    if np.isscalar(v) or np.ndim(v) == 0:
        v = float(v) * np.ones(X.shape)
    v = np.asarray(v, dtype=float)
    if v.shape != X.shape:
        v = v.reshape(X.shape)
    return v


def _set_u0(H, pp, ks=None):
    """setU0 (ebpInit.m:39-51): u0 at cell centres, analytically."""
    for k in (ks if ks is not None else range(1, H.nlev + 1)):
        lev = H.L(k)
        for p in range(lev.np):
            X, Y = lev.cell_centers(pp.domain, p)
            v = pp.u0(X, Y)
            if np.isscalar(v) or np.ndim(v) == 0:
                v = float(v) * np.ones(X.shape)
            lev.Q[p][1:-1, 1:-1] = v


def _interiors(H):
    return [[H.L(k).Q[p][1:-1, 1:-1].copy() for p in range(H.L(k).np)] for k in range(1, H.nlev + 1)]


# ---------------------------------------------------------------- state
class _State:
    """Host record of H.par (ebpInit.m:31-35) and the device handle."""

    def __init__(self, H, G, aDt, key):
        self.H = H
        self.G = G
        self.aDt = aDt
        self.stateKey = key
        self.src_key = None


def _add_level(H, G, boxes):
    """ebAddLevel on the host tables; the new device handle receives the
    pulled state (the prolonged initial guess is overwritten by the callers)."""
    G.pull_q(H)
    G.close()
    kNew = hierarchy.add_level(H, boxes)
    return kNew, GpuHierarchy(H)


def ebp_init(pp: ParabolicProblem, opts: ParabolicOptions, aDt1):
    """ebpInit.m:7-35: hierarchy adapted to u0, history Uo = u0."""
    H = hierarchy.setup(eprob(pp, 0.0, aDt1), opts.eb)
    _set_u0(H, pp)
    G = GpuHierarchy(H)
    for k in range(1, H.nlev + 1):
        G.fill_ghosts(k)
    G.pull_q(H)
    while H.nlev - H.rootIdx + 1 < opts.eb.MaxLevels:
        boxes = tagcluster.tag_cluster(H, device=G)
        if not boxes:
            break
        kNew, G = _add_level(H, G, boxes)
        _set_u0(H, pp, [kNew])
        G.set_field(kNew, _lib.BSAM_Q, H.L(kNew).Q)
        G.fill_ghosts(kNew)
        G.pull_q(H)
    G.transfer(0, "filldown")
    G.parabolic_init()
    td = pp.tdD or pp.tdC or pp.tdBC
    return _State(H, G, aDt1, (0.0 * td, aDt1))


def set_time(S: _State, pp: ParabolicProblem, t, aDt, force=False):
    """ebpSetTime.m:14-55: re-evaluate CC = C(t) + aDt (and D, boundary data
    when they depend on t), coefficient fill-down on the host, then the
    device refresh of everything derived from them."""
    key = (t * (pp.tdD or pp.tdC or pp.tdBC), aDt)
    if not force and S.stateKey == key:
        return
    H = S.H
    dom = pp.domain
    for k in range(1, H.nlev + 1):
        lev = H.L(k)
        hx, hy = lev.h
        for p in range(lev.np):
            b = lev.box[p]
            xc = dom[0] + (np.arange(b[0], b[1] + 1, dtype=float) - 0.5) * hx
            yc = dom[2] + (np.arange(b[2], b[3] + 1, dtype=float) - 0.5) * hy
            if pp.tdD or force:
                xf = dom[0] + np.arange(b[0] - 1, b[1] + 1, dtype=float) * hx
                yf = dom[2] + np.arange(b[2] - 1, b[3] + 1, dtype=float) * hy
                XF, YC = hierarchy.ndgrid(xf, yc)
                lev.DeW[p] = _eval(pp.D, XF, YC, t)
                XC, YF = hierarchy.ndgrid(xc, yf)
                lev.DnS[p] = _eval(pp.D, XC, YF, t)
            X, Y = hierarchy.ndgrid(xc, yc)
            lev.CC[p] = _eval(pp.C, X, Y, t) + aDt
            if pp.tdBC or force:
                fint = lev.faceInternal[p]
                gs = lev.bc_g[p]
                if not fint[0]:
                    gs[0] = np.ravel(_eval(pp.bc[0].g, dom[0] * np.ones(yc.shape), yc, t))
                if not fint[1]:
                    gs[1] = np.ravel(_eval(pp.bc[1].g, dom[1] * np.ones(yc.shape), yc, t))
                if not fint[2]:
                    gs[2] = np.ravel(_eval(pp.bc[2].g, xc, dom[2] * np.ones(xc.shape), t))
                if not fint[3]:
                    gs[3] = np.ravel(_eval(pp.bc[3].g, xc, dom[3] * np.ones(xc.shape), t))
    hierarchy.coeff_filldown(H)
    S.G.update_coeffs(H)
    S.stateKey = key
    S.aDt = aDt


def _source(S: _State, pp: ParabolicProblem, t1):
    """f(X, Y, t1) at cell centres of levels >= root (ebpRhs.m:28-36), kept
    resident on the device; re-sent only when f depends on t or the state
    (hierarchy) is new."""
    key = t1 if pp.tdF else "constant"
    if S.src_key == key:
        return
    H = S.H
    F = []
    zero = True
    for k in range(1, H.nlev + 1):
        lev = H.L(k)
        row = []
        for p in range(lev.np):
            X, Y = lev.cell_centers(pp.domain, p)
            v = _eval(pp.f, X, Y, t1) if k >= H.rootIdx else np.zeros(X.shape)
            zero = zero and not np.any(v) and not np.any(np.signbit(v))
            row.append(v)
        F.append(row)
    S.G.set_source(None if zero else F)
    S.src_key = key


def cn_rhs(S: _State, pp: ParabolicProblem, t1, Dt):
    """ebpRhs.m cn branch: the explicit half with the operator of the old
    time, so it runs before set_time moves the coefficients to t1; the
    source is taken at the midpoint t1 - Dt/2."""
    _source(S, pp, t1 - Dt / 2)
    S.G.parabolic_set_shift(S.aDt)
    S.G.parabolic_rhs(Dt, "cn")


def _cn_step(S: _State, pp: ParabolicProblem, t1, Dt, eopts):
    """ebpSolve.m:46-58 for cn: rhs, operator of t1, solve, history, mass."""
    cn_rhs(S, pp, t1, Dt)
    set_time(S, pp, t1, ALPHA["cn"] / Dt)
    st = S.G.solve_cycles(eopts)
    S.G.parabolic_advance()
    m, a = S.G.composite_mean()
    return st, m * a


# ---------------------------------------------------------------- regrid
def ebp_regrid(S: _State, pp: ParabolicProblem, t, aDtNext, opts: ParabolicOptions):
    """ebpRegrid.m:15-59: rebuild around the current solution and transfer
    the state conservatively; returns the new state.  The old hierarchy stays
    on the device until the transfer is done: the root copy, the sub-root
    restriction, the per-level conservative-linear transfer of Q and Uoo
    (bsam_regrid_level), tagging and the history all run there; the host
    builds the new levels' boxes and coefficients."""
    Hold, Gold = S.H, S.G
    Hn = hierarchy.setup(eprob(pp, t, aDtNext), opts.eb)
    r = Hn.rootIdx
    G = GpuHierarchy(Hn)
    G.regrid_level(r, Gold)
    for k in range(r, 1, -1):  # sub-root levels: restriction of the root
        G.transfer(k, "restrict")
    for k in range(1, r + 1):
        G.fill_ghosts(k)
    while Hn.nlev - Hn.rootIdx + 1 < opts.eb.MaxLevels:
        boxes = tagcluster.tag_cluster(Hn, device=G)
        if not boxes:
            break
        kN = hierarchy.add_level(Hn, boxes)
        Gn = GpuHierarchy(Hn)
        Gn.inherit(G)
        G.close()
        G = Gn
        G.regrid_level(kN, Gold)
        G.fill_ghosts(kN)
    Gold.close()
    G.transfer(0, "filldown")
    G.parabolic_uo_from_q()
    G.pull_q(Hn)
    Hn.stats = Hold.stats
    td = pp.tdD or pp.tdC or pp.tdBC
    return _State(Hn, G, aDtNext, (t * td, aDtNext))


# ---------------------------------------------------------------- solve
class Solution(dict):
    __getattr__ = dict.__getitem__


def _matlab_round(x):
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def ebp_solve(pp: ParabolicProblem, opts: ParabolicOptions | None = None, **kw) -> Solution:
    """Implicit time stepping, one AFAS solve per step (ebpSolve.m:17-98)."""
    opts = opts if opts is not None else parabolic_options(**kw)
    nsteps = max(1, _matlab_round(opts.TFinal / opts.Dt))
    Dt = opts.TFinal / nsteps
    scheme1 = "be" if opts.Scheme == "bdf2" else opts.Scheme
    S = ebp_init(pp, opts, ALPHA[scheme1] / Dt)
    m, a = S.G.composite_mean()
    mass = [m * a]
    cycles, relres, stages, levels_hist = [], [], [], []
    nregrids = 0
    for n in range(1, nsteps + 1):
        t1 = n * Dt
        scheme = scheme1 if n == 1 else opts.Scheme
        if scheme == "cn":
            st, mn = _cn_step(S, pp, t1, Dt, opts.eb)
        else:
            _source(S, pp, t1)
            # be/bdf2 right-hand sides do not read the operator, so the
            # refresh may precede them (the reference refreshes after ebpRhs)
            set_time(S, pp, t1, ALPHA[scheme] / Dt)
            st, mn = S.G.parabolic_step(Dt, scheme, opts.eb)
        S.H.stats["stages"].append(st)
        stages.append(st)
        cycles.append(st["vcycles"])
        relres.append(st["finalRel"])
        mass.append(mn)
        levels_hist.append((S.H.nlev - S.H.rootIdx + 1, int(st["totalCells"])))
        if opts.Callback is not None:
            opts.Callback(S, t1, n)
        if opts.StepVerbose > 0 and (n % opts.StepVerbose == 0 or n == nsteps):
            print(f"[ebsam-p] step {n:4d}/{nsteps}  t={t1:.4f}  V={st['vcycles']}  rel={st['finalRel']:.1e}  "
                  f"levels={S.H.nlev - S.H.rootIdx + 1}  cells={st['totalCells']}  mass={mn:.6e}")
        if opts.RegridEvery > 0 and n % opts.RegridEvery == 0 and n < nsteps and opts.eb.MaxLevels > 1:
            S = ebp_regrid(S, pp, t1, ALPHA[opts.Scheme] / Dt, opts)
            nregrids += 1
    S.G.pull_q(S.H)
    H = S.H
    sol = Solution(H=H, device=S.G, t=[i * Dt for i in range(nsteps + 1)], mass=mass, cycles=cycles,
                   relres=relres, nregrids=nregrids, stats=H.stats, stages=stages, levels_hist=levels_hist,
                   Dt=Dt)
    sol["levels"] = [{"n": H.L(k).n, "np": H.L(k).np,
                      "cells": sum(int(np.prod(H.L(k).dims(p))) for p in range(H.L(k).np))}
                     for k in range(H.rootIdx, H.nlev + 1)]
    if pp.exact is not None:
        from . import diagnostics
        T = opts.TFinal
        sol["err"] = diagnostics.error_norms(H, lambda x, y: pp.exact(x, y, T))
    else:
        sol["err"] = None
    return sol
