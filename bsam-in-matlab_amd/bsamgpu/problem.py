"""Problem and option records of the host layer.

Mirrors the reference's host-side value types:
  * ``Problem``  <- EllipticBSAM/src/ebProblem.m:1-63
  * ``Options``  <- EllipticBSAM/src/ebOptions.m:1-74 (names and defaults kept)

Coefficients, sources and boundary data are plain Python callables
``fn(X, Y) -> ndarray | scalar`` evaluated on the host; the device library only
ever sees the sampled FP64 arrays.
"""
from __future__ import annotations

import dataclasses
import inspect
from typing import Any, Callable, Optional

import numpy as np

SIDES = ("west", "east", "south", "north")


class BsamError(ValueError):
    """Error carrying a reference-style identifier such as ``eb_problem:bc``."""

    def __init__(self, ident: str, msg: str):
        super().__init__(f"{ident}: {msg}")
        self.ident = ident


def _tofun(v, label):
    # ebProblem.m:57-63 -- scalar -> constant handle, handle passes through
    if callable(v):
        return v
    if np.isscalar(v) and np.isreal(v):
        vv = float(v)
        return lambda X, Y, _v=vv: _v * np.ones(np.shape(X))
    raise BsamError("eb_problem:badinput", f"{label} must be a function handle or scalar.")


@dataclasses.dataclass
class BcSide:
    type: str  # 'dirichlet' | 'neumann'
    g: Callable


class Problem:
    """-div(D grad u) + C u = f on domain=[a b c d]  (ebProblem.m:1-55)."""

    def __init__(self, domain, D, C, f, bc, exact=None):
        domain = [float(v) for v in np.ravel(domain)]
        if len(domain) != 4 or domain[1] <= domain[0] or domain[3] <= domain[2]:
            raise BsamError("eb_problem:domain", "domain must be [a b c d] with b>a, d>c.")
        self.domain = domain
        self.D = _tofun(D, "D")
        self.C = _tofun(C, "C")
        self.f = _tofun(f, "f")
        # ebProblem.m:29-48: {type, g} shorthand or per-side mapping
        if isinstance(bc, (tuple, list)) and len(bc) == 2 and isinstance(bc[0], str):
            bc = {s: bc for s in SIDES}
        self.bc = []
        for s in SIDES:
            if s not in bc:
                raise BsamError("eb_problem:bc", f"bc missing side {s}")
            t, g = bc[s]
            t = t.lower()
            if t not in ("dirichlet", "neumann"):
                raise BsamError("eb_problem:bc", "bc type must be dirichlet or neumann")
            self.bc.append(BcSide(t, _tofun(g, "g_" + s)))
        self.exact = _tofun(exact, "exact") if exact is not None else None


@dataclasses.dataclass
class Options:
    """ebOptions.m:44-66 name/value pairs with the reference defaults."""

    BaseCells: Any = (64, 64)
    MaxLevels: int = 1
    CoarsestCells: int = 16
    GhostInterp: str = "qif"
    MassCorrection: bool = True
    NPre: int = 3
    NPost: int = 3
    Omega: float = 1.0
    MaxVCycles: int = 60
    RelTol: float = 1e-9
    Indicator: str = "ulap"
    TagFunction: Optional[Callable] = None
    TagThreshold: Any = 1e-3
    TagBuffer: int = 2
    MinBoxWidth: int = 4
    MinFillRatio: float = 0.70
    Verbose: int = 1
    InitialGuess: Optional[Callable] = None
    MeanZero: Any = "auto"
    SourceQuad: int = 1
    # device-path extras (no reference counterpart): arithmetic of the inner
    # V-cycles; 'fp64' is the plain path, 'mixed' runs FP32 V-cycles inside an
    # FP64 defect-correction loop.
    Precision: str = "fp64"
    # smoother form: 'exact' reproduces ebRelaxLevel's sequence bit for bit;
    # 'tiled' runs every colour pass of a relax call in one LDS-staged launch
    # (same converged solution, sibling ghosts frozen within a call); 'sweep'
    # is 'exact' with pairs of red+black sweeps fused into one launch each on
    # levels whose tiles fit one round of workgroups (bit-identical); 'passes'
    # runs colour passes everywhere ('exact' streams whole sweeps on large
    # one-patch levels with physical sides, bit-identical)
    Smoother: str = "exact"

    def __post_init__(self):
        bc = np.round(np.ravel(self.BaseCells)).astype(int)
        if bc.size == 1:
            bc = np.array([bc[0], bc[0]])
        self.BaseCells = (int(bc[0]), int(bc[1]))
        self.GhostInterp = self.GhostInterp.lower()
        if self.GhostInterp not in ("qif", "lif"):
            raise BsamError("eb_options:ghost", "GhostInterp must be qif or lif")
        if self.Precision not in ("fp64", "mixed"):
            raise BsamError("eb_options:precision", "Precision must be fp64 or mixed")

    def replace(self, **kw) -> "Options":
        return dataclasses.replace(self, **kw)


def call_tagfun(fn, X, Y, level):
    """TagFunction with 2 or 3 arguments (ebTagCluster.m:44-49)."""
    try:
        nargs = len(inspect.signature(fn).parameters)
    except (TypeError, ValueError):
        nargs = 2
    if nargs >= 3:
        return np.asarray(fn(X, Y, level), dtype=bool)
    return np.asarray(fn(X, Y), dtype=bool)
