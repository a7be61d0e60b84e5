"""Hierarchies used by the parity tests (host layer only, no solves).

Fixed box layouts in the spirit of the reference's injected-layout stress
tests (EllipticBSAM/tests/stressInterfaces.m:1-40): same-level faces, coarse-
fine corners, reentrant corners, patches on the physical boundary (ring
overhang into parent ghosts), plus the README elliptic layout.
"""
from __future__ import annotations

import numpy as np

from bsamgpu import Options, Problem, configs, hierarchy


def _randomize(H, seed):
    rng = np.random.default_rng(seed)
    for k in range(1, H.nlev + 1):
        L = H.L(k)
        for p in range(L.np):
            L.Q[p] = rng.standard_normal(L.Q[p].shape)
            L.FTrue[p] = L.FTrue[p] + 0.1 * rng.standard_normal(L.FTrue[p].shape)


def _vardc_problem(bc):
    D = lambda x, y: 1 + 0.5 * np.sin(2 * np.pi * x) * np.cos(np.pi * y)  # noqa: E731
    C = lambda x, y: 1 + x * y  # noqa: E731
    f = lambda x, y: np.sin(3 * x) * np.cos(2 * y) + 1  # noqa: E731
    return Problem([0, 1, 0, 1], D, C, f, bc)


def plain(seed=None, **kw):
    """README elliptic hierarchy: 16, 32 (root), 64 (1 patch), 4 patches at 128."""
    prob = configs.plain_problem()
    opts = configs.plain_options(**kw)
    H = hierarchy.setup(prob, opts)
    hierarchy.add_level(H, [[1, 32, 1, 32]])
    hierarchy.add_level(H, [[7, 26, 7, 26], [7, 26, 39, 58], [39, 58, 7, 26], [39, 58, 39, 58]])
    if seed is not None:
        _randomize(H, seed)
    return H


def lshape(seed=None, bc=("dirichlet", lambda x, y: x - 2 * y), interp="qif", **kw):
    """Two adaptive levels with an L-shaped union, an abutting patch, a
    domain-corner patch and a nested level with same-level faces."""
    prob = _vardc_problem(bc)
    o = dict(BaseCells=32, MaxLevels=3, GhostInterp=interp, RelTol=1e-10, Verbose=0)
    o.update(kw)
    opts = Options(**o)
    H = hierarchy.setup(prob, opts)
    hierarchy.add_level(H, [[5, 12, 5, 8], [5, 8, 9, 12], [13, 16, 5, 6], [1, 4, 25, 32], [27, 32, 13, 20]])
    hierarchy.add_level(H, [[10, 15, 10, 22], [16, 23, 10, 15], [1, 6, 51, 62], [55, 62, 27, 38]])
    if seed is not None:
        _randomize(H, seed)
    return H


def neumann(seed=None, **kw):
    """Singular all-Neumann, C = 0 problem (KKT coarsest system)."""
    f = lambda x, y: np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)  # noqa: E731
    prob = Problem([0, 1, 0, 1], 1, 0, f, ("neumann", 0))
    o = dict(BaseCells=32, MaxLevels=2, RelTol=1e-10, Verbose=0)
    o.update(kw)
    H = hierarchy.setup(prob, Options(**o))
    hierarchy.add_level(H, [[9, 24, 9, 20], [25, 28, 9, 16]])
    if seed is not None:
        rng = np.random.default_rng(seed)
        for k in range(1, H.nlev + 1):
            L = H.L(k)
            for p in range(L.np):
                L.Q[p] = rng.standard_normal(L.Q[p].shape)
    return H


def mixed_bc(seed=None, **kw):
    bc = {"west": ("dirichlet", lambda x, y: np.sin(3 * y)), "east": ("neumann", 0.5),
          "south": ("neumann", lambda x, y: x), "north": ("dirichlet", 1.0)}
    return lshape(seed=seed, bc=bc, **kw)


CASES = {
    "plain": lambda seed=None: plain(seed),
    "lshape": lambda seed=None: lshape(seed),
    "lshape_lif": lambda seed=None: lshape(seed, interp="lif"),
    "neumann": lambda seed=None: neumann(seed),
    "mixed_bc": lambda seed=None: mixed_bc(seed),
}
