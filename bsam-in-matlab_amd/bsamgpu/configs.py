"""The benchmark and parity workloads, as problem + option pairs.

1. plain      -- README elliptic example, -Lap u + u = f on [0,1]^2, Dirichlet 0
                 (reference README.md:50-54); BaseCells 32, MaxLevels 3.
2. varcoef    -- seeded variable-D Dirichlet problem, BaseCells 256, MaxLevels 4.
3. starfish   -- diffuse-domain starfish transmission problem
                 (examples/demoStarfish.m:11-23), 64 -> 1024, 5 AMR levels.
4. heat       -- BDF2 heat diffusion (ParabolicBSAM README), BaseCells 64.
5. served     -- per-request random elliptic problems for the serving loop.

Random parameters are drawn with NumPy generators from fixed seeds and are
also written to JSON fixtures under tests/golden so both sides read them.
"""
from __future__ import annotations

import numpy as np

from .problem import Options, Problem


def plain_problem():
    u = lambda x, y: np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)  # noqa: E731
    f = lambda x, y: (1 + 8 * np.pi ** 2) * u(x, y)  # noqa: E731
    return Problem([0, 1, 0, 1], 1, 1, f, ("dirichlet", 0), u)


def plain_options(**kw):
    base = dict(BaseCells=32, MaxLevels=3, TagThreshold=1e-2, RelTol=1e-10, Verbose=0)
    base.update(kw)
    return Options(**base)


def varcoef_params(seed=2):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-0.25, 0.25, 4)
    m = rng.integers(1, 5, 4)
    n = rng.integers(1, 5, 4)
    c = rng.standard_normal(8)
    p = rng.integers(1, 9, 8)
    q = rng.integers(1, 9, 8)
    return {"a": a.tolist(), "m": m.tolist(), "n": n.tolist(), "c": c.tolist(), "p": p.tolist(), "q": q.tolist()}


def varcoef_problem(par=None):
    par = par or varcoef_params()
    a, m, n = np.array(par["a"]), np.array(par["m"]), np.array(par["n"])
    c, p, q = np.array(par["c"]), np.array(par["p"]), np.array(par["q"])

    def D(x, y):
        s = 0.0
        for k in range(4):
            s = s + a[k] * np.sin(2 * np.pi * m[k] * x) * np.cos(2 * np.pi * n[k] * y)
        return 1 + 0.5 * s

    def f(x, y):
        s = 0.0 * x
        for j in range(8):
            s = s + c[j] * np.sin(np.pi * p[j] * x) * np.sin(np.pi * q[j] * y)
        return s

    return Problem([0, 1, 0, 1], D, 1, f, ("dirichlet", 0))


def varcoef_options(**kw):
    base = dict(BaseCells=256, MaxLevels=4, TagThreshold=1e-3, RelTol=1e-10, Verbose=0)
    base.update(kw)
    return Options(**base)


# The ulap indicator is h^2 |lap u| (ebTagCluster.m:53-55); on the 256^2 root of
# the variable-coefficient problem it peaks at 2.6e-5 and drops about 4x per
# level, so the specified threshold 1e-3 tags nothing.  At 1e-6 the stages
# refine to the 2048^2 spacing the configuration names: 8 levels, a 1024^2
# level as one patch and 20 patches at 2048^2, 1,747,976 cells in all.
VARCOEF_REFINED_THRESHOLD = 1e-6


def varcoef_refined_options(**kw):
    base = dict(TagThreshold=VARCOEF_REFINED_THRESHOLD)
    base.update(kw)
    return varcoef_options(**base)


STARFISH_DATA = dict(alpha=3.0, beta=2.0, gamma=1.0, kappa=0.01, g=4.0)


def starfish_problem(epsilon=0.05, npts=3000, device=None):
    """``device``: evaluate the distance and the diffuse-domain fields on that
    HIP device (bsam_polygon_fields) instead of the host."""
    from .geometry import starfish_geometry, transmission_problem

    g = starfish_geometry(npts, device)
    data = dict(STARFISH_DATA)
    data["q"] = lambda x, y: 15 - x ** 2
    data["h"] = lambda x, y: 2.5 * np.sin(x) + np.exp(np.cos(y))
    prob, aux = transmission_problem(g, epsilon, [-2, 2, -2, 2], data)
    return prob, aux, g


def starfish_options(geom, base=64, levels=5, epsilon=0.05, **kw):
    sd = geom.signed_distance

    def tag(x, y, L):
        return np.abs(sd(x, y)) < max(3 * epsilon, 3 * 4 / (base * 2 ** (L - 1)))

    o = dict(BaseCells=base, MaxLevels=levels, Indicator="fun", TagFunction=tag, TagBuffer=2,
             RelTol=1e-10, MaxVCycles=100, Verbose=0)
    o.update(kw)
    return Options(**o)


def served_params(nreq=512, seed=5):
    """Per-request parameters of the served workload: BaseCells ~ U{32,64,128},
    MaxLevels ~ U{2..5}, D = 1 + sum_k a_k sin(2 pi m_k x) cos(2 pi n_k y) with
    sum |a_k| <= 0.5, C ~ U(0,2), f = sum_j c_j sin(pi p_j x) sin(pi q_j y) with
    c_j ~ N(0,1), each side Dirichlet 0 or Neumann 0 with probability 1/2."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nreq):
        out.append({
            "id": i,
            "base": int(rng.choice([32, 64, 128])),
            "levels": int(rng.integers(2, 6)),
            "a": rng.uniform(-0.125, 0.125, 4).tolist(),
            "m": rng.integers(1, 5, 4).tolist(),
            "n": rng.integers(1, 5, 4).tolist(),
            "C": float(rng.uniform(0.0, 2.0)),
            "c": rng.standard_normal(8).tolist(),
            "p": rng.integers(1, 9, 8).tolist(),
            "q": rng.integers(1, 9, 8).tolist(),
            "bc": ["dirichlet" if b else "neumann" for b in rng.integers(0, 2, 4)],
        })
    return out


def served_arrivals(nreq, rate, seed=55):
    """Poisson arrival times (s) at `rate` requests per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, nreq)).tolist()


def served_problem(par):
    a, m, n = np.array(par["a"]), np.array(par["m"]), np.array(par["n"])
    c, p, q = np.array(par["c"]), np.array(par["p"]), np.array(par["q"])

    def D(x, y):
        s = 0.0
        for k in range(4):
            s = s + a[k] * np.sin(2 * np.pi * m[k] * x) * np.cos(2 * np.pi * n[k] * y)
        return 1 + s

    def f(x, y):
        s = 0.0 * x
        for j in range(8):
            s = s + c[j] * np.sin(np.pi * p[j] * x) * np.sin(np.pi * q[j] * y)
        return s

    bc = {side: (t, 0) for side, t in zip(("west", "east", "south", "north"), par["bc"])}
    return Problem([0, 1, 0, 1], D, par["C"], f, bc)


def served_options(par, **kw):
    base = dict(BaseCells=par["base"], MaxLevels=par["levels"], TagThreshold=1e-3, RelTol=1e-8, Verbose=0)
    base.update(kw)
    return Options(**base)


def heat_initial():
    return lambda x, y: 1 + 0.2 * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)
