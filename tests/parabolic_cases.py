"""Parabolic problems of the reference's parabolic suites, sized for tests.

  heat_readme   README.md:61-69 (published log verification/readme-examples.txt:23-27)
  heat_config   the heat problem at BaseCells 64 / MaxLevels 4, per-step RelTol 1e-12
  mms_bump      ParabolicBSAM/tests/runParabolicTests.m:21-33 (P1: traveling
                Gaussian, time-dependent Dirichlet data and source, regrid every 2 steps)
  mms_bump_64   the same at the next resolution of P1b (runParabolicTests.m:68-84,
                BaseCells 64, Dt 0.025): its 64^2 root, with Dirichlet data that
                change every step, runs region launches that write its ghosts
  reaction_t    runParabolicRegressions.m:6-8 (C = 1 + t, f = t e^-t, Neumann)
  diffusion_t   runParabolicRegressions.m:22-25 (D = 1 + t, time-dependent Dirichlet data)
  fixed_patch   runParabolicRegressions.m:43-50 (D = 1 + 0.2x, a fixed fine patch)
  *_cn          the last three with Scheme 'cn', as runParabolicRegressions.m runs them
"""
import numpy as np

from bsamgpu import hierarchy, tagcluster
from bsamgpu import parabolic as par


class Host:
    """Host decisions handed to the oracle (shared with the product layer)."""
    setup = staticmethod(hierarchy.setup)
    add_level = staticmethod(hierarchy.add_level)
    tag_cluster = staticmethod(tagcluster.tag_cluster)
    coeff_filldown = staticmethod(hierarchy.coeff_filldown)


def heat_u0(x, y):
    return 1 + 0.2 * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)


def heat_readme(**kw):
    pp = par.ParabolicProblem([0, 1, 0, 1], 1, 0, 0, ("neumann", 0), heat_u0)
    o = dict(Scheme="bdf2", Dt=0.002, TFinal=0.1, BaseCells=32, MaxLevels=3, RegridEvery=5, StepVerbose=0)
    o.update(kw)
    return pp, par.parabolic_options(**o)


def heat_config(**kw):
    o = dict(BaseCells=64, MaxLevels=4, RelTol=1e-12)
    o.update(kw)
    return heat_readme(**o)


def mms_bump(**kw):
    sg, vx, vy = 0.06, 0.4, 0.3

    def xc(t):
        return 0.30 + vx * t

    def yc(t):
        return 0.35 + vy * t

    def uex(x, y, t):
        return np.exp(-((x - xc(t)) ** 2 + (y - yc(t)) ** 2) / (2 * sg ** 2))

    def ff(x, y, t):
        u = uex(x, y, t)
        ut = u * ((x - xc(t)) * vx + (y - yc(t)) * vy) / sg ** 2
        lap = u * (((x - xc(t)) ** 2 + (y - yc(t)) ** 2) / sg ** 4 - 2 / sg ** 2)
        return ut - lap + u

    pp = par.ParabolicProblem([0, 1, 0, 1], 1, 1, ff, ("dirichlet", uex), lambda x, y: uex(x, y, 0), uex)
    thr = 5e-2
    o = dict(Scheme="bdf2", Dt=0.05, TFinal=0.4, RegridEvery=2, StepVerbose=0, BaseCells=32, MaxLevels=2,
             Indicator="ulap", TagThreshold=lambda L: thr / 4 ** (L - 1), TagBuffer=3, RelTol=1e-9)
    o.update(kw)
    return pp, par.parabolic_options(**o)


def reaction_t(**kw):
    pp = par.ParabolicProblem([0, 1, 0, 1], 1, lambda x, y, t: 1 + t + 0 * x, lambda x, y, t: t * np.exp(-t) + 0 * x,
                              ("neumann", 0), lambda x, y: np.ones(np.shape(x)),
                              lambda x, y, t: np.exp(-t) * np.ones(np.shape(x)))
    o = dict(Scheme="bdf2", Dt=0.05, TFinal=0.4, BaseCells=8, MaxLevels=1, CoarsestCells=4, StepVerbose=0,
             RegridEvery=0, RelTol=1e-12)
    o.update(kw)
    return pp, par.parabolic_options(**o)


def diffusion_t(**kw):
    def u(x, y, t):
        return np.exp(-t) * (1 + x ** 2 + y ** 2)

    def f(x, y, t):
        return -u(x, y, t) - 4 * (1 + t) * np.exp(-t)

    pp = par.ParabolicProblem([0, 1, 0, 1], lambda x, y, t: (1 + t) + 0 * x, 0, f, ("dirichlet", u),
                              lambda x, y: u(x, y, 0), u)
    o = dict(Scheme="bdf2", Dt=0.05, TFinal=0.4, BaseCells=16, MaxLevels=1, CoarsestCells=4, StepVerbose=0,
             RegridEvery=0, RelTol=1e-12)
    o.update(kw)
    return pp, par.parabolic_options(**o)


def fixed_patch(**kw):
    pp = par.ParabolicProblem([0, 1, 0, 1], lambda x, y: 1 + 0.2 * x, 0, 0, ("neumann", 0),
                              lambda x, y: 1 + 0.2 * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y))
    o = dict(Scheme="bdf2", Dt=0.002, TFinal=0.02, BaseCells=16, MaxLevels=2, CoarsestCells=4,
             Indicator="fun", TagFunction=lambda x, y: (x > 0.25) & (x < 0.625) & (y > 0.25) & (y < 0.75),
             TagBuffer=1, StepVerbose=0, RegridEvery=2, RelTol=1e-11)
    o.update(kw)
    return pp, par.parabolic_options(**o)


def reaction_t_cn(**kw):
    return reaction_t(**dict(dict(Scheme="cn"), **kw))


def diffusion_t_cn(**kw):
    return diffusion_t(**dict(dict(Scheme="cn"), **kw))


def fixed_patch_cn(**kw):
    return fixed_patch(**dict(dict(Scheme="cn"), **kw))


def mms_bump_64(**kw):
    thr = 5e-2 / 4
    o = dict(Dt=0.025, BaseCells=64, TagThreshold=lambda L: thr / 4 ** (L - 1))
    o.update(kw)
    return mms_bump(**o)


CASES = {"heat_readme": heat_readme, "mms_bump": mms_bump, "mms_bump_64": mms_bump_64, "reaction_t": reaction_t,
         "diffusion_t": diffusion_t, "fixed_patch": fixed_patch, "reaction_t_cn": reaction_t_cn,
         "diffusion_t_cn": diffusion_t_cn, "fixed_patch_cn": fixed_patch_cn}
