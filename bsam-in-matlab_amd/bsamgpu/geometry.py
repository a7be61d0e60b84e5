"""Diffuse-domain starfish geometry and transmission coefficients (host side).

* ``starfish_geometry``       <- DiffuseDomainBSAM/src/ddStarfishGeometry.m:1-73
* ``transmission_problem``    <- DiffuseDomainBSAM/src/ddTransmissionProblem.m:1-30

The signed distance is computed by the compiled helper ``libbsamgeom.so``
(bsam_geom.c): exact point-to-segment distance with bounding-box pruning.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .problem import Problem, BsamError

_HERE = os.path.dirname(os.path.abspath(__file__))
_GEOM_SO = os.path.join(_HERE, "..", "csrc", "libbsamgeom.so")
_geom = None


def _lib():
    global _geom
    if _geom is None:
        if not os.path.exists(_GEOM_SO):
            raise BsamError("bsamgpu:geom", f"{_GEOM_SO} not built (run __graft_entry__.build())")
        lib = ctypes.CDLL(_GEOM_SO)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.bsam_polygon_signed_distance.argtypes = [dp, dp, ctypes.c_int, dp, dp, ctypes.c_long, dp]
        lib.bsam_polygon_signed_distance.restype = ctypes.c_int
        _geom = lib
    return _geom


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class PolygonGeometry:
    """A closed, radially ordered polygon.  ``device``: a HIP device index;
    then distances and the diffuse-domain fields are evaluated by the device
    library (bsam_polygon_fields), else by the host helper."""

    def __init__(self, vertices, name="", device=None):
        self.vertices = np.ascontiguousarray(vertices, dtype=float)
        self.name = name
        self.device = device
        self._vx = np.ascontiguousarray(self.vertices[:, 0])
        self._vy = np.ascontiguousarray(self.vertices[:, 1])

    @staticmethod
    def _flat(X, Y):
        X = np.asarray(X, dtype=float)
        Y = np.asarray(Y, dtype=float)
        shape = np.broadcast(X, Y).shape
        x = np.ascontiguousarray(np.broadcast_to(X, shape).reshape(-1, order="F"))
        y = np.ascontiguousarray(np.broadcast_to(Y, shape).reshape(-1, order="F"))
        return x, y, shape

    def fields(self, X, Y, epsilon=None):
        """(d, phi, delta) at the points; phi and delta are None unless
        ``epsilon`` is given (ddTransmissionProblem.m:14-15)."""
        x, y, shape = self._flat(X, Y)
        d = np.empty(x.size)
        phi = delta = None
        if self.device is None:
            rc = _lib().bsam_polygon_signed_distance(_ptr(self._vx), _ptr(self._vy), self._vx.size,
                                                     _ptr(x), _ptr(y), x.size, _ptr(d))
            if rc != 0:
                raise BsamError("bsamgpu:geom", f"signed distance failed ({rc})")
            if epsilon is not None:
                phi = (1 + np.tanh(d / epsilon)) / 2
                delta = (1.0 / np.cosh(d / epsilon)) ** 2 / (2 * epsilon)
        else:
            from . import _lib as dev

            if epsilon is not None:
                phi = np.empty(x.size)
                delta = np.empty(x.size)
            null = ctypes.POINTER(ctypes.c_double)()
            dev.check(dev.load().bsam_polygon_fields(
                int(self.device), _ptr(self._vx), _ptr(self._vy), self._vx.size, _ptr(x), _ptr(y), x.size,
                float(epsilon or 0.0), _ptr(d), _ptr(phi) if phi is not None else null,
                _ptr(delta) if delta is not None else null), "bsam_polygon_fields")
        out = [d.reshape(shape, order="F")]
        out += [a.reshape(shape, order="F") if a is not None else None for a in (phi, delta)]
        return tuple(out)

    def signed_distance(self, X, Y):
        return self.fields(X, Y)[0]

    def inside(self, X, Y):
        return self.signed_distance(X, Y) >= 0


def starfish_geometry(n=3000, device=None):
    """r(theta) = 0.9 (1.2 + 0.7 sin 5 theta), n points incl. the repeated
    endpoint (ddStarfishGeometry.m:8-21); ``device`` as for PolygonGeometry."""
    if n < 32:
        raise BsamError("dd_starfish:n", "n must be an integer >= 32")
    theta = np.linspace(0.0, 2.0 * np.pi, n)
    r = 0.9 * (1.2 + 0.7 * np.sin(5 * theta))
    v = np.stack([r * np.cos(theta), r * np.sin(theta)], axis=1)
    v[-1, :] = v[0, :]
    return PolygonGeometry(v, "Helmholtz starfish: r = 0.9(1.2 + 0.7 sin(5 theta))", device)


def _tofun(v):
    if callable(v):
        return v
    return lambda X, Y, _v=float(v): _v + np.zeros(np.shape(X))


def transmission_problem(geom, epsilon, box, data):
    """ddTransmissionProblem.m:1-30:
    -div(D grad u) + (c + kappa delta) u = h(1-phi) + q phi - g delta,
    phi = (1+tanh(d/eps))/2, delta = sech(d/eps)^2/(2 eps)."""
    if not (epsilon > 0):
        raise BsamError("dd_transmission:epsilon", "epsilon must be positive")
    if not (data["alpha"] > 0):
        raise BsamError("dd_transmission:alpha", "alpha must be positive")
    for f in ("beta", "gamma", "kappa"):
        if data[f] < 0:
            raise BsamError("dd_transmission:" + f, f + " must be nonnegative")
    sd = geom.signed_distance
    fields = getattr(geom, "fields", None)

    def phi(X, Y):
        if fields is not None:
            return fields(X, Y, epsilon)[1]
        return (1 + np.tanh(sd(X, Y) / epsilon)) / 2

    def delta(X, Y):
        if fields is not None:
            return fields(X, Y, epsilon)[2]
        return (1.0 / np.cosh(sd(X, Y) / epsilon)) ** 2 / (2 * epsilon)

    q = _tofun(data["q"])
    h = _tofun(data["h"])
    g = _tofun(data["g"])
    al, be, ga, ka = data["alpha"], data["beta"], data["gamma"], data["kappa"]

    def D(X, Y):
        return al + (1 - al) * phi(X, Y)

    def C(X, Y):
        return be + (ga - be) * phi(X, Y) + ka * delta(X, Y)

    def F(X, Y):
        return h(X, Y) * (1 - phi(X, Y)) + q(X, Y) * phi(X, Y) - g(X, Y) * delta(X, Y)

    bc = data.get("OuterBC", ("neumann", 0))
    prob = Problem(box, D, C, F, bc, data.get("exact"))
    aux = {"phi": phi, "delta": delta, "epsilon": epsilon, "signedDistance": sd,
           "vertices": geom.vertices}
    return prob, aux
