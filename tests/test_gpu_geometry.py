"""Diffuse-domain starfish geometry on the device (bsam_polygon_fields)
against the host helper (DiffuseDomainBSAM/src/ddStarfishGeometry.m:35-72,
ddTransmissionProblem.m:14-15).

Stated tolerance: the distance magnitude is bit-identical (the same
point-to-segment expressions, minimum over every segment); the sign agrees
wherever |d| > 1e-12 (it comes from atan2/cos/sin, whose device and host
library results may differ in the last bit); phi within 1e-15 absolute and
delta within 1e-13 relative (tanh and cosh of the device math library).  The
host helper itself is pinned against brute force and the polygon interior in
the CPU suite (testStarfish.m:5-16).
"""
import numpy as np
import pytest

from bsamgpu import configs, geometry, solve

pytestmark = pytest.mark.gpu


def _points():
    rng = np.random.default_rng(1127)
    pts = [rng.uniform(-2, 2, (300, 2))]
    h = 4.0 / 64
    c = -2 + (np.arange(64) + 0.5) * h
    X, Y = np.meshgrid(c, c, indexing="ij")
    pts.append(np.stack([X.ravel(), Y.ravel()], axis=1))
    g = geometry.starfish_geometry()
    v = g.vertices[:-1]
    pts.append(v * (1 + 1e-3 * rng.standard_normal((v.shape[0], 1))))  # close to the boundary
    pts.append(np.array([[0.0, 0.0], [1e-300, 0.0], [0.0, -1.5], [2.0, 2.0]]))
    return np.concatenate(pts)


def test_device_distance_and_fields_match_host():
    P = _points()
    host = geometry.starfish_geometry()
    dev = geometry.starfish_geometry(device=0)
    eps = 0.05
    dh, ph, eh = host.fields(P[:, 0], P[:, 1], eps)
    dd, pd, ed = dev.fields(P[:, 0], P[:, 1], eps)
    print(f"{P.shape[0]} points; max |d| {np.abs(dh).max():.3f}", flush=True)
    assert np.array_equal(np.abs(dd), np.abs(dh))
    far = np.abs(dh) > 1e-12
    assert np.array_equal(np.sign(dd[far]), np.sign(dh[far]))
    assert np.max(np.abs(pd - ph)) <= 1e-15
    assert np.all(np.abs(ed - eh) <= 1e-13 * np.abs(eh) + 1e-300)
    assert np.array_equal(dev.signed_distance(P[:, 0], P[:, 1]), dd)


def test_starfish_built_with_device_geometry():
    """The same hierarchy (tags come from |d|), coefficients within the
    stated tolerance (1e-12 relative to max(|value|, 1): F carries g * delta
    with delta up to 1/(2 eps)), and the solve of the host-built hierarchy
    (10 cycles, final residual within 10%)."""
    import time

    t = time.perf_counter()
    prob, _, g = configs.starfish_problem()
    Hh = solve.build_hierarchy(prob, configs.starfish_options(g))
    th = time.perf_counter() - t
    t = time.perf_counter()
    prob_d, _, g_d = configs.starfish_problem(device=0)
    Hd = solve.build_hierarchy(prob_d, configs.starfish_options(g_d))
    td = time.perf_counter() - t
    print(f"starfish hierarchy build: host geometry {th:.2f} s, device geometry {td:.2f} s", flush=True)
    assert Hd.nlev == Hh.nlev
    worst = 0.0
    for k in range(1, Hh.nlev + 1):
        assert np.array_equal(np.array(Hd.L(k).box), np.array(Hh.L(k).box)), k
        for p in range(Hh.L(k).np):
            for name in ("DeW", "DnS", "CC", "FTrue"):
                a = getattr(Hd.L(k), name)[p]
                b = getattr(Hh.L(k), name)[p]
                worst = max(worst, float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0))))
    print(f"max coefficient difference (relative to max(|value|, 1)): {worst:.2e}", flush=True)
    assert worst < 1e-12
    from bsamgpu.device import GpuHierarchy

    out = []
    for H in (Hh, Hd):
        G = GpuHierarchy(H)
        out.append(G.solve_cycles())
        G.close()
    sh, sd = out
    print(f"starfish solve: host geometry {sh['vcycles']} cycles rel {sh['finalRel']:.4e}; device geometry "
          f"{sd['vcycles']} cycles rel {sd['finalRel']:.4e}", flush=True)
    # the last cycle's residual sits a few times above FP64 round-off, so
    # round-off-level coefficient changes move it by a few percent
    assert sd["vcycles"] == sh["vcycles"] == 10
    assert sd["finalRel"] <= 1e-10 and abs(sd["finalRel"] - sh["finalRel"]) <= 0.1 * sh["finalRel"]


def test_distance_throughput():
    """Device kernel against the host helper on the cell centres of a 1024^2
    grid over the starfish box (timings printed, not asserted)."""
    import time

    n = 1024
    c = -2 + (np.arange(n) + 0.5) * (4.0 / n)
    X, Y = np.meshgrid(c, c, indexing="ij")
    host = geometry.starfish_geometry()
    dev = geometry.starfish_geometry(device=0)
    dev.fields(X[:4], Y[:4], 0.05)  # initialise the device
    t = time.perf_counter()
    dd = dev.fields(X, Y, 0.05)[0]
    td = time.perf_counter() - t
    t = time.perf_counter()
    dh = host.fields(X, Y, 0.05)[0]
    th = time.perf_counter() - t
    print(f"signed distance + fields on {n * n} points: device {td * 1e3:.1f} ms (incl. transfers), "
          f"host helper {th * 1e3:.1f} ms", flush=True)
    assert np.array_equal(np.abs(dd), np.abs(dh))


def test_polygon_fields_errors():
    from bsamgpu import _lib

    lib = _lib.load()
    v = np.zeros(3)
    P = _lib.P_double
    null = P()
    x = np.zeros(2)
    d = np.zeros(2)
    dp = lambda a: a.ctypes.data_as(P)  # noqa: E731
    assert lib.bsam_polygon_fields(0, dp(v), dp(v), 3, dp(x), dp(x), 2, 0.1, dp(d), null, null) == -1
    g = geometry.starfish_geometry()
    vx = np.ascontiguousarray(g.vertices[:, 0])
    vy = np.ascontiguousarray(g.vertices[:, 1])
    assert lib.bsam_polygon_fields(0, dp(vx), dp(vy), vx.size, dp(x), dp(x), 2, 0.0, dp(d), dp(d), null) == -1
    assert lib.bsam_polygon_fields(99, dp(vx), dp(vy), vx.size, dp(x), dp(x), 2, 0.1, dp(d), null, null) == -1
    assert lib.bsam_polygon_fields(0, dp(vx), dp(vy), vx.size, dp(x), dp(x), 2, 0.1, dp(d), null, null) == 0
