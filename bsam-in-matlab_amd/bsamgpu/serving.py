"""Serving many independent elliptic solves on one GPU with continuous batching.

Each request is one ebSolve (EllipticBSAM/src/ebSolve.m:22-77): solve stages
separated by host tagging and level construction, or one ebpSolve
(ParabolicBSAM/src/ebpSolve.m:17-98): one stage per time step, with the BDF2
history resident on the device between steps.  The device advances every
request that is in a solve stage with one launch per tick
(``bsam_batch_step``: one workgroup per hierarchy, its stage preamble or one
whole cycle), so requests that arrive while others are on the GPU join the
next tick, and a request whose stage ends leaves the batch for host work and
comes back with its next stage.  Host work (hierarchy setup, uploads,
tagging, level construction) runs on a thread pool while the device ticks;
the foreign calls release the interpreter lock.

Admission: at most ``slots`` requests in flight and an estimated device
footprint of at most ``mem_budget`` bytes; arrivals wait in FIFO order.

A request's result does not depend on what else is in the batch: each
workgroup runs exactly the arithmetic of the hierarchy's own solve.
"""
from __future__ import annotations

import ctypes
import threading
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib, diagnostics, hierarchy, parabolic, solve, tagcluster
from ._lib import check
from .device import GpuHierarchy, context, solve_opts

REALMIN = np.finfo(float).tiny


class Batch:
    """A ``bsam_batch``: one launch advances up to ``capacity`` hierarchies."""

    def __init__(self, capacity, device=0):
        self.lib = _lib.load()
        self.capacity = int(capacity)
        self.b = ctypes.c_void_p()
        check(self.lib.bsam_batch_create(context(device), self.capacity, ctypes.byref(self.b)), "bsam_batch_create")
        self._hs = (ctypes.c_void_p * self.capacity)()
        self._ph = (ctypes.c_int * self.capacity)()
        self._out = (_lib.Residual * self.capacity)()

    def step(self, items, phases, so):
        n = len(items)
        for i, (G, ph) in enumerate(zip(items, phases)):
            self._hs[i] = G.h
            self._ph[i] = int(ph)
        check(self.lib.bsam_batch_step(self.b, self._hs, self._ph, n, ctypes.byref(so), self._out), "bsam_batch_step")
        return [{"linf": r.linf, "l2": r.l2, "fnorm": r.fnorm, "rel": r.rel, "ncell": r.ncell}
                for r in self._out[:n]]

    def close(self):
        if getattr(self, "b", None):
            self.lib.bsam_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_bytes(H):
    """Estimated device footprint of a hierarchy (fields, tables, coarsest inverse)."""
    b = 0
    for k in range(1, H.nlev + 1):
        L = H.L(k)
        for p in range(L.np):
            mx, my = L.dims(p)
            cells = mx * my
            b += 8 * ((mx + 2) * (my + 2) + 6 * cells + (mx + 1) * my + mx * (my + 1)
                      + (mx // 2 + 2) * (my // 2 + 2)) + 64 * (cells + 2 * (mx + my))
    n1 = H.L(1).n[0] * H.L(1).n[1] + 1
    return b + 8 * n1 * n1


class Request:
    """State of one served ebSolve."""

    def __init__(self, rid, prob, opts, arrival=0.0):
        self.id = rid
        self.prob = prob
        self.opts = opts
        self.arrival = arrival
        self.H = None
        self.G = None
        self.bytes = 0
        self.phase = 0
        self.it = 0
        self.hist = []
        self.rel = float("nan")
        self.ncell = 0
        self.stages = []
        self.t_admit = None
        self.t_done = None
        self.Q = None
        self.error = None
        self.charged = 0

    # ---- stage bookkeeping (ebSolveCycles.m:15-41 stopping rule)
    def take(self, res):
        """Apply one batch result; True when the stage has ended."""
        if self.phase == 0:
            self.hist = [res["linf"]]
            self.it = 0
        else:
            self.it += 1
            self.hist.append(res["linf"])
        self.rel = res["rel"]
        self.ncell = res["ncell"]
        more = self.it < self.opts.MaxVCycles and self.rel > self.opts.RelTol and np.isfinite(self.rel)
        self.phase = 1
        return not more

    def stage_record(self):
        H = self.H
        hist = self.hist
        tot = sum(int(np.prod(H.L(k).dims(p))) for k in range(H.rootIdx, H.nlev + 1) for p in range(H.L(k).np))
        return {"amrLevels": H.nlev - H.rootIdx + 1, "nlev": H.nlev, "vcycles": self.it, "resHist": list(hist),
                "finalRel": self.rel, "visibleCells": self.ncell, "totalCells": tot,
                "avgFactor": (hist[-1] / max(hist[0], REALMIN)) ** (1.0 / (len(hist) - 1))
                if len(hist) >= 2 else float("nan")}


    # ---- host steps (run on the server's thread pool)
    def prepare(self):
        """ebSolve.m:22-36: hierarchy setup and upload; the first stage follows."""
        self.H = hierarchy.setup(self.prob, self.opts)
        self.G = GpuHierarchy(self.H)
        if self.opts.Indicator.lower() == "fun":  # geometric tags: every level before the stage
            while self.H.nlev - self.H.rootIdx + 1 < self.opts.MaxLevels:
                boxes = tagcluster.tag_cluster(self.H)
                if not boxes:
                    break
                self.G.pull_q(self.H)
                self.G.close()
                kNew = hierarchy.add_level(self.H, boxes)
                self.G = GpuHierarchy(self.H)
                self.G.init_level_prolong(kNew)
        self.bytes = device_bytes(self.H)
        self.phase = 0
        return True

    def after_stage(self):
        """ebSolve.m:45-53: tag on the stage's solution and add a level, or
        finish.  Returns True when another stage follows."""
        self.stages.append(self.stage_record())
        H, opts = self.H, self.opts
        self.G.pull_q(H)
        if opts.Indicator.lower() != "fun" and H.nlev - H.rootIdx + 1 < opts.MaxLevels:
            boxes = tagcluster.tag_cluster(H, device=self.G)
            if boxes:
                self.G.close()
                kNew = hierarchy.add_level(H, boxes)
                self.G = GpuHierarchy(H)
                self.G.init_level_prolong(kNew)
                self.bytes = device_bytes(H)
                self.phase = 0
                return True
        mz = opts.MeanZero  # ebSolve.m:56-68
        if isinstance(mz, str) and mz.lower() == "auto":
            mz = H.singular
        if mz:
            m, _ = diagnostics.composite_mean(H)
            for k in range(1, H.nlev + 1):
                for p in range(H.L(k).np):
                    H.L(k).Q[p] = H.L(k).Q[p] - m
        self.Q = [[q.copy() for q in H.L(k).Q] for k in range(1, H.nlev + 1)]
        self.G.close()
        self.G = None
        return False


class ParabolicRequest(Request):
    """One served ebpSolve (ParabolicBSAM/src/ebpSolve.m:17-98): every time
    step is one stage of the batch (rhs from the resident history, stage
    preamble and cycles); history rotation, composite mass, coefficient
    refresh and regrids run as host steps between stages."""

    def __init__(self, rid, pp, popts, arrival=0.0):
        super().__init__(rid, None, popts.eb, arrival)
        self.pp = pp
        self.popts = popts
        self.S = None
        self.n = 0
        self.mass = []
        self.nregrids = 0

    def _begin_step(self):
        """ebpSolve.m:41-49: source f(t1), implicit shift, right-hand side."""
        self.n += 1
        t1 = self.n * self.dt
        scheme = self.scheme1 if self.n == 1 else self.popts.Scheme
        if scheme == "cn":  # explicit half with the old operator first
            parabolic.cn_rhs(self.S, self.pp, t1, self.dt)
            parabolic.set_time(self.S, self.pp, t1, parabolic.ALPHA[scheme] / self.dt)
        else:
            parabolic._source(self.S, self.pp, t1)
            parabolic.set_time(self.S, self.pp, t1, parabolic.ALPHA[scheme] / self.dt)
            self.S.G.parabolic_rhs(self.dt, scheme)
        self.G = self.S.G
        self.H = self.S.H
        self.phase = 0

    def prepare(self):
        po = self.popts
        self.nsteps = max(1, parabolic._matlab_round(po.TFinal / po.Dt))
        self.dt = po.TFinal / self.nsteps
        self.scheme1 = "be" if po.Scheme == "bdf2" else po.Scheme
        self.S = parabolic.ebp_init(self.pp, po, parabolic.ALPHA[self.scheme1] / self.dt)
        m, a = self.S.G.composite_mean()
        self.mass = [m * a]
        self.bytes = device_bytes(self.S.H)
        self._begin_step()
        return True

    def after_stage(self):
        """ebpSolve.m:55-77: history, mass, regrid, next step or finish."""
        st = self.stage_record()
        self.stages.append(st)
        G = self.S.G
        G.parabolic_advance()
        m, a = G.composite_mean()
        self.mass.append(m * a)
        po = self.popts
        if po.RegridEvery > 0 and self.n % po.RegridEvery == 0 and self.n < self.nsteps and po.eb.MaxLevels > 1:
            self.S = parabolic.ebp_regrid(self.S, self.pp, self.n * self.dt, parabolic.ALPHA[po.Scheme] / self.dt, po)
            self.nregrids += 1
            self.bytes = device_bytes(self.S.H)
        if self.n < self.nsteps:
            self._begin_step()
            return True
        self.S.G.pull_q(self.S.H)
        self.H = self.S.H
        self.Q = [[q.copy() for q in self.H.L(k).Q] for k in range(1, self.H.nlev + 1)]
        self.S.G.close()
        self.G = None
        return False


class Server:
    """Continuous-batching request loop on one GPU.

    ``opts`` supplies the smoothing parameters shared by every request of a
    batch (NPre, NPost, Omega, MassCorrection); RelTol, MaxVCycles and the
    refinement options are per request."""

    def __init__(self, slots=256, mem_budget=16 << 30, precision="fp64", host_workers=8, device=0, opts=None):
        from .problem import Options

        self.slots = int(slots)
        self.mem_budget = int(mem_budget)
        self.precision = precision
        self.batch = Batch(self.slots, device)
        self.pool = ThreadPoolExecutor(max_workers=host_workers)
        self.so = solve_opts(opts or Options(), precision=precision)
        self.ticks = 0
        self.tick_items = 0
        self.peak_bytes = 0

    def close(self):
        self.pool.shutdown(wait=True)
        self.batch.close()

    def run(self, requests, realtime=True):
        """Serve ``requests`` (each with .arrival in seconds from the start;
        with realtime=False every request is available at once).  Returns
        the requests in completion order with t_admit, t_done (seconds from
        the start), stages and the result Q set."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        total = len(pending)
        lock = threading.Lock()
        ready = deque()
        finished = deque()
        wake = threading.Event()
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        def host_done(fut):
            req = fut.req
            try:
                more = fut.result()
            except Exception as e:  # the request ends with its host-side error
                req.error = e
                more = False
            with lock:
                if more:
                    ready.append(req)
                else:
                    req.t_done = now()
                    finished.append(req)
            wake.set()

        def submit(fn, req):
            fut = self.pool.submit(fn, req)
            fut.req = req
            fut.add_done_callback(host_done)

        def prep(req):
            return req.prepare()

        inflight = 0
        used = 0
        done = []
        active = []
        while len(done) < total:
            with lock:
                back = list(ready)
                ready.clear()
                out = list(finished)
                finished.clear()
            for r in back:  # returned from host work: charge its actual footprint
                used += r.bytes - r.charged
                r.charged = r.bytes
                active.append(r)
            for r in out:
                inflight -= 1
                used -= r.charged
                done.append(r)
            self.peak_bytes = max(self.peak_bytes, used)
            # admission: FIFO under the slot count and the memory budget
            while pending and (not realtime or pending[0].arrival <= now()) and inflight < self.slots:
                req = pending[0]
                est = device_bytes_estimate(req)
                if used + est > self.mem_budget and inflight > 0:
                    break
                pending.popleft()
                req.t_admit = now()
                req.charged = est
                inflight += 1
                used += est
                submit(prep, req)
            if not active:
                timeout = 0.05
                if pending and realtime and inflight < self.slots:
                    timeout = min(timeout, max(0.0, pending[0].arrival - now()))
                wake.wait(timeout)
                wake.clear()
                continue
            res = self.batch.step([r.G for r in active], [r.phase for r in active], self.so)
            self.ticks += 1
            self.tick_items += len(active)
            keep = []
            for r, rr in zip(active, res):
                if r.take(rr):
                    submit(lambda q: q.after_stage(), r)
                else:
                    keep.append(r)
            active = keep
        return done


def device_bytes_estimate(req):
    """Footprint of a request's first stage from its options (ladder + root)."""
    n0 = np.array(req.opts.BaseCells)
    sizes = hierarchy.ladder_sizes(req.opts)
    cells = sum(int(s[0] * s[1]) for s in sizes)
    return 8 * 12 * cells + 64 * 4 * int(n0.sum()) + 8 * 257 * 257


def serial_solve(req, precision="fp64"):
    """One-at-a-time form: ebSolve of this request alone with the plain
    multi-launch stage solves; returns (stages, Q)."""
    sol = solve.eb_solve(req.prob, req.opts.replace(Precision=precision))
    H = sol["H"]
    sol["device"].close()
    return list(H.stats["stages"]), [[q.copy() for q in H.L(k).Q] for k in range(1, H.nlev + 1)]


def latency_stats(reqs):
    lat = np.array([r.t_done - r.arrival for r in reqs])
    return {"n": len(reqs), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3)}
