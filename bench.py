#!/usr/bin/env python3
"""Benchmark: starfish 64 -> 1024 composite solve to 1e-10 on one MI355X.

Workload (the diffuse-domain starfish transmission problem, 5 AMR levels,
64^2 root refined to 1024^2 spacing near the interface, 7 multigrid levels,
544,664 stored cells, 408,562 visible): one step = one composite solve from
the zero initial guess to relative composite residual <= 1e-10 (ebSolveCycles:
fill, baseline residual, V-cycles + filldown + residual per cycle) with the
hierarchy resident in HBM.

Reported: solves/s (value), seconds per V-cycle, p50/p99 solve latency, the
smoother's achieved bandwidth against 8 TB/s (roofline of the dominant kernel,
the finest-level smoother, timed with HIP events on the library stream), the
FP32-inner / FP64-defect-correction form with its error against the FP64
solution, and the C++ restatement of the solve stage timed on this host (one
thread and every core) as the baseline.

  python bench.py [--steps K] [--warmup W] [--no-cpu-baseline]
  python bench.py --kernel-only      # only the timed smoother launches (for rocprofv3)
"""
import argparse
import json
import os
import statistics
import sys

import numpy as np
import time
import warnings

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bsam-in-matlab_amd"))
sys.path.insert(0, ROOT)
warnings.filterwarnings("ignore")

HBM_PEAK_GBS = 8000.0


def pmc_traffic(kernel_tag):
    """HBM-side bytes per launch of the profiled kernel, as measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (copy-calibrated) recorded in
    profiles/; None when no such record exists."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{kernel_tag}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("TRAFFIC_BYTES_PER_LAUNCH="):
                return int(line.split("=", 1)[1])
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_starfish():
    from bsamgpu import configs, solve

    prob, aux, geom = configs.starfish_problem()
    opts = configs.starfish_options(geom)
    return solve.build_hierarchy(prob, opts), opts


def time_solves(G, steps, warmup, precision, **kw):
    for _ in range(warmup):
        G.reset()
        G.solve_cycles(precision=precision, **kw)
    lat = []
    stages = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        G.reset()
        st = G.solve_cycles(precision=precision, **kw)
        lat.append((time.perf_counter() - t) * 1e3)
        stages.append(st)
    total = time.perf_counter() - t0
    return total, lat, stages


def pct(xs, q):
    xs = sorted(xs)
    i = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[i]


def visible_rel_l2(H, Qa, Qb):
    num = 0.0
    den = 0.0
    for k in range(H.rootIdx, H.nlev + 1):
        L = H.L(k)
        da = L.h[0] * L.h[1]
        for p in range(L.np):
            vis = ~L.covered[p]
            a = Qa[k][p][1:-1, 1:-1][vis]
            b = Qb[k][p][1:-1, 1:-1][vis]
            num += float(((a - b) ** 2).sum()) * da
            den += float((b ** 2).sum()) * da
    return (num / den) ** 0.5


def cpu_baseline(H, cores, gpu_hist=None, seconds=6.0):
    """The C++ restatement of the solve stage (oracle/bsam_cpu.cpp, -O3,
    OpenMP over patches) timed on this host: whole starfish solves repeated
    for about ``seconds`` each at one thread and at ``cores`` threads.  Its
    residual history is bit-identical to the NumPy oracle's and is checked
    against the device's FP64 history."""
    from oracle import bsam_oracle as O
    from oracle import bsam_cpu

    c = bsam_cpu.CpuSolve(O.OracleHierarchy(H))
    out = {}
    for label, th in (("single", 1), ("all", cores)):
        n = 0
        t0 = time.perf_counter()
        while True:
            c.reset()
            hist, it, rel = c.solve(threads=th)
            n += 1
            dt = time.perf_counter() - t0
            if dt >= seconds:
                break
        out[label] = {"solves": n, "seconds_per_solve": dt / n, "vcycles": it, "final_rel": rel}
    same = None if gpu_hist is None else bool(np.array_equal(np.asarray(gpu_hist), hist))
    a, s1 = out["all"], out["single"]
    return {"value": round(1.0 / a["seconds_per_solve"], 3), "unit": "solves/s", "cores": cores,
            "kind": "port",
            "sample": f"{a['solves']} whole starfish solves ({a['vcycles']} V-cycles to rel {a['final_rel']:.2e}) "
                      f"by the C++ restatement of the solve stage on {cores} threads, "
                      f"{a['seconds_per_solve'] * 1e3:.1f} ms each; tables prebuilt",
            "seconds_per_solve": a["seconds_per_solve"],
            "seconds_per_vcycle": a["seconds_per_solve"] / max(a["vcycles"], 1),
            "single_thread": {"value": round(1.0 / s1["seconds_per_solve"], 3), "cores": 1,
                              "solves": s1["solves"], "seconds_per_solve": s1["seconds_per_solve"]},
            "residual_history_equals_device_fp64": same}


def bench_large_smoother(n, launches=40, solves=True):
    """The smoother on one uniform n x n level (D = C = 1), large enough to be
    HBM-bound rather than launch-bound: the colour pass (half a sweep per
    launch) and the streaming sweep (a whole red+black sweep per launch),
    which the solve uses on such levels."""
    from bsamgpu import Options, Problem, _lib, hierarchy
    from bsamgpu.device import GpuHierarchy

    prob = Problem([0, 1, 0, 1], 1, 1, 1, ("dirichlet", 0))
    # RelTol 1e-8: FP64 round-off bounds the relative residual near 1e-9 at
    # this spacing (|L| ~ 8/h^2 = 1.3e8 times |u| ~ 0.07 times 2^-52)
    H = hierarchy.setup(prob, Options(BaseCells=n, MaxLevels=1, Verbose=0, RelTol=1e-8))
    G = GpuHierarchy(H)
    k = H.nlev
    out = {"cells": n * n}
    for form in ("passes", "exact"):
        G.solve_cycles(max_vcycles=0, smoother=form)  # selects the timed form
        for name, mixed, b in (("fp64", False, 8), ("fp32", True, 4)):
            G.time_smoother(k, 5, mixed=mixed)
            ms = G.time_smoother(k, launches, mixed=mixed)
            byt = G.smoother_bytes(k, b)
            key = f"{name}_colour_pass" if form == "passes" else f"{name}_streaming_sweep"
            out[key] = {"us_per_launch": round(ms * 1e3, 2), "bytes_per_launch": byt,
                        "GBps": round(byt / ms / 1e6, 1), "frac": round(byt / ms / 1e6 / HBM_PEAK_GBS, 4)}
    # the streaming sweep at each row-block height (0: the automatic choice)
    from bsamgpu import autotune

    rows = {}
    for r in (0,) + autotune.STREAM_ROWS:
        G.solve_cycles(max_vcycles=0, smoother="exact", stream_rows=r)
        rows[str(r)] = {}
        for name, mixed, b in (("fp64", False, 8), ("fp32", True, 4)):
            G.time_smoother(k, 5, mixed=mixed)
            ms = G.time_smoother(k, launches, mixed=mixed)
            rows[str(r)][name] = {"us_per_launch": round(ms * 1e3, 2),
                                  "frac": round(G.smoother_bytes(k, b) / ms / 1e6 / HBM_PEAK_GBS, 4)}
    out["streaming_sweep_by_rows"] = rows
    if not solves:
        G.close()
        return out
    # whole solves from zero to RelTol, the option search over precision,
    # FP32 cycles per FP64 residual and row-block height (coarse cycle off:
    # one level of 2^24 cells and its sub-root levels); the fastest mixed
    # form's error against the FP64 solution
    cands = [dict(coarse_cells=-1, smoother="exact", precision="fp64")] + [
        dict(coarse_cells=-1, smoother="exact", precision="mixed", inner_cycles=i) for i in (1, 2)]
    best, table = autotune.autotune(G, bit_exact_only=False, repeats=3, candidates=cands)
    r64 = min((r for r in table if r["precision"] == "fp64"), key=lambda r: r["seconds"])
    rmx = min((r for r in table if r["precision"] == "mixed" and r["finalRel"] <= G.opts.RelTol),
              key=lambda r: r["seconds"])
    G.reset()
    G.solve_cycles(precision="fp64", coarse_cells=-1)
    q64 = G.get_field(k, _lib.BSAM_Q)[0][1:-1, 1:-1].copy()
    G.reset()
    G.solve_cycles(**{x: rmx[x] for x in rmx if x not in ("seconds", "vcycles", "finalRel")})
    qmx = G.get_field(k, _lib.BSAM_Q)[0][1:-1, 1:-1]
    rel_l2 = float(np.linalg.norm(qmx - q64) / np.linalg.norm(q64))
    out["solve"] = {"rel_tol": G.opts.RelTol, "autotune_pick": best, "fp64_ms": round(r64["seconds"] * 1e3, 3),
                    "mixed_ms": round(rmx["seconds"] * 1e3, 3),
                    "mixed_inner_cycles": rmx.get("inner_cycles", 1), "mixed_stream_rows": rmx.get("stream_rows", 0),
                    "speedup_mixed_vs_fp64": round(r64["seconds"] / rmx["seconds"], 3),
                    "rel_l2_mixed_vs_fp64": rel_l2,
                    "table": [{k: (float(f"{v:.6g}") if isinstance(v, float) else v) for k, v in r.items()}
                              for r in table]}
    G.close()
    return out


def bench_uniform_solves(n, repeats=3):
    """Only the whole solves of the uniform n x n level (the problem of
    bench_large_smoother): FP64 and FP32-inner with two FP32 cycles per FP64
    residual, `repeats` of each after a warm-up (for kernel traces)."""
    from bsamgpu import Options, Problem, hierarchy
    from bsamgpu.device import GpuHierarchy

    prob = Problem([0, 1, 0, 1], 1, 1, 1, ("dirichlet", 0))
    H = hierarchy.setup(prob, Options(BaseCells=n, MaxLevels=1, Verbose=0, RelTol=1e-8))
    G = GpuHierarchy(H)
    out = {"cells": n * n}
    for name, kw in (("fp64", dict(precision="fp64")), ("mixed", dict(precision="mixed", inner_cycles=2))):
        G.reset()
        G.solve_cycles(coarse_cells=-1, **kw)
        ts = []
        for _ in range(repeats):
            G.reset()
            t = time.perf_counter()
            st = G.solve_cycles(coarse_cells=-1, **kw)
            ts.append(time.perf_counter() - t)
        out[name] = {"ms": round(min(ts) * 1e3, 3), "vcycles": st["vcycles"], "final_rel": st["finalRel"]}
    G.close()
    return out


def bench_parabolic():
    """Heat BDF2 (BaseCells 64, MaxLevels 4, Dt 0.002, T 0.1, RegridEvery 5,
    RelTol 1e-12 per step): wall time of ebp_solve with the history resident,
    V-cycles, composite mass drift; FP32-inner against FP64, and the same
    problem stepped by Crank-Nicolson."""
    from bsamgpu import parabolic as par

    def run(precision, scheme="bdf2"):
        pp = par.ParabolicProblem([0, 1, 0, 1], 1, 0, 0, ("neumann", 0),
                                  lambda x, y: 1 + 0.2 * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y))
        opts = par.parabolic_options(Scheme=scheme, Dt=0.002, TFinal=0.1, BaseCells=64, MaxLevels=4,
                                     RegridEvery=5, StepVerbose=0, RelTol=1e-12, Precision=precision)
        t = time.perf_counter()
        sol = par.ebp_solve(pp, opts)
        wall = time.perf_counter() - t
        m = np.array(sol["mass"])
        gpu_ms = sum(st["gpu_ms"] for st in sol["stages"])
        return sol, {"wall_s": round(wall, 4), "steps": len(sol["cycles"]), "vcycles": int(sum(sol["cycles"])),
                     "regrids": sol["nregrids"], "mass_drift_rel": float(np.max(np.abs(m - m[0])) / abs(m[0])),
                     "solve_gpu_ms": round(gpu_ms, 3), "steps_per_s": round(len(sol["cycles"]) / wall, 2)}

    run("fp64")  # warm-up (first-call library and graph setup)
    s64, r64 = run("fp64")
    smx, rmx = run("mixed")
    H64, Hmx = s64["H"], smx["H"]
    if H64.nlev == Hmx.nlev:
        num = den = 0.0
        for k in range(H64.rootIdx, H64.nlev + 1):
            for p in range(H64.L(k).np):
                a = Hmx.L(k).Q[p][1:-1, 1:-1]
                b = H64.L(k).Q[p][1:-1, 1:-1]
                num += float(((a - b) ** 2).sum())
                den += float((b ** 2).sum())
        rmx["rel_l2_vs_fp64"] = (num / den) ** 0.5
    _, rcn = run("fp64", "cn")
    return {"workload": "heat BDF2 64^2 root, MaxLevels 4, 50 steps, 9 regrids, RelTol 1e-12",
            "fp64": r64, "mixed": rmx, "cn_fp64": rcn}


def bench_varcoef_refined(steps=5, warmup=2):
    """Variable-coefficient configuration with the ulap threshold lowered to
    1e-6 so that the stages refine to the 2048^2 spacing (1.75M cells, a
    1024^2 single-patch level): a bandwidth-bound hierarchy.  Times the
    final stage's solve from zero (FP64 and FP32-inner), the smoother on its
    largest level, and reports the mixed form's error against FP64."""
    from bsamgpu import _lib, configs, solve
    from bsamgpu.device import GpuHierarchy

    with open(os.path.join(ROOT, "tests", "golden", "config2_params.json")) as f:
        par = json.load(f)
    t = time.perf_counter()
    sol = solve.eb_solve(configs.varcoef_problem(par), configs.varcoef_refined_options())
    build_s = time.perf_counter() - t
    sol["device"].close()
    H = sol["H"]
    for k in range(1, H.nlev + 1):  # the final stage from the zero initial guess
        for p in range(H.L(k).np):
            H.L(k).Q[p][:] = 0.0
    G = GpuHierarchy(H)
    tot64, lat64, st64 = time_solves(G, steps, warmup, "fp64")
    Q64 = {k: G.get_field(k, _lib.BSAM_Q) for k in range(1, H.nlev + 1)}
    totmx, latmx, stmx = time_solves(G, steps, warmup, "mixed")
    Qmx = {k: G.get_field(k, _lib.BSAM_Q) for k in range(1, H.nlev + 1)}
    cells = [sum(H.L(k).dims(p)[0] * H.L(k).dims(p)[1] for p in range(H.L(k).np)) for k in range(1, H.nlev + 1)]
    kb = max(range(1, H.nlev + 1), key=lambda k: cells[k - 1])
    G.time_smoother(kb, 10)
    ms64 = G.time_smoother(kb, 100)
    G.time_smoother(kb, 10, mixed=True)
    ms32 = G.time_smoother(kb, 100, mixed=True)
    b64, b32 = G.smoother_bytes(kb, 8), G.smoother_bytes(kb, 4)
    G.close()
    return {"workload": "variable D, BaseCells 256, MaxLevels 4, ulap threshold 1e-6 (stated change from 1e-3, "
                        "which tags nothing): final stage solved from zero to rel 1e-10",
            "levels": H.nlev, "patches_per_level": [H.L(k).np for k in range(1, H.nlev + 1)],
            "cells": int(sum(cells)), "build_s": round(build_s, 2),
            "fp64": {"ms_per_solve": round(tot64 / steps * 1e3, 3), "vcycles": st64[-1]["vcycles"],
                     "final_rel": st64[-1]["finalRel"], "p50_ms": round(pct(lat64, 50), 3),
                     "p99_ms": round(pct(lat64, 99), 3)},
            "mixed": {"ms_per_solve": round(totmx / steps * 1e3, 3), "vcycles": stmx[-1]["vcycles"],
                      "final_rel": stmx[-1]["finalRel"], "speedup_vs_fp64": round(tot64 / totmx, 3),
                      "rel_l2_vs_fp64": visible_rel_l2(H, Qmx, Q64)},
            "smoother_largest_level": {"level": kb, "cells": int(cells[kb - 1]),
                                       "fp64": {"us_per_launch": round(ms64 * 1e3, 2),
                                                "GBps": round(b64 / ms64 / 1e6, 1),
                                                "frac": round(b64 / ms64 / 1e6 / HBM_PEAK_GBS, 4)},
                                       "fp32": {"us_per_launch": round(ms32 * 1e3, 2),
                                                "GBps": round(b32 / ms32 / 1e6, 1)}}}


def _served_rel_l2(a, b):
    """Visible-cell relative L2 difference of two served results on the same
    hierarchy (None when the hierarchies differ)."""
    if [list(map(list, a.H.L(k).box)) for k in range(1, a.H.nlev + 1)] != \
            [list(map(list, b.H.L(k).box)) for k in range(1, b.H.nlev + 1)]:
        return None
    num = den = 0.0
    for k in range(a.H.rootIdx, a.H.nlev + 1):
        L = a.H.L(k)
        for p in range(L.np):
            vis = ~L.covered[p]
            x = a.Q[k - 1][p][1:-1, 1:-1][vis]
            y = b.Q[k - 1][p][1:-1, 1:-1][vis]
            num += float(((x - y) ** 2).sum())
            den += float((y ** 2).sum())
    return (num / den) ** 0.5


def bench_serving(nreq, rates):
    """Served stream: nreq requests of the seeded fixture with Poisson
    arrivals at each rate (req/s), continuous batching (FP32-inner and FP64)
    against one-at-a-time plain FP64 ebSolve on the same arrivals."""
    import json as _json

    from bsamgpu import configs, serving

    with open(os.path.join(ROOT, "tests", "golden", "config5_requests.json")) as f:
        fx = _json.load(f)

    def reqs(rate):
        return [serving.Request(i, configs.served_problem(fx["requests"][i]),
                                configs.served_options(fx["requests"][i]), fx["arrivals_at_rate_1"][i] / rate)
                for i in range(nreq)]

    out = {"requests": nreq, "tolerance": 1e-8, "rates": {}}
    # warm-up of the batched kernels and the host pool
    srv = serving.Server(slots=256, mem_budget=32 << 30, precision="mixed")
    srv.run(reqs(1.0)[:8], realtime=False)
    srv.close()
    # one at a time: service times of plain FP64 solves, FIFO on the arrivals
    service = []
    t0 = time.perf_counter()
    for r in reqs(1.0):
        t = time.perf_counter()
        serving.serial_solve(r, "fp64")
        service.append(time.perf_counter() - t)
    total = time.perf_counter() - t0
    log(f"serving: one-at-a-time mean service {1e3 * total / nreq:.2f} ms")
    for rate in rates:
        res = {}
        done = {}
        for prec in ("mixed", "fp64"):
            srv = serving.Server(slots=256, mem_budget=32 << 30, precision=prec)
            t = time.perf_counter()
            d = srv.run(reqs(rate), realtime=True)
            wall = time.perf_counter() - t
            srv.close()
            done[prec] = {r.id: r for r in d}
            bad = [r.id for r in d if r.error is not None or r.stages[-1]["finalRel"] > r.opts.RelTol]
            st = serving.latency_stats(d)
            res[f"batched_{prec}"] = {"p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2),
                                      "throughput_per_s": round(nreq / max(r.t_done for r in d), 2),
                                      "wall_s": round(wall, 3), "ticks": srv.ticks,
                                      "mean_batch": round(srv.tick_items / max(srv.ticks, 1), 1),
                                      "peak_budget_bytes": srv.peak_bytes, "unconverged": bad}
            log(f"serving at {rate}/s {prec}: {res[f'batched_{prec}']}")
        errs = [e for e in (_served_rel_l2(done["mixed"][i], done["fp64"][i]) for i in done["fp64"]) if e is not None]
        res["batched_mixed"]["rel_l2_vs_fp64_max"] = float(max(errs)) if errs else None
        res["batched_mixed"]["rel_l2_vs_fp64_median"] = float(np.median(errs)) if errs else None
        res["batched_mixed"]["same_hierarchy_requests"] = len(errs)
        fin = 0.0
        lat = []
        for r, sv in zip(reqs(rate), service):
            fin = max(fin, r.arrival) + sv
            lat.append(fin - r.arrival)
        res["one_at_a_time_fp64"] = {"p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
                                     "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
                                     "throughput_per_s": round(nreq / fin, 2),
                                     "mean_service_ms": round(1e3 * total / nreq, 2),
                                     "utilisation": round(rate * total / nreq, 2)}
        out["rates"][str(rate)] = res
    return out


def _rates(text):
    return [float(x) for x in str(text).split(",") if x.strip()]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-only", action="store_true")
    ap.add_argument("--smoother-launches", type=int, default=400)
    ap.add_argument("--fp64-only", action="store_true", help="time only the FP64 solves (profiling)")
    ap.add_argument("--smoother", default=None, choices=["exact", "sweep", "tiled"],
                    help="smoother form of the headline solves (default: the autotuner's bit-exact pick)")
    ap.add_argument("--coarse-cells", type=int, default=0)
    ap.add_argument("--sweep", action="store_true", help="tuning sweep of the coarse-cycle cut")
    ap.add_argument("--no-parabolic", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="skip the per-hierarchy option search")
    ap.add_argument("--large-smoother", type=int, default=4096, metavar="N",
                    help="also time the smoother on one uniform N x N level (0: skip)")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-varcoef", action="store_true", help="skip the refined variable-coefficient workload")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-rates", default="50,400", help="comma-separated Poisson arrival rates (req/s)")
    ap.add_argument("--only", choices=["parabolic", "serving", "varcoef", "uniform", "uniform_solve"],
                    help="run one secondary workload and print it")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="launch a plain device copy of MIB MiB 20 times (scale for the PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    # host threads for the CPU baseline: the process's CPU share (16 per GPU
    # on a shared node) unless OMP_NUM_THREADS says otherwise
    host_threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or min(16, len(os.sched_getaffinity(0)))
    os.environ.setdefault("OMP_NUM_THREADS", "1")

    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(20):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"copy calibration: 20 copies of {args.copy_calibration} MiB (read + write each)")
        return

    from bsamgpu import _lib
    from bsamgpu.device import GpuHierarchy

    if args.only == "parabolic":
        print(json.dumps(bench_parabolic()), flush=True)
        return
    if args.only == "serving":
        print(json.dumps(bench_serving(args.serve_requests, _rates(args.serve_rates))), flush=True)
        return
    if args.only == "varcoef":
        print(json.dumps(bench_varcoef_refined()), flush=True)
        return
    if args.only == "uniform":
        # --no-autotune: the smoother launches only, no whole solves
        print(json.dumps(bench_large_smoother(args.large_smoother or 4096, solves=not args.no_autotune)),
              flush=True)
        return
    if args.only == "uniform_solve":
        print(json.dumps(bench_uniform_solves(args.large_smoother or 4096)), flush=True)
        return

    log("building starfish hierarchy on the host")
    t = time.perf_counter()
    H, opts = build_starfish()
    log(f"hierarchy: {H.nlev} levels, patches {[H.L(k).np for k in range(1, H.nlev + 1)]} "
        f"({time.perf_counter() - t:.1f} s)")
    t = time.perf_counter()
    G = GpuHierarchy(H)
    upload_s = time.perf_counter() - t
    log(f"device hierarchy (tables + upload): {upload_s:.3f} s")
    kf = H.nlev
    sm_bytes = G.smoother_bytes(kf, 8)
    if args.kernel_only:
        G.time_smoother(kf, 20)
        ms = G.time_smoother(kf, args.smoother_launches)
        log(f"smoother level {kf}: {ms * 1e3:.2f} us/launch, {sm_bytes / ms / 1e6:.1f} GB/s")
        return

    if args.sweep:
        for prec in ("fp64", "mixed"):
            for sm in ("exact", "sweep", "tiled"):
                for cc in (-1, 300, 1100, 4100):
                    tot, _, st = time_solves(G, args.steps, args.warmup, prec, coarse_cells=cc, smoother=sm)
                    log(f"sweep {prec} {sm} coarse_cells={cc}: {tot / args.steps * 1e3:.3f} ms/solve "
                        f"({st[-1]['vcycles']} cycles, rel {st[-1]['finalRel']:.2e})")
        return

    tune = None
    if not args.no_autotune:
        from bsamgpu.autotune import autotune

        best, table = autotune(G, bit_exact_only=False, repeats=3)
        exact = [r for r in table if r["smoother"] in ("exact", "sweep") and r["precision"] == "fp64" and
                 r["finalRel"] <= opts.RelTol]
        pick = min(exact, key=lambda r: r["seconds"])
        if args.coarse_cells == 0:
            args.coarse_cells = pick["coarse_cells"]
        if args.smoother is None:
            args.smoother = pick["smoother"]
        tune = {"fastest": best, "fastest_bit_exact": {k: pick[k] for k in ("coarse_cells", "smoother", "precision")},
                "table": [{k: (float(f"{v:.6g}") if isinstance(v, float) else v) for k, v in r.items()} for r in table]}
        log(f"autotune: fastest {best}, fastest bit-exact {tune['fastest_bit_exact']}")

    args.smoother = args.smoother or "exact"
    log(f"FP64 solves (smoother {args.smoother}, coarse_cells {args.coarse_cells})")
    tot64, lat64, st64 = time_solves(G, args.steps, args.warmup, "fp64", smoother=args.smoother, coarse_cells=args.coarse_cells)
    s64 = st64[-1]
    log(f"fp64: {s64['vcycles']} V-cycles, rel {s64['finalRel']:.3e}, {tot64 / args.steps * 1e3:.2f} ms/solve")
    if args.fp64_only:
        return
    Q64 = {k: G.get_field(k, _lib.BSAM_Q) for k in range(1, H.nlev + 1)}

    log("mixed-precision solves")
    totmx, latmx, stmx = time_solves(G, args.steps, args.warmup, "mixed", smoother=args.smoother, coarse_cells=args.coarse_cells)
    smx = stmx[-1]
    Qmx = {k: G.get_field(k, _lib.BSAM_Q) for k in range(1, H.nlev + 1)}
    err_mx = visible_rel_l2(H, Qmx, Q64)
    log(f"mixed: {smx['vcycles']} V-cycles, rel {smx['finalRel']:.3e}, {totmx / args.steps * 1e3:.2f} ms/solve, "
        f"rel L2 vs fp64 {err_mx:.2e}")

    # dominant kernel: the finest level's smoother launch in the form the
    # solve uses, HIP events on the library stream.  A region launch runs a
    # whole relax call (NPre red+black sweeps) over one pass of the level's
    # data; a colour pass is half a sweep.
    G.reset()
    G.solve_cycles(smoother=args.smoother, coarse_cells=args.coarse_cells)  # the options the timing uses
    sm_bytes = G.smoother_bytes(kf, 8)
    region = args.smoother == "exact" and int(G.export(kf, 5)[0]) > 0
    sweeps = opts.NPre if region else 0.5
    G.time_smoother(kf, 20)
    ms64 = G.time_smoother(kf, args.smoother_launches)
    G.time_smoother(kf, 20, mixed=True)
    ms32 = G.time_smoother(kf, args.smoother_launches, mixed=True)
    gbs64 = sm_bytes / ms64 / 1e6
    gbs32 = G.smoother_bytes(kf, 4) / ms32 / 1e6
    per_level = {}
    per_level_us = {}
    for k in range(2, H.nlev + 1):
        msk = G.time_smoother(k, 100)
        per_level[str(k)] = round(G.smoother_bytes(k, 8) / msk / 1e6, 1)
        per_level_us[str(k)] = round(msk * 1e3, 2)
    # the colour-pass form on the same level, for comparison
    G.reset()
    G.solve_cycles(smoother="classic", coarse_cells=args.coarse_cells)
    G.time_smoother(kf, 20)
    ms_pass = G.time_smoother(kf, args.smoother_launches)
    pass_bytes = G.smoother_bytes(kf, 8)
    relax_call_pass_us = ms_pass * 1e3 * 2 * opts.NPre
    log(f"smoother fp64 {ms64 * 1e3:.2f} us/launch {gbs64:.0f} GB/s; fp32 {ms32 * 1e3:.2f} us {gbs32:.0f} GB/s; "
        f"colour pass {ms_pass * 1e3:.2f} us ({pass_bytes / ms_pass / 1e6:.0f} GB/s)")

    large = None
    if args.large_smoother:
        log(f"smoother on a uniform {args.large_smoother}^2 level")
        large = bench_large_smoother(args.large_smoother)
        log(f"large smoother: {large}")

    varc = None
    if not args.no_varcoef:
        log("variable-coefficient workload refined to 2048^2 spacing")
        varc = bench_varcoef_refined()
        log(f"varcoef: fp64 {varc['fp64']}, mixed {varc['mixed']}")

    parab = None
    if not args.no_parabolic:
        log("parabolic heat workload")
        parab = bench_parabolic()
        log(f"parabolic: {parab['fp64']}")
    serv = None
    if not args.no_serving:
        log(f"serving {args.serve_requests} requests at {args.serve_rates} req/s")
        serv = bench_serving(args.serve_requests, _rates(args.serve_rates))
        log(f"serving: {serv}")

    base = None
    if not args.no_cpu_baseline:
        log("CPU baseline (C++ restatement of the solve stage, 1 thread and all cores)")
        base = cpu_baseline(H, host_threads, s64["resHist"])
        log(f"cpu: {base['seconds_per_solve'] * 1e3:.1f} ms/solve on {base['cores']} threads, "
            f"{base['single_thread']['seconds_per_solve'] * 1e3:.1f} ms on one")

    ms_step = tot64 / args.steps * 1e3
    out = {
        "metric": "V-cycle s & smoother HBM GB/s at 1e-10 rel. residual; p50/p99 solve latency",
        "value": round(args.steps / tot64, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: starfish transmission coefficients sampled on the host (no dataset)",
        "config": {"workload": "starfish 64->1024 composite solve (5 AMR levels, 7 MG levels, "
                               "544664 stored / 408562 visible cells) from zero to rel 1e-10",
                   "levels": H.nlev, "patches_per_level": [H.L(k).np for k in range(1, H.nlev + 1)],
                   "rel_tol": opts.RelTol},
        "vcycles": s64["vcycles"],
        "hierarchy_create_s": round(upload_s, 4),
        "coarse_cells": args.coarse_cells,
        "autotune": tune,
        "final_rel": s64["finalRel"],
        "vcycle_s": round(ms_step / 1e3 / s64["vcycles"], 7),
        "p50_ms": round(pct(lat64, 50), 3),
        "p99_ms": round(pct(lat64, 99), 3),
        "smoother_GBps": round(gbs64, 1),
        "smoother_GBps_per_level": per_level,
        "smoother_us_per_level": per_level_us,
        "smoother_uniform_level": large,
        "mixed": {"value": round(args.steps / totmx, 3), "ms_per_step": round(totmx / args.steps * 1e3, 3),
                  "vcycles": smx["vcycles"], "final_rel": smx["finalRel"],
                  "speedup_vs_fp64": round(tot64 / totmx, 3), "rel_l2_vs_fp64": err_mx,
                  "p50_ms": round(pct(latmx, 50), 3), "p99_ms": round(pct(latmx, 99), 3),
                  "smoother_fp32_GBps": round(gbs32, 1)},
        "roofline": {"bound": "hbm", "achieved": round(gbs64, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs64 / HBM_PEAK_GBS, 4),
                     "traffic": pmc_traffic("k_relax_region_starfish_level7" if region
                                            else "k_relax_c_starfish_level7"),
                     "kernel": (f"k_relax_region<double,1024,1,1> (a whole relax call: {opts.NPre} red+black sweeps "
                                f"per launch on LDS canvases)" if region else
                                "k_relax_c<double,10> (colour pass, neighbour-assembled ghosts + sibling push)")
                               + f" level {kf} ({H.L(kf).np} patches, "
                               f"{sum(H.L(kf).dims(p)[0] * H.L(kf).dims(p)[1] for p in range(H.L(kf).np))} cells)",
                     "bytes_per_launch": sm_bytes, "us_per_launch": round(ms64 * 1e3, 3),
                     "sweeps_per_launch": sweeps,
                     # the algorithmic bytes of the sweeps a launch performs, per second
                     "sweep_bytes_rate_GBps": round(sweeps * gbs64 if region else gbs64, 1),
                     "sweep_bytes_rate_frac": round((sweeps * gbs64 if region else gbs64) / HBM_PEAK_GBS, 4),
                     "colour_pass_us": round(ms_pass * 1e3, 3),
                     "relax_call_us": {"launch_form": round(ms64 * 1e3 * (1 if region else 2 * opts.NPre), 2),
                                       "colour_passes": round(relax_call_pass_us, 2)}},
        "cpu_baseline": base,
        "varcoef_refined": varc,
        "parabolic": parab,
        "serving": serv,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
