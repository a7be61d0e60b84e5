"""Per-hierarchy choice of the device execution options.

The device path has knobs that change speed but not the converged solution:
- how many of the coarsest levels run their V-cycle inside one LDS-staged
  workgroup (``coarse_cells``);
- the smoother form (``exact``: one launch per colour pass, bit-identical to
  the reference sequence; ``sweep``: the same sequence with pairs of sweeps
  fused into one launch each on levels whose 32x32 tiles fit one round of
  workgroups, also bit-identical; ``tiled``: every pass of a relax call in
  one launch on LDS tiles with sibling ghosts frozen);
- the arithmetic of the inner V-cycles (``fp64`` or ``mixed``) and, for
  ``mixed``, how many FP32 V-cycles run per FP64 residual (``inner_cycles``);
- on levels that run streaming sweeps, the row-block height of a sweep
  launch (``stream_rows``: the tile shape is 124 columns by that many rows,
  one 64-lane wave each), which changes the number of waves per CU and the
  share of recomputed halo rows but not the result.

``autotune`` times complete solves from the hierarchy's uploaded state for
each candidate and returns the fastest one whose final relative residual
meets RelTol (an error when none does); when the hierarchy has a streaming level it then times the
winner at each row-block height.  ``bit_exact_only`` restricts the search to
the candidates that reproduce the reference's FP64 sequence bit for bit.
"""
from __future__ import annotations

import itertools
import statistics
import time

EXACT_CANDIDATES = [
    dict(coarse_cells=c, smoother=s, precision="fp64") for s in ("exact", "sweep") for c in (-1, 300, 1100, 4100)
]
STREAM_ROWS = (8, 16, 32, 64)
ALL_CANDIDATES = EXACT_CANDIDATES + [
    dict(coarse_cells=c, smoother=s, precision=p)
    for c, s, p in itertools.product((300, 1100), ("exact", "tiled"), ("fp64", "mixed"))
    if not (s == "exact" and p == "fp64")
]


def time_solve(G, repeats=3, **kw):
    """Median wall time (s) of G.reset() + G.solve_cycles(**kw) and the last stage."""
    G.reset()
    G.solve_cycles(**kw)  # warm-up: graph capture, FP32 copies
    ts = []
    st = None
    for _ in range(repeats):
        t = time.perf_counter()
        G.reset()
        st = G.solve_cycles(**kw)
        ts.append(time.perf_counter() - t)
    return statistics.median(ts), st


def autotune(G, rel_tol=None, bit_exact_only=True, repeats=3, candidates=None):
    """Returns (best candidate dict, table of {candidate, seconds, vcycles, finalRel})."""
    rel_tol = G.opts.RelTol if rel_tol is None else rel_tol
    cands = candidates or (EXACT_CANDIDATES if bit_exact_only else ALL_CANDIDATES)
    table = []
    for kw in cands:
        sec, st = time_solve(G, repeats=repeats, **kw)
        table.append(dict(kw, seconds=sec, vcycles=st["vcycles"], finalRel=st["finalRel"]))
    ok = [r for r in table if r["finalRel"] <= rel_tol]
    if not ok:
        raise RuntimeError(f"autotune: no candidate met RelTol {rel_tol:g} "
                           f"(best final rel {min(r['finalRel'] for r in table):.3e})")
    best = min(ok, key=lambda r: r["seconds"])
    keys = tuple(k for k in best if k not in ("seconds", "vcycles", "finalRel", "stream_rows"))
    best = dict({k: best[k] for k in keys}, stream_rows=best.get("stream_rows", 0), seconds=best["seconds"])
    if best["smoother"] in ("exact", "sweep") and G.streaming_levels():
        for r in STREAM_ROWS:
            kw = dict({k: best[k] for k in keys}, stream_rows=r)
            sec, st = time_solve(G, repeats=repeats, **kw)
            table.append(dict(kw, seconds=sec, vcycles=st["vcycles"], finalRel=st["finalRel"]))
            if st["finalRel"] <= rel_tol and sec < best["seconds"]:
                best = dict(kw, seconds=sec)
    G.reset()
    return {k: best[k] for k in keys + ("stream_rows",)}, table
