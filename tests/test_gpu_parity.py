"""Parity of the HIP library (through the C ABI) with the CPU oracle.

FP64 device results must equal the oracle bit for bit: both evaluate the
reference's expressions in the same operation order without fused
multiply-add, and the coarsest solve uses the same explicit inverse
arithmetic.  Only the composite L2 norm (a parallel sum) is compared with a
tolerance (1e-13 relative).
"""
import numpy as np
import pytest

from cases import CASES
from bsamgpu.device import GpuHierarchy
from bsamgpu import _lib
from oracle import bsam_oracle as O

pytestmark = pytest.mark.gpu


def _cmp_level(Orc, G, k, fields=("Q", "F", "QC"), label=""):
    L = Orc.L(k)
    for name in fields:
        if name == "QC" and k == 1:
            continue
        code = {"Q": _lib.BSAM_Q, "F": _lib.BSAM_F, "QC": _lib.BSAM_QC, "FTrue": _lib.BSAM_FTRUE}[name]
        got = G.get_field(k, code)
        ref = {"Q": L.Q, "F": L.F, "QC": getattr(L, "QC", None), "FTrue": L.FTrue}[name]
        for p in range(L.np):
            if not np.array_equal(got[p], ref[p]):
                d = np.abs(got[p] - ref[p])
                idx = np.unravel_index(np.argmax(d), d.shape)
                raise AssertionError(f"{label}: level {k} patch {p} field {name} differs: max {d.max():.3e} at {idx} "
                                     f"(gpu {got[p][idx]!r} oracle {ref[p][idx]!r})")


def _cmp_all(Orc, G, label):
    for k in range(1, Orc.nlev + 1):
        _cmp_level(Orc, G, k, label=label)


@pytest.mark.parametrize("name", list(CASES))
def test_operations_bit_exact(name):
    H = CASES[name](seed=11)
    print(f"case {name}", flush=True)
    Orc = O.OracleHierarchy(H)
    G = GpuHierarchy(H)
    _cmp_all(Orc, G, "upload")
    for k in range(1, Orc.nlev + 1):
        Orc.fill_ghosts(k)
        G.fill_ghosts(k)
        _cmp_all(Orc, G, f"fill {k}")
    for k in range(2, Orc.nlev + 1):
        Orc.transfer(k, "ring")
        G.transfer(k, "ring")
    _cmp_all(Orc, G, "ring")
    for k in range(Orc.nlev, 0, -1):
        Orc.relax_level(k, 1)
        G.relax_level(k, 1)
        _cmp_all(Orc, G, f"relax {k}")
    for k in range(Orc.nlev, 1, -1):
        Orc.transfer(k, "restrict")
        G.transfer(k, "restrict")
        _cmp_all(Orc, G, f"restrict {k}")
        Orc.fill_ghosts(k - 1)
        G.fill_ghosts(k - 1)
        Orc.transfer(k, "ring")
        G.transfer(k, "ring")
        Orc.coarse_load(k)
        G.coarse_load(k)
        _cmp_all(Orc, G, f"coarse load {k}")
    Orc.coarsest_solve()
    G.coarsest_solve()
    _cmp_all(Orc, G, "coarsest")
    for k in range(2, Orc.nlev + 1):
        Orc.transfer(k, "correct")
        G.transfer(k, "correct")
        _cmp_all(Orc, G, f"correct {k}")
    Orc.transfer(0, "filldown")
    G.transfer(0, "filldown")
    _cmp_all(Orc, G, "filldown")
    ro = Orc.comp_residual()
    rg = G.comp_residual()
    _cmp_all(Orc, G, "residual state")
    assert rg["linf"] == ro["linf"] and rg["fnorm"] == ro["fnorm"] and rg["ncell"] == ro["ncell"]
    assert rg["rel"] == ro["rel"]
    assert abs(rg["l2"] - ro["l2"]) <= 1e-13 * ro["l2"]
    for k in range(1, Orc.nlev + 1):
        for p in range(Orc.L(k).np):
            Orc.L(k).F[p] = Orc.L(k).FTrue[p].copy()
    G.reset_force()
    Orc.vcycle(Orc.nlev)
    G.vcycle()
    _cmp_all(Orc, G, "vcycle")


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("graph,coarse,smoother", [(False, -1, "exact"), (True, -1, "exact"), (True, 0, "exact"),
                                                    (True, 100000, "exact"), (False, -1, "sweep"),
                                                    (True, 0, "sweep")])
def test_solve_cycles_bit_exact(name, graph, coarse, smoother):
    """coarse = -1: every level as multi-workgroup launches; 0: default cut;
    100000: the whole V-cycle inside one workgroup.  smoother 'sweep': pairs
    of red+black sweeps as fused launches where the level's tiles allow."""
    H = CASES[name](seed=None)
    print(f"case {name} graph={graph} coarse={coarse} smoother={smoother}", flush=True)
    Orc = O.OracleHierarchy(H)
    G = GpuHierarchy(H)
    so = Orc.solve_cycles()
    sg = G.solve_cycles(use_graph=graph, coarse_cells=coarse, smoother=smoother)
    assert sg["vcycles"] == so["vcycles"]
    assert sg["resHist"] == so["resHist"], (sg["resHist"], so["resHist"])
    assert sg["finalRel"] == so["finalRel"]
    assert sg["visibleCells"] == so["visibleCells"]
    assert sg["totalCells"] == so["totalCells"]
    _cmp_all(Orc, G, "solve")


def _visible_rel_l2(Orc, G):
    """Composite relative L2 difference over visible cells; for the singular
    all-Neumann case (solution defined up to a constant, ebSolve.m:56-68
    pins the mean afterwards) both sides are compared with their composite
    means removed."""
    got = {k: G.get_field(k, _lib.BSAM_Q) for k in range(Orc.rootIdx, Orc.nlev + 1)}
    ma = mb = 0.0
    if Orc.coarsest["singular"]:
        area = 0.0
        for k in got:
            L = Orc.L(k)
            da = L.h[0] * L.h[1]
            for p in range(L.np):
                vis = ~L.covered[p]
                ma += float(got[k][p][1:-1, 1:-1][vis].sum()) * da
                mb += float(L.Q[p][1:-1, 1:-1][vis].sum()) * da
                area += vis.sum() * da
        ma /= area
        mb /= area
    num = den = 0.0
    for k in got:
        L = Orc.L(k)
        da = L.h[0] * L.h[1]
        for p in range(L.np):
            vis = ~L.covered[p]
            a = got[k][p][1:-1, 1:-1][vis] - ma
            b = L.Q[p][1:-1, 1:-1][vis] - mb
            num += float(((a - b) ** 2).sum()) * da
            den += float((b ** 2).sum()) * da
    return (num / den) ** 0.5


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tiled_and_mixed_reach_the_same_solution(name, precision):
    """The LDS-tiled smoother and the FP32-inner defect-correction form change
    the iteration, not the fixed point: converged to the round-off floor
    (rel 1e-11, or the oracle's own floor) they agree with the oracle's FP64
    solution to 1e-10 relative L2 (the bound stated for FP64 results)."""
    H = CASES[name](seed=None)
    H.opts = H.opts.replace(RelTol=1e-11, MaxVCycles=40)
    Orc = O.OracleHierarchy(H)
    so = Orc.solve_cycles()
    for smoother in ("exact", "tiled", "sweep"):
        if precision == "fp64" and smoother != "tiled":
            continue  # bit-identical forms: covered by test_solve_cycles_bit_exact
        G = GpuHierarchy(H)
        sg = G.solve_cycles(smoother=smoother, precision=precision)
        err = _visible_rel_l2(Orc, G)
        print(f"case {name} {precision}/{smoother}: {sg['vcycles']} cycles (oracle {so['vcycles']}), "
              f"rel {sg['finalRel']:.2e} (oracle {so['finalRel']:.2e}), rel L2 vs oracle {err:.2e}", flush=True)
        assert sg["finalRel"] <= max(1e-11, 3 * so["finalRel"])
        assert err < 1e-10


def test_first_residual_after_hierarchies_were_freed():
    """The host waits for a residual by polling a sequence number that the
    reduction writes into the hierarchy's pinned result after the scalars.  A
    freed hierarchy's pinned block can be handed to the next one still holding
    its last number (1 after a single residual); the next hierarchy's first
    read must still return its own residual, not the scalars left there."""
    for seed in (21, 22, 23):
        old = GpuHierarchy(CASES["lshape"](seed=seed))
        for k in range(1, old.nlev + 1):
            old.fill_ghosts(k)
        old.comp_residual()
        old.close()
        H = CASES["lshape"](seed=seed + 10)
        Orc = O.OracleHierarchy(H)
        G = GpuHierarchy(H)
        for k in range(1, Orc.nlev + 1):
            Orc.fill_ghosts(k)
            G.fill_ghosts(k)
        ro, rg = Orc.comp_residual(), G.comp_residual()
        assert rg["linf"] == ro["linf"] and rg["fnorm"] == ro["fnorm"] and rg["ncell"] == ro["ncell"], seed
        st = G.solve_cycles(max_vcycles=2)  # its baseline residual is the one above
        assert st["resHist"][0] == ro["linf"] and 1 <= st["vcycles"] <= 2
        G.close()


def test_relax_omega_bit_exact():
    H = CASES["lshape"](seed=5)
    H.opts = H.opts.replace(Omega=1.3)
    Orc = O.OracleHierarchy(H)
    G = GpuHierarchy(H)
    for k in range(1, Orc.nlev + 1):
        Orc.fill_ghosts(k)
        G.fill_ghosts(k)
    for k in range(Orc.nlev, 0, -1):
        Orc.relax_level(k, 2)
        G.relax_level(k, 2)
    _cmp_all(Orc, G, "omega relax")


def test_tables_match_reference_descriptors():
    """Ghost classes (land/cornerClass), parent window sources (ring/winrect)
    and reflux targets derived by the library equal the oracle's restatement
    of ebBuildOps.m."""
    for name in ("plain", "lshape", "neumann"):
        H = CASES[name](seed=None)
        Orc = O.OracleHierarchy(H)
        G = GpuHierarchy(H)
        for k in range(2, Orc.nlev + 1):
            L = Orc.L(k)
            klass = G.export(k, 0)
            ref = []
            for p in range(L.np):
                for s in "WESN":
                    ref += [int(v) for v in L.land[p][s]]
                ref += [int(c) for c in L.cornerClass[p]]
            assert klass.tolist() == ref, (name, k)
            wm = G.export(k, 1)
            off = 0
            for p in range(L.np):
                W = np.full(L.QC[p].shape, -1, dtype=np.int64)
                Lc = Orc.L(k - 1)
                for q, si, sj, di, dj in L.ops[p]["ring"]:
                    W[di, dj] = si + sj * (Lc.Q[q].shape[0]) + 1000000 * q
                for q, s, d in L.ops[p]["winrect"]:
                    for a in range(d[0], d[1] + 1):
                        for b in range(d[2], d[3] + 1):
                            W[a, b] = (s[0] + a - d[0]) + (s[2] + b - d[2]) * Lc.Q[q].shape[0] + 1000000 * q
                n = W.size
                assert wm[off:off + n].tolist() == W.reshape(-1, order="F").tolist(), (name, k, p)
                off += n
            tg = G.export(k, 2).reshape(-1, 2)
            Lc = Orc.L(k - 1)
            coff = np.cumsum([0] + [Lc.FTrue[q].size for q in range(Lc.np)])
            ref = []
            for p in range(L.np):
                for f in range(4):
                    ent = []
                    for q, src, di, dj in L.ops[p]["corr"][f]:
                        mq = Lc.FTrue[q].shape[0]
                        for s_, a, b in zip(src, di, dj):
                            ent.append((int(s_), int(coff[q] + a + b * mq)))
                    ref += [(t, f) for _, t in sorted(ent)]
            assert [tuple(r) for r in tg.tolist()] == ref, (name, k)


def test_init_level_prolong_bit_exact():
    H = CASES["lshape"](seed=3)
    Orc = O.OracleHierarchy(H)
    G = GpuHierarchy(H)
    for k in range(2, Orc.nlev + 1):
        Orc.init_level_prolong(k)
        G.init_level_prolong(k)
        _cmp_all(Orc, G, f"prolong init {k}")


def test_errors_are_reported():
    H = CASES["plain"](seed=None)
    H.L(4).box[0] = [14, 53, 13, 52]  # misaligned (ebBuildOps.m:38-42)
    with pytest.raises(Exception) as ei:
        GpuHierarchy(H)
    assert "eb_build_ops:align" in str(ei.value)
    H = CASES["plain"](seed=None)
    H.L(4).DeW[1][3, 3] = -1.0
    with pytest.raises(Exception) as ei:
        GpuHierarchy(H)
    assert "eb_make_level:D" in str(ei.value)


@pytest.mark.parametrize("omega,cells", [(1.0, [2048, 1024]), (1.3, [2048, 1024]), (1.0, [4096, 2048])])
def test_streaming_sweeps_on_a_uniform_level(omega, cells):
    """A one-patch level with physical sides on every face and at least 2^20
    cells runs whole red+black sweeps as streaming launches by default: the
    solve equals the colour-pass solve ('passes') bit for bit (inhomogeneous
    Dirichlet and Neumann sides, variable D and C, a 2048 x 1024 finest level:
    16 full strips and a partial one, several row blocks; at 4096 x 2048 the
    first sub-root level streams too, and its odd sweep counts take their
    copy from the restriction and the correction), in FP64 and in the
    FP32-inner form."""
    from bsamgpu import Options, Problem, hierarchy

    D = lambda x, y: 1 + 0.4 * np.sin(3 * x) * np.cos(2 * y)  # noqa: E731
    C = lambda x, y: 0.5 + x * y  # noqa: E731
    f = lambda x, y: np.sin(4 * x) + y * y  # noqa: E731
    bc = {"west": ("dirichlet", lambda x, y: 1 + y), "east": ("neumann", lambda x, y: 0.5 * y),
          "south": ("dirichlet", lambda x, y: x * x), "north": ("neumann", lambda x, y: 0 * x - 1)}
    H = hierarchy.setup(Problem([0, 1, 0, 1], D, C, f, bc),
                        Options(BaseCells=cells, MaxLevels=1, Omega=omega, RelTol=1e-10, MaxVCycles=4,
                                Verbose=0))
    G = GpuHierarchy(H)
    streams = [H.nlev] if cells[0] == 2048 else [H.nlev - 1, H.nlev]
    for prec in ("fp64", "mixed"):
        out = {}
        for sm in ("passes", "exact"):
            G.reset()
            st = G.solve_cycles(smoother=sm, precision=prec)
            out[sm] = (st["resHist"], G.get_field(H.nlev, _lib.BSAM_Q))
        print(f"omega {omega} {prec}: {len(out['exact'][0]) - 1} cycles", flush=True)
        assert out["exact"][0] == out["passes"][0]
        assert all(np.array_equal(a, b) for a, b in zip(out["exact"][1], out["passes"][1]))
        # every row-block height of the streaming launch gives the same solve
        for rows in (8, 16, 32, 64):
            G.reset()
            st = G.solve_cycles(smoother="exact", precision=prec, stream_rows=rows)
            assert st["resHist"] == out["passes"][0], rows
            assert all(np.array_equal(a, b) for a, b in zip(G.get_field(H.nlev, _lib.BSAM_Q), out["passes"][1]))
    assert G.streaming_levels() == streams
    with pytest.raises(Exception) as ei:
        G.solve_cycles(stream_rows=12)
    assert "stream_rows" in str(ei.value)
    G.close()


def test_streaming_sweep_solve_equals_oracle():
    """The streaming sweep (the default smoother of a one-patch physical level
    of >= 2^20 cells) against the CPU oracle directly: two FP64 V-cycles on a
    2048 x 1024 level with inhomogeneous Dirichlet/Neumann sides and variable
    D, C reproduce the oracle's residual history and Q on every level bit
    for bit."""
    from bsamgpu import Options, Problem, hierarchy

    D = lambda x, y: 1 + 0.4 * np.sin(3 * x) * np.cos(2 * y)  # noqa: E731
    C = lambda x, y: 0.5 + x * y  # noqa: E731
    f = lambda x, y: np.sin(4 * x) + y * y  # noqa: E731
    bc = {"west": ("dirichlet", lambda x, y: 1 + y), "east": ("neumann", lambda x, y: 0.5 * y),
          "south": ("dirichlet", lambda x, y: x * x), "north": ("neumann", lambda x, y: 0 * x - 1)}
    H = hierarchy.setup(Problem([0, 1, 0, 1], D, C, f, bc),
                        Options(BaseCells=[2048, 1024], MaxLevels=1, RelTol=1e-12, MaxVCycles=2, Verbose=0))
    G = GpuHierarchy(H)
    assert G.streaming_levels() == [H.nlev]
    st = G.solve_cycles(smoother="exact")
    Orc = O.OracleHierarchy(H)
    so = Orc.solve_cycles()
    print(f"2048x1024: {st['vcycles']} cycles, history {st['resHist']}", flush=True)
    assert st["vcycles"] == so["vcycles"] == 2
    assert st["resHist"] == so["resHist"]
    _cmp_all(Orc, G, "streaming solve")
    G.close()


@pytest.mark.parametrize("name", ["plain", "neumann"])
def test_device_coarsest_inverse_is_the_host_inverse(name):
    """The coarsest inverse the device applies is the host routine's, bit
    for bit (the upload path), for a regular and a KKT system."""
    H = CASES[name](seed=None)
    print(f"case {name}", flush=True)
    G = GpuHierarchy(H)
    Hh = GpuHierarchy(H, device=-1)
    dev, host = G.export(1, 6), Hh.export(1, 6)
    assert dev.size == host.size and dev.size >= 256 * 256
    assert np.array_equal(dev, host)
    G.close()
    Hh.close()
