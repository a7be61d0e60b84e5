"""Tables of the whole-relax-call smoother (region tiles), checked on the CPU
against the oracle's ebRelaxLevel (EllipticBSAM/src/ebRelaxLevel.m:28-38).

A region tile owns a rectangle of one patch and lists every cell of the
level whose value the own cells' values after a relax call of `rdepth`
passes depend on, with its depth (the last pass after which it must be
exact).  Checked here:
- every cell of the level is owned by exactly one tile, and the own cells
  carry the full depth;
- the canvas holds every listed cell with a slot of margin (the kernel's
  neighbour reads stay in it);
- the region is complete: perturbing every cell of the level OUTSIDE a
  tile's region and running the oracle's relax call (3 sweeps, ghosts
  refreshed after every colour) leaves the tile's own cells bit-identical,
  whereas perturbing one listed cell of depth 0 next to the own cells
  changes them (so the check has teeth);
- depths are consistent with the colour of each pass: a cell updated in
  pass p (colour (p-1) mod 2, depth >= p) has every same-patch neighbour
  in the region with depth >= p-1.
"""
import numpy as np
import pytest

from cases import CASES
from bsamgpu.device import GpuHierarchy
from oracle import bsam_oracle as O


def parse(v):
    v = [int(x) for x in v]
    ntile, rdepth = v[0], v[1]
    pos = 2
    tiles = []
    for _ in range(ntile):
        cx0, cy0, cw, ch, np_, nload, nred, nblack = v[pos:pos + 8]
        pos += 8
        pats = [tuple(v[pos + 3 * q:pos + 3 * q + 3]) for q in range(np_)]
        pos += 3 * np_
        cells = []
        for e in v[pos:pos + nload + nred + nblack]:
            ps, i, j, d = e & 15, (e >> 4) & 4095, (e >> 16) & 4095, (e >> 28) & 15
            cells.append((pats[ps][0], i, j, d))
        pos += nload + nred + nblack
        ngh = v[pos]
        pos += 1
        ghosts = [(v[pos + 7 * g], v[pos + 7 * g + 1:pos + 7 * g + 7]) for g in range(ngh)]
        pos += 7 * ngh
        nrs = v[pos]
        pos += 1
        rs = [(v[pos + 2 * e], v[pos + 2 * e + 1]) for e in range(nrs)]
        pos += 2 * nrs
        tiles.append(dict(cx0=cx0, cy0=cy0, cw=cw, ch=ch, pats=pats, cells=cells, nload=nload, nred=nred,
                          nblack=nblack, ghosts=ghosts, rs=rs))
    assert pos == len(v)
    return rdepth, tiles


def _tables(H):
    G = GpuHierarchy(H, device=-1)
    out = {k: parse(G.export(k, 5)) for k in range(2, H.nlev + 1)}
    G.close()
    return out


@pytest.mark.parametrize("name", sorted(CASES))
def test_region_tiles_cover_and_fit(name):
    H = CASES[name](seed=None)
    for k, (rdepth, tiles) in _tables(H).items():
        L = H.L(k)
        assert tiles and rdepth == 6, k
        owned = {}
        for t, T in enumerate(tiles):
            slots = set()
            pslot = {pid: (ox, oy) for pid, ox, oy in T["pats"]}
            for n, (p, i, j, d) in enumerate(T["cells"]):
                ox, oy = pslot[p]
                x, y = ox + i, oy + j
                assert 1 <= x <= T["cw"] - 2 and 1 <= y <= T["ch"] - 2
                assert (x, y) not in slots
                slots.add((x, y))
                mx, my = L.dims(p)
                assert 0 <= i < mx and 0 <= j < my
                colour = (i + j + L.par0[p]) & 1
                if n < T["nload"]:
                    assert d == 0
                else:
                    assert d >= 1 and colour == (0 if n < T["nload"] + T["nred"] else 1)
                if d == rdepth:
                    assert (p, i, j) not in owned
                    owned[(p, i, j)] = t
        assert len(owned) == sum(int(np.prod(L.dims(p))) for p in range(L.np)), k


def _perturbed_relax(Orc, k, keep, rng):
    """The oracle's relax call after every interior cell of level k outside
    `keep` (a set of (p, i, j)) is replaced by a random value."""
    A = Orc.copy()
    L = A.L(k)
    for p in range(L.np):
        mx, my = L.Q[p].shape[0] - 2, L.Q[p].shape[1] - 2
        for j in range(my):
            for i in range(mx):
                if (p, i, j) not in keep:
                    L.Q[p][i + 1, j + 1] = rng.standard_normal()
    A.relax_level(k, 3)
    return A


@pytest.mark.parametrize("name", ["lshape", "lshape_lif", "plain", "neumann", "mixed_bc"])
def test_region_is_complete_against_the_oracle_relax(name):
    H = CASES[name](seed=3)
    Orc = O.OracleHierarchy(H)
    for k in range(1, Orc.nlev + 1):
        Orc.fill_ghosts(k)
    rng = np.random.default_rng(11)
    for k, (rdepth, tiles) in _tables(H).items():
        ref = Orc.copy()
        ref.relax_level(k, 3)
        for T in tiles:
            region = {(p, i, j) for p, i, j, d in T["cells"]}
            own = [(p, i, j) for p, i, j, d in T["cells"] if d == rdepth]
            A = _perturbed_relax(Orc, k, region, rng)
            for p, i, j in own:
                assert A.L(k).Q[p][i + 1, j + 1] == ref.L(k).Q[p][i + 1, j + 1], (name, k, p, i, j)
            # teeth: perturbing the region's depth-0 cells as well changes the own cells
            d0 = {(p, i, j) for p, i, j, d in T["cells"] if d == 0}
            if d0:
                B = _perturbed_relax(Orc, k, region - d0, rng)
                assert any(B.L(k).Q[p][i + 1, j + 1] != ref.L(k).Q[p][i + 1, j + 1] for p, i, j in own)


@pytest.mark.parametrize("name", ["lshape", "plain"])
def test_region_depths_follow_the_pass_colours(name):
    H = CASES[name](seed=None)
    for k, (rdepth, tiles) in _tables(H).items():
        L = H.L(k)
        for T in tiles:
            # kRegionVirtual (15): a cell another tile owns, updated in every pass
            depth = {(p, i, j): (rdepth if d == 15 else d) for p, i, j, d in T["cells"]}
            for (p, i, j), d in depth.items():
                colour = (i + j + L.par0[p]) & 1
                mx, my = L.dims(p)
                for q in range(1, d + 1):
                    if (q - 1) % 2 != colour:
                        continue
                    for ni, nj in ((i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)):
                        if 0 <= ni < mx and 0 <= nj < my:
                            assert depth.get((p, ni, nj), -1) >= q - 1, (k, p, i, j, q)


def test_tiles_merge_pieces_of_several_patches():
    """Adjacent pieces share a tile when the merged region still fits one
    workgroup: on the L-shaped union a tile owns cells of two patches, and
    the starfish-like packing keeps every tile within the workgroup's lanes."""
    H = CASES["lshape"](seed=None)
    merged = 0
    for k, (rdepth, tiles) in _tables(H).items():
        for T in tiles:
            own_patches = {p for p, i, j, d in T["cells"] if d == rdepth}
            merged += len(own_patches) > 1
            assert T["nred"] <= 1024 and T["nblack"] <= 1024 and T["nload"] <= 1024
            assert T["cw"] * T["ch"] <= 4096
    assert merged >= 2
