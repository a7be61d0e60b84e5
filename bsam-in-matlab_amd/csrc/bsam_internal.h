// Internal layout of the device-resident patch hierarchy.
//
// Every level stores its patches packed back to back in one buffer per field,
// each patch column-major (x index fastest), with a per-patch offset record:
//   Q     (mx+2)(my+2)  padded solution            PatchInfo::qoff
//   F, FT, CC, ID, vis/tgt maps   mx*my            PatchInfo::coff
//   DeW   (mx+1)*my     x-faces                    PatchInfo::xoff
//   DnS   mx*(my+1)     y-faces                    PatchInfo::yoff
//   QC    (mx/2+2)(my/2+2) coarse shadow + ring    PatchInfo::qcoff
//   G     boundary data W(my) E(my) S(mx) N(mx)    PatchInfo::goff
// All table indices are absolute element indices into a level's buffers, so a
// launch sweeps every patch of a level through one offset table.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

namespace bsam {

constexpr int kTile = 256;  // cells per work tile (4 wave64)

struct PatchInfo {
  int mx, my, par0, cf;   // cf: patch has an internal face or CF corner
  int bc[4];              // W E S N: 0 internal, 1 Dirichlet, 2 Neumann
  int qoff, coff, xoff, yoff, qcoff, goff, ccoff;
  int gbase;              // first ghost record: W(my) E(my) S(mx) N(mx) SW SE NW NE
  unsigned magic;         // c / mx == umulhi(c, magic) for every cell c (0: divide)
};

// ghost record kinds (one record per ghost cell of a level)
enum GhostKind : int {
  G_COPY = 0,      // Q[s0] (same-level sibling interior)
  G_QIF = 1,       // (a1 Q[s0] + a2 Q[s1]) + a3 P[s2]
  G_QIF_END = 2,   // (a1 Q[s0] + a2 Q[s1]) + a3 ((b1 P[s2] + b2 P[s3]) + b3 P[s4])
  G_LIF = 3,       // (2/3 P[s2] + 2/3 Q[s0]) - 1/3 Q[s1]
  G_LIF_CORNER = 4,// 2/3 P[s2] + 1/3 Q[s0]
  G_DIR = 5,       // (8/3 g - 2 Q[s0]) + 1/3 Q[s1]
  G_NEU = 6,       // Q[s0] + (h g)
  G_PCORNER = 7,   // (side(s0) + side(s1)) - Q[s2]; s0,s1 are record ids
};

struct GhostRec {
  int dst, kind;
  int s[5];
  int gi;    // G_DIR / G_NEU: boundary-data index of c; G_COPY on a face: the
             // sibling's ghost slot that copies this record's own boundary cell
             // (the cell a smoother pass pushes its new value to); else -1
  double c;  // g (Dirichlet) or h*g (Neumann)
};

// per ghost record, what the smoother's neighbour-assembled passes need in
// one 16-byte load: the parent part of a coarse-fine value (rewritten by
// fills and keeping passes), the record kind and the sibling push target
template <typename T>
struct alignas(16) FaceAux {
  T cf;
  int kind;
  int push;  // GhostRec::gi of a G_COPY face record, else -1
};

struct RingRec {
  int dst, src;  // QC index <- parent Q index
};

// A 256-cell work tile with its patch record inline, so that a kernel reaches
// every offset it needs with one (scalar) load instead of a tile -> patch
// chain of dependent loads.
// 96 bytes, 16-aligned: a kernel reads a whole record with wide loads
struct alignas(16) TileRec {
  PatchInfo pi;
  int p, start, pad0, pad1;
};

// per coarse footprint cell: parent indices for restriction / FAS load, with
// the fine patch's offsets inline (one dependent load fewer)
struct FootRec {
  int p, ci, cj;       // fine patch and coarse cell (0-based, interior)
  int pq;              // parent Q index (padded) of the coarse cell
  int pc;              // parent cell index (F/CC layout)
  int dew_w, dew_e;    // parent DeW indices of the cell's west/east faces
  int dns_s, dns_n;    // parent DnS indices of the cell's south/north faces
  int qoff, coff, xoff, yoff, qcoff, mx, my;  // fine patch offsets and size
};

// flux-balance (reflux) contribution of one coarse CF face
struct CorrRec {
  int d1, d2;          // face diffusivities (DeW or DnS)
  int in1, g1, in2, g2;// fine interior / ghost Q indices
  int ring, qcin;      // QC ring / QC interior indices
  int face;
  int cfidx;           // parent DeW (W/E) or DnS (S/N) index of cf
  double cf;           // coarse interface-face diffusivity
};

// smoother tile: interior cells [i0,i1] x [j0,j1] (0-based) of patch p
struct STile {
  int p, i0, i1, j0, j1;
};

// Fused red+black sweep tiles (levels with L.arith).  A tile is a rectangle
// of one patch, at most kSweepTW x kSweepTH cells.  Its halo lists the
// colour-1 cells outside the tile that the tile's colour-2 cells read in the
// second half of a sweep (same-patch neighbours and coarse-fine
// interpolation sources, sibling cells across copy faces), each with the
// slot of its position in the tile's (tw+4) x (th+4) region and its patch
// (0: the tile's patch, s >= 1: sib[s-1]) and local cell coordinates.
constexpr int kSweepTW = 32, kSweepTH = 32;
constexpr int kSweepMaxSib = 8;
constexpr int kSweepMaxHalo = 160;
constexpr int kSweepRegion = (kSweepTW + 4) * (kSweepTH + 4);
struct SweepHalo {
  int slot_ps;  // region slot | (patch slot << 16)
  int ij;       // i | (j << 16), local to the cell's patch
};
struct alignas(16) SweepTile {
  int p, i0, j0, tw, th, nhalo, nsib, pad;
  PatchInfo pi;
  PatchInfo sib[kSweepMaxSib];
  SweepHalo halo[kSweepMaxHalo];
};

// Whole relax calls in one launch (levels with L.arith): a region tile owns
// the cells of a rectangle of one patch and stages, in an LDS canvas laid out
// in the level's index space, every cell whose value after some pass the
// owned cells' final values depend on (same-patch neighbours, sibling cells
// across copy faces, coarse-fine interpolation sources).  Each such cell has a
// depth d: its value is exact after passes 1..d of a call of `rdepth` passes,
// the tile's own cells have depth rdepth, and a pass p updates the cells of
// its colour with depth >= p (a cell of depth 0 is only read).
// Cell entries: patch slot (bits 0-3) | i (4-15) | j (16-27) | depth (28-31),
// (i, j) 0-based in the patch; the lists of a tile are its depth-0 cells,
// then its colour-1 cells, then its colour-2 cells.
constexpr int kRegionMaxPatches = 16;
constexpr int kRegionMaxSlots = 4096;
constexpr int kRegionMaxDepth = 14;
constexpr int kRegionThreads = 1024;  // lanes of a region workgroup
constexpr int kRegionK = 1;           // updated cells per lane and colour
constexpr int kRegionKL = 1;          // depth-0 cells per lane
// entries per tile: [colour 1: NT*K][colour 2: NT*K][depth 0: NT*KL], the
// unused ones kRegionNone; the patches of a tile take kRegionMaxPatches slots
constexpr int kRegionStride = kRegionThreads * (2 * kRegionK + kRegionKL);
constexpr unsigned kRegionNone = 0xffffffffu;
// depth code of a cell updated in every pass but written by another tile: a
// source of a ghost or restriction that this tile writes
constexpr int kRegionVirtual = 15;
struct alignas(16) RegionTile {
  int cx0, cy0, cw, ch;   // canvas origin (0-based level index) and size
  int np, nload, nred, nblack;
  int gofs, ngh;          // the tile's ghost records in the level's rghost list
  int rofs, nrs;          // the coarse cells it restricts in the level's rrs list
  int wofs, nwin;         // its parent-window cells in the level's rwin list (fused correction)
};
// A ghost record the tile evaluates after the last pass (ebFillGhosts from
// its final values), with everything but the parent parts resolved on the
// host: the record's kind and boundary value (for a physical corner, those of
// its two side records), the aux entries holding the parent parts (-1: none),
// and the canvas slots of its fine sources -- [0..1] the record's own; for a
// physical corner [0..1] and [2..3] those of its side records and [4] the
// diagonal cell
constexpr unsigned short kRegionNoSlot = 0xffff;
struct alignas(16) RegionGhost {
  double c, ca, cb;   // boundary values (G_DIR: g, G_NEU: h g) of the record / its sides
  int dst;            // Q index written
  int kinds;          // kind | side A's kind << 4 | side B's kind << 8
  int rec, reca, recb;// aux entries of the parent parts (-1: none)
  unsigned s01, s23, s45;  // slots 0..5, two per word (low half first)
};
__host__ __device__ inline unsigned region_slot(const RegionGhost& g, int k) {
  const unsigned w = k < 2 ? g.s01 : k < 4 ? g.s23 : g.s45;
  return (k & 1) ? (w >> 16) : (w & 0xffffu);
}
// A coarse cell the tile restricts: QC index, parent Q index, canvas slot of
// the block's first fine cell
struct alignas(16) RegionRestrict {
  int qc, pq, slot, pad;
};
struct RegionPatch {
  PatchInfo pi;  // the patch record
  int ox, oy;    // canvas position of the patch's cell (0, 0)
  // the rectangle of the patch's QC window [ex0, ex0 + ew) x [ey0, ...) that
  // the tile stages (parent value and QC) at window slots from eoff, row by row
  int eoff, ex0, ey0, ew;
};
// A parent-window cell a tile stages for the correction folded into its
// post-smoothing launch: the QC index and the parent Q index it maps to
struct RegionWin {
  int w, wm;
};
// The window slots of the parent sources of a region ghost record's
// coarse-fine parts: [0..2] the record's own (or a physical corner's side A),
// [3..5] a physical corner's side B; kRegionNoSlot where unused
struct alignas(16) RegionGhostW {
  unsigned short s[8];
};
constexpr int kRegionMaxWin = 2048;  // window cells per tile (two per lane)
__host__ __device__ inline unsigned region_code(int ps, int i, int j, int d) {
  return (unsigned)ps | ((unsigned)i << 4) | ((unsigned)j << 16) | ((unsigned)d << 28);
}

// device tables and buffers of one level
template <typename T>
struct LevelDev {
  int n[2];
  double h[2];
  double ih2[2];     // 1/h^2 (exact when pow2)
  int pow2;          // h[0], h[1] are powers of two: q/h^2 == q*(1/h^2) bit for bit
  int even_mx;       // every patch has an even width (colour-compacted passes)
  int arith;         // face records follow the fixed interpolation stencil and
                     // sibling faces carry push targets: passes after the first
                     // of a relax call read materialised sibling ghosts and
                     // evaluate coarse-fine ghosts from neighbours + cf
  int np, ncell, nq, nx, ny, nqc, ng;
  int ntile, nctile, nghost, nring_all, nring_cf, nfoot, ncorr, ntgt;
  int nstile, stile_lds;  // smoother tiles and their largest LDS region (cells)
  STile* stiles;
  int stream;             // one patch of even width, physical on every side: streaming sweeps apply
  int nsweep;             // fused sweep tiles (0: the level runs colour passes)
  SweepTile* sweep;
  T* Qs;                  // the sweep's second copy of Q (same layout)
  int nregion;            // region tiles (0: the level's relax calls take other forms)
  int rdepth;             // passes per relax call the region tables cover
  int rslots;             // largest canvas of the level's region tiles
  RegionTile* region;
  RegionPatch* rpatch;    // kRegionMaxPatches per tile
  unsigned* rcell;        // kRegionStride per tile
  RegionGhost* rghost;    // per tile (RegionTile::gofs, ngh)
  RegionGhostW* rgw;      // parallel to rghost: window slots of their parent parts (rcorr)
  RegionWin* rwin;        // per tile (RegionTile::wofs, nwin)
  int rcorr;              // the post-smoothing launch can apply the coarse-grid correction itself
  int rwinmax;            // the most window cells of one tile
  RegionRestrict* rrs;    // per tile (RegionTile::rofs, nrs)
  int rghmax, rrsmax;     // the most ghosts / restricted cells of one tile
  PatchInfo* P;
  TileRec* tiles;    // patch record + first cell per 256-cell tile
  int2* ctiles;      // (patch, first coarse cell) per tile of the footprint
  GhostRec* ghost;
  RingRec* ring_all;
  RingRec* ring_cf;
  FootRec* foot;
  CorrRec* corr;     // this level's CF faces (targets live on level k-1)
  int* tgt_cell;     // targets on THIS level (from level k+1's corrections)
  int* tgt_ptr;
  int* tgt_ent;      // entries into level k+1's corr array
  int* tgt_of_cell;  // per cell: target row or -1
  int* wmap;         // per QC cell: parent Q index
  int* wgr;          // per QC cell: the parent ghost record at wmap (-1: interior)
  int2* pcov;        // per parent Q index: (Q index of the first fine cell of the 2 x 2 block
                     // that covers it, fine row stride), x = -1 where no patch covers it
  int* ring_rec;     // per ring_all item: the parent ghost record at its source (-1: interior)
  int pgh_ok;        // every such record is physical (the load may evaluate it)
  uint8_t* vis;      // per cell: 1 if not covered by the next level
  T *Q, *F, *FT, *CC, *ID, *DeW, *DnS, *QC, *G;
  FaceAux<T>* aux;   // per ghost record (null: parent parts not kept)
};

// A level-1 ghost record whose source cells are all rows of one 64-row block
// of the coarsest solve: rows[0..1] the record's sources (a Neumann record
// repeats rows[0]); for a physical corner rows[0..1] and [2..3] those of its
// two side records and rows[4] the diagonal cell; else rows[2..4] = rows[0]
struct CoarsestGhost {
  int rec;
  int rows[5];
};
struct CoarsestDev {
  int N, M;          // unknowns, matrix order (N or N+1 with the KKT row)
  double* Ainv;      // M x M column-major
  double* bcRhs;     // N
  // level 1's ghost records by 64-row block (gofs: CSR offsets), evaluated
  // by the block that computes their sources; null: a fill launch follows
  CoarsestGhost* gtab;
  int* gofs;
};

// scalars of the composite residual
struct ResidualOut {
  double linf, l2sq, fnorm, pad;
  long long ncell, pad2[3];
};

// one hierarchy of a batched launch: every workgroup runs one item's whole
// preamble (phase 0: fill + baseline residual) or cycle (phase 1)
struct BatchItem {
  const LevelDev<double>* LV64;  // device array of the FP64 level descriptors
  const LevelDev<float>* LV32;   // FP32 levels (mixed precision) or null
  float* const* rout32;          // per-level FP32 defect arrays (mixed) or null
  ResidualOut* out;              // residual scalars of this item
  ResidualOut* parts;            // per-workgroup residual partials (nwg > 1)
  CoarsestDev cst;
  int nlev, root, phase;
  int wg0, nwg;                  // the item's workgroups in its launch
  int pad;
};

}  // namespace bsam
