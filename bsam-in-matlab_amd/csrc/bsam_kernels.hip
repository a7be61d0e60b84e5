// CDNA4 (gfx950) kernels of the adaptive FAS V-cycle.
//
// One launch sweeps every patch of a level: work is described by per-level
// tables (256-cell tiles, ghost records, ring gathers, footprint cells,
// reflux targets) whose indices are absolute positions in the level's packed
// buffers.  Each operation is written once as a per-item device function; the
// multi-workgroup kernels map items to threads, and the single-workgroup
// coarse-cycle kernel runs the bottom levels' whole V-cycle with workgroup
// barriers between dependent steps.
//
// All kernels are templated on the storage type T: double for the plain path
// (built with -ffp-contract=off and evaluated in the reference's operation
// order, so FP64 results equal the CPU oracle bit for bit) and float for the
// inner V-cycles of the mixed-precision path.  Reflux sums, the composite
// residual and the coarsest solve accumulate in FP64.
//
// Formulas follow:
//   relax            EllipticBSAM/src/ebKernels.m:37-95
//   ghost records    EllipticBSAM/src/ebFillGhosts.m:18-92, ebInterpCf.m:41-153
//   restrict         EllipticBSAM/src/ebTransfer.m:10-33, ebKernels.m:98-102
//   coarse load      EllipticBSAM/src/ebCoarseLoad.m:17-32, ebKernels.m:13-25
//   reflux           EllipticBSAM/src/ebFaceCorrections.m:29-55, ebCoarseLoad.m:37-45
//   correct          EllipticBSAM/src/ebTransfer.m:57-71, ebKernels.m:105-123
//   coarsest         EllipticBSAM/src/ebCoarsest.m:79-91
//   residual         EllipticBSAM/src/ebCompResidual.m:23-84
//   composite mean   EllipticBSAM/src/ebCompositeMean.m:4-23
#include <algorithm>

#include "bsam_kernels.h"

namespace bsam {

#define A1_ (-1.0 / 5.0)
#define A2_ (2.0 / 3.0)
#define A3_ (8.0 / 15.0)
#define B1_ (5.0 / 32.0)
#define B2_ (15.0 / 16.0)
#define B3_ (-3.0 / 32.0)
#define C83_ (8.0 / 3.0)
#define C13_ (1.0 / 3.0)
#define C23_ (2.0 / 3.0)

// q / h^2 of a level.  When h is a power of two so is h^2, and q * (1/h^2) is
// the same correctly rounded number as q / h^2, at a fraction of the cost.
template <typename T>
__device__ __forceinline__ T div_h2(T q, T h2, T ih2, int pow2) {
  if (pow2) return q * ih2;
  return q / h2;
}

// row of cell c of a patch (c / mx), by a multiply-high when the host found
// a magic number exact for every cell of the patch
__device__ __forceinline__ int cell_row(int c, int mx, unsigned magic) {
  return magic ? (int)__umulhi((unsigned)c, magic) : c / mx;
}

// ---------------------------------------------------------------- ghosts
__device__ __forceinline__ bool is_cf_kind(int kind) {
  return kind == G_QIF || kind == G_QIF_END || kind == G_LIF || kind == G_LIF_CORNER;
}

// The parent part of a coarse-fine ghost value: the last-added term of the
// QIF forms and the first term of the LIF forms, so a value assembled from a
// stored copy of it rounds exactly as the whole expression does.
template <typename T>
__device__ __forceinline__ T cf_part(const GhostRec& r, const T* __restrict__ Pq) {
  switch (r.kind) {
    case G_QIF:
      return T(A3_) * Pq[r.s[2]];
    case G_QIF_END:
      return T(A3_) * ((T(B1_) * Pq[r.s[2]] + T(B2_) * Pq[r.s[3]]) + T(B3_) * Pq[r.s[4]]);
    case G_LIF:
    case G_LIF_CORNER:
      return T(C23_) * Pq[r.s[2]];
    default:
      return T(0);
  }
}

template <typename T>
__device__ __forceinline__ T ghost_value(const GhostRec& r, const T* __restrict__ Q,
                                         const T* __restrict__ Pq, int homog) {
  switch (r.kind) {
    case G_COPY:
      return Q[r.s[0]];
    case G_QIF:
    case G_QIF_END:
      return (T(A1_) * Q[r.s[0]] + T(A2_) * Q[r.s[1]]) + cf_part<T>(r, Pq);
    case G_LIF:
      return (cf_part<T>(r, Pq) + T(C23_) * Q[r.s[0]]) - T(C13_) * Q[r.s[1]];
    case G_LIF_CORNER:
      return cf_part<T>(r, Pq) + T(C13_) * Q[r.s[0]];
    case G_DIR: {
      const T g = homog ? T(0) : T(r.c);
      return (T(C83_) * g - T(2) * Q[r.s[0]]) + T(C13_) * Q[r.s[1]];
    }
    case G_NEU: {
      const T hg = homog ? T(0) : T(r.c);
      return Q[r.s[0]] + hg;
    }
    default:
      return T(0);
  }
}

// ghost_value with the parent part of a coarse-fine record taken from its
// stored copy (the same number cf_part computes from the same parent)
template <typename T>
__device__ __forceinline__ T ghost_value_cf(const GhostRec& r, const T* __restrict__ Q, T cf, int homog) {
  switch (r.kind) {
    case G_COPY:
      return Q[r.s[0]];
    case G_QIF:
    case G_QIF_END:
      return (T(A1_) * Q[r.s[0]] + T(A2_) * Q[r.s[1]]) + cf;
    case G_LIF:
      return (cf + T(C23_) * Q[r.s[0]]) - T(C13_) * Q[r.s[1]];
    case G_LIF_CORNER:
      return cf + T(C13_) * Q[r.s[0]];
    case G_DIR: {
      const T g = homog ? T(0) : T(r.c);
      return (T(C83_) * g - T(2) * Q[r.s[0]]) + T(C13_) * Q[r.s[1]];
    }
    case G_NEU: {
      const T hg = homog ? T(0) : T(r.c);
      return Q[r.s[0]] + hg;
    }
    default:
      return T(0);
  }
}

// item t of ebFillGhosts: t < nghost -> one ghost cell; then (when the level
// has a parent) one QC ring refresh of a patch with coarse-fine faces.  A
// coarse-fine record also stores its parent part in L.aux for the smoother
// passes that assemble ghosts from neighbours.
// Qsrc: where the interior values are read (L.Q, or a second copy that
// holds the level's new interior values)
template <typename T>
__device__ __forceinline__ void fill_item(const LevelDev<T>& L, const T* __restrict__ Pq, int t,
                                          int homog, const T* __restrict__ Qsrc = nullptr) {
  if (t < L.nghost) {
    const T* __restrict__ Qr = Qsrc ? Qsrc : L.Q;
    const GhostRec r = L.ghost[t];
    T v;
    if (r.kind == G_PCORNER) {
      // physical corner from the two adjacent side ghosts (ebFillGhosts.m:85-89)
      const T a = ghost_value<T>(L.ghost[r.s[0]], Qr, Pq, homog);
      const T b = ghost_value<T>(L.ghost[r.s[1]], Qr, Pq, homog);
      v = (a + b) - Qr[r.s[2]];
    } else {
      v = ghost_value<T>(r, Qr, Pq, homog);
      if (L.aux && is_cf_kind(r.kind)) L.aux[t].cf = cf_part<T>(r, Pq);
    }
    L.Q[r.dst] = v;
  } else if (Pq && t < L.nghost + L.nring_cf) {
    const RingRec rr = L.ring_cf[t - L.nghost];
    L.QC[rr.dst] = Pq[rr.src];
  }
}

// A level's interior values as they stand after the restriction from its
// child: a cell the child covers reads the 2 x 2 average of the child's cells
// (restrict_item's expression), any other cell its own value
template <typename T>
struct CoveredQ {
  const T* __restrict__ Q;
  const T* __restrict__ Qf;
  const int2* __restrict__ cov;
  __device__ __forceinline__ T operator()(int q) const {
    const int2 c = cov[q];
    if (c.x < 0) return Q[q];
    return T(0.25) * (((Qf[c.x] + Qf[c.x + 1]) + Qf[c.x + c.y]) + Qf[c.x + c.y + 1]);
  }
};

// ghost_value with the interior read through V
template <typename T, typename V>
__device__ __forceinline__ T ghost_value_via(const GhostRec& r, const V& Q, const T* __restrict__ Pq, int homog) {
  switch (r.kind) {
    case G_COPY:
      return Q(r.s[0]);
    case G_QIF:
    case G_QIF_END:
      return (T(A1_) * Q(r.s[0]) + T(A2_) * Q(r.s[1])) + cf_part<T>(r, Pq);
    case G_LIF:
      return (cf_part<T>(r, Pq) + T(C23_) * Q(r.s[0])) - T(C13_) * Q(r.s[1]);
    case G_LIF_CORNER:
      return cf_part<T>(r, Pq) + T(C13_) * Q(r.s[0]);
    case G_DIR: {
      const T g = homog ? T(0) : T(r.c);
      return (T(C83_) * g - T(2) * Q(r.s[0])) + T(C13_) * Q(r.s[1]);
    }
    case G_NEU: {
      const T hg = homog ? T(0) : T(r.c);
      return Q(r.s[0]) + hg;
    }
    default:
      return T(0);
  }
}

// fill_item with the interior read through V (its writes are fill_item's)
template <typename T, typename V>
__device__ __forceinline__ void fill_item_via(const LevelDev<T>& L, const T* __restrict__ Pq, int t, int homog,
                                              const V& Q) {
  if (t < L.nghost) {
    const GhostRec r = L.ghost[t];
    T v;
    if (r.kind == G_PCORNER) {
      const T a = ghost_value_via<T>(L.ghost[r.s[0]], Q, Pq, homog);
      const T b = ghost_value_via<T>(L.ghost[r.s[1]], Q, Pq, homog);
      v = (a + b) - Q(r.s[2]);
    } else {
      v = ghost_value_via<T>(r, Q, Pq, homog);
      if (L.aux && is_cf_kind(r.kind)) L.aux[t].cf = cf_part<T>(r, Pq);
    }
    L.Q[r.dst] = v;
  } else if (Pq && t < L.nghost + L.nring_cf) {
    const RingRec rr = L.ring_cf[t - L.nghost];
    L.QC[rr.dst] = Pq[rr.src];
  }
}

template <typename T>
__device__ __forceinline__ int fill_items(const LevelDev<T>& L, const T* Pq) {
  return L.nghost + (Pq ? L.nring_cf : 0);
}

// ---------------------------------------------------------------- smoother
// relax pass modes: kInlRecords / kInlArith / kKeepCf / kPush (bsam_kernels.h)

// parent part of a face cell's coarse-fine record, for the passes that
// assemble the value from neighbours
template <typename T>
__device__ __forceinline__ void keep_face_cf(const LevelDev<T>& L, const T* __restrict__ Pq, int idx) {
  const GhostRec r = L.ghost[idx];
  if (is_cf_kind(r.kind)) L.aux[idx].cf = cf_part<T>(r, Pq);
}

template <typename T>
__device__ __forceinline__ T record_ghost(const LevelDev<T>& L, const T* __restrict__ Q, const T* __restrict__ Pq,
                                          int idx, int homog, int keep) {
  const GhostRec r = L.ghost[idx];
  if (keep && is_cf_kind(r.kind)) L.aux[idx].cf = cf_part<T>(r, Pq);
  return ghost_value<T>(r, Q, Pq, homog);
}

// Coarse-fine ghost of the face cell x at 1-based position u along a face of
// length len, from x's own value, its inward neighbour, its neighbours along
// the face (am: u-1, ap: u+1), the cell two inward-along steps away (diag: at
// u+2 for even u, u-2 for odd u) and the parent part cf.  These are exactly
// the cells the interpolation record names (ebInterpCf.m:45-121), in the
// record's operation order.
template <typename T>
__device__ __forceinline__ T cf_face_ghost(int kind, int u, T self, T inward, T am, T ap, T diag, T cf) {
  const bool even = (u & 1) == 0;
  switch (kind) {
    case G_QIF:
      return (T(A1_) * diag + T(A2_) * (even ? ap : am)) + cf;
    case G_QIF_END:
      return (T(A1_) * inward + T(A2_) * self) + cf;
    default:  // G_LIF
      return (cf + T(C23_) * self) - T(C13_) * (even ? am : ap);
  }
}

// along-face offset (in cells) of the diagonal interpolation source
__device__ __forceinline__ int cf_diag_step(int u, int len) {
  return (u == 1 || u == len) ? 0 : ((u & 1) ? -2 : 2);
}

// a tile record as one block of wide loads
__device__ __forceinline__ TileRec load_tile(const TileRec* tiles, int idx) {
  TileRec r;
  __builtin_memcpy(&r, tiles + idx, sizeof(TileRec));
  return r;
}

// the tile fields a pass reads, requested together (see relax_args_early):
// one round trip for the record
__device__ __forceinline__ void tile_fields_early(const TileRec& t) {
  const PatchInfo& p = t.pi;
  asm volatile("" ::"s"(p.mx), "s"(p.my), "s"(p.par0), "s"(t.start), "s"(p.magic), "s"(p.gbase), "s"(p.qoff),
               "s"(p.coff), "s"(p.xoff), "s"(p.yoff), "s"(p.goff));
  asm volatile("" ::"s"(p.bc[0]), "s"(p.bc[1]), "s"(p.bc[2]), "s"(p.bc[3]));
}

// One cell of a colour pass (ebKernels.m:37-95).  Inside one relax call the
// ghost a colour-c cell reads depends only on other-colour cells, on the cell
// itself (read before it is written) and on the parent level, so every ghost
// source mode yields what ebFillGhosts would have written after the previous
// pass.  Every load the cell needs is issued before any value is used, so a
// pass costs one memory round trip after the tile record (two in record
// mode, whose sources come from the records).
// The update of ebKernels.m:37-95 from resolved neighbour values (ghosts
// already substituted), in the reference's operation order.
template <typename T>
__device__ __forceinline__ T relax_combine(const LevelDev<T>& L, T qe, T qw, T qn, T qs, T f, T de, T dw, T dn,
                                           T ds, T id, T self, T omega) {
  const T hx = (T)L.h[0], hy = (T)L.h[1];
  const T hx2 = hx * hx, hy2 = hy * hy;
  const T ihx2 = (T)L.ih2[0], ihy2 = (T)L.ih2[1];
  const int p2 = L.pow2;
  T tmp = f + div_h2<T>(qe * de, hx2, ihx2, p2);
  tmp = tmp + div_h2<T>(qw * dw, hx2, ihx2, p2);
  tmp = tmp + div_h2<T>(qn * dn, hy2, ihy2, p2);
  tmp = tmp + div_h2<T>(qs * ds, hy2, ihy2, p2);
  tmp = tmp * id;
  return omega == T(1) ? tmp : omega * tmp + (T(1) - omega) * self;
}

// relax_combine with the level's power-of-two spacing (P2: q / h^2 taken as
// q * (1 / h^2), the same number) and omega == 1 (W1) fixed at compile time,
// so the region passes carry no per-update branches for them
template <typename T, bool P2, bool W1>
__device__ __forceinline__ T relax_combine_s(const LevelDev<T>& L, T qe, T qw, T qn, T qs, T f, T de, T dw, T dn,
                                             T ds, T id, T self, T omega) {
  const T hx = (T)L.h[0], hy = (T)L.h[1];
  const T hx2 = hx * hx, hy2 = hy * hy;
  const T ihx2 = (T)L.ih2[0], ihy2 = (T)L.ih2[1];
  T tmp = f + (P2 ? (qe * de) * ihx2 : (qe * de) / hx2);
  tmp = tmp + (P2 ? (qw * dw) * ihx2 : (qw * dw) / hx2);
  tmp = tmp + (P2 ? (qn * dn) * ihy2 : (qn * dn) / hy2);
  tmp = tmp + (P2 ? (qs * ds) * ihy2 : (qs * ds) / hy2);
  tmp = tmp * id;
  if (W1) return tmp;
  return omega * tmp + (T(1) - omega) * self;
}

// Neighbour values of cell (i, j) with its ghosts substituted: partial
// physical ghosts (Dirichlet (8/3)g + (1/3)u_2, Neumann h g; g: 0 when
// homogeneous) and, when assembled, coarse-fine ghosts from the cell's own
// value, its neighbours, the diagonal sources x[] and the face records a[]
// (sibling ghosts keep the frame value).
template <typename T>
__device__ __forceinline__ void resolve_ghosts(const LevelDev<T>& L, const PatchInfo& pi, int i, int j, T re, T rw,
                                               T rn, T rs, T self, const T* g, const T* x, const FaceAux<T>* a,
                                               bool assembled, T& qe, T& qw, T& qn, T& qs) {
  const int mx = pi.mx, my = pi.my;
  const T hx = (T)L.h[0], hy = (T)L.h[1];
  qe = re; qw = rw; qn = rn; qs = rs;
  if (i == 0 && pi.bc[0]) qw = pi.bc[0] == 1 ? T(C83_) * g[0] + T(C13_) * re : hx * g[0];
  if (i == mx - 1 && pi.bc[1]) qe = pi.bc[1] == 1 ? T(C83_) * g[1] + T(C13_) * rw : hx * g[1];
  if (j == 0 && pi.bc[2]) qs = pi.bc[2] == 1 ? T(C83_) * g[2] + T(C13_) * rn : hy * g[2];
  if (j == my - 1 && pi.bc[3]) qn = pi.bc[3] == 1 ? T(C83_) * g[3] + T(C13_) * rs : hy * g[3];
  if (assembled) {
    if (i == 0 && !pi.bc[0] && a[0].kind != G_COPY) qw = cf_face_ghost<T>(a[0].kind, j + 1, self, re, rs, rn, x[0], a[0].cf);
    if (i == mx - 1 && !pi.bc[1] && a[1].kind != G_COPY)
      qe = cf_face_ghost<T>(a[1].kind, j + 1, self, rw, rs, rn, x[1], a[1].cf);
    if (j == 0 && !pi.bc[2] && a[2].kind != G_COPY) qs = cf_face_ghost<T>(a[2].kind, i + 1, self, rn, rw, re, x[2], a[2].cf);
    if (j == my - 1 && !pi.bc[3] && a[3].kind != G_COPY)
      qn = cf_face_ghost<T>(a[3].kind, i + 1, self, rs, rw, re, x[3], a[3].cf);
  }
}

// Every load of one cell for a frame or assembled update, issued at in-bounds
// addresses whatever the cell's position, so that several cells' loads are in
// flight at once.  A cell lies on at most one x-face and one y-face (patches
// are at least 2 wide), so one boundary value, face record and diagonal
// source per axis suffice: off the faces they stand in for something never
// read.
template <typename T>
struct CellIn {
  T re, rw, rn, rs, self, f, de, dw, dn, ds, id;
  T gx, gy, xx, xy;
  FaceAux<T> ax, ay;
};

template <typename T>
__device__ __forceinline__ void cell_load(const LevelDev<T>& L, const T* __restrict__ Q, const PatchInfo& pi, int i,
                                          int j, bool assembled, CellIn<T>& c) {
  const int mx = pi.mx, my = pi.my, ldq = mx + 2;
  const int qc = pi.qoff + (i + 1) + (j + 1) * ldq;
  const int cc = pi.coff + i + j * mx;
  const int fx = pi.xoff + i + j * (mx + 1);
  const int fy = pi.yoff + i + j * mx;
  const bool ee = i == mx - 1, nn = j == my - 1;
  c.re = Q[qc + 1]; c.rw = Q[qc - 1]; c.rn = Q[qc + ldq]; c.rs = Q[qc - ldq]; c.self = Q[qc];
  c.f = L.F[cc]; c.id = L.ID[cc];
  c.de = L.DeW[fx + 1]; c.dw = L.DeW[fx]; c.dn = L.DnS[fy + mx]; c.ds = L.DnS[fy];
  c.gx = L.G[pi.goff + (ee ? my : 0) + j];
  c.gy = L.G[pi.goff + 2 * my + (nn ? mx : 0) + i];
  c.ax = L.aux[pi.gbase + (ee ? my : 0) + j];
  c.ay = L.aux[pi.gbase + 2 * my + (nn ? mx : 0) + i];
  if (assembled) {
    const bool onx = i == 0 || ee, ony = j == 0 || nn;
    c.xx = Q[qc + (ee ? -1 : 1) + (onx ? cf_diag_step(j + 1, my) : 0) * ldq];
    c.xy = Q[qc + (nn ? -ldq : ldq) + (ony ? cf_diag_step(i + 1, mx) : 0)];
  } else {
    c.xx = c.xy = T(0);
  }
}

template <typename T>
__device__ __forceinline__ T cell_relax(const LevelDev<T>& L, const PatchInfo& pi, int i, int j, const CellIn<T>& c,
                                        int homog, bool assembled, T omega) {
  const T gx = homog ? T(0) : c.gx, gy = homog ? T(0) : c.gy;
  const T g[4] = {gx, gx, gy, gy}, x[4] = {c.xx, c.xx, c.xy, c.xy};
  const FaceAux<T> a[4] = {c.ax, c.ax, c.ay, c.ay};
  T qe, qw, qn, qs;
  resolve_ghosts<T>(L, pi, i, j, c.re, c.rw, c.rn, c.rs, c.self, g, x, a, assembled, qe, qw, qn, qs);
  return relax_combine<T>(L, qe, qw, qn, qs, c.f, c.de, c.dw, c.dn, c.ds, c.id, c.self, omega);
}

// EARLY (multi-workgroup passes): boundary data are loaded for every cell of
// a patch with that physical side, with the other loads, instead of behind
// the boundary-cell branch (latency over instruction count; the workgroup
// passes on LDS-staged levels prefer the opposite).
template <typename T, bool EARLY = true>
__device__ __forceinline__ T relax_value(const LevelDev<T>& L, const T* __restrict__ Q, const T* __restrict__ Pq,
                                         const PatchInfo& pi, int i, int j, T omega, int homog, int mode,
                                         FaceAux<T>* aux_out = nullptr) {
  const int mx = pi.mx, my = pi.my;
  const int ldq = mx + 2;
  const int qc = pi.qoff + (i + 1) + (j + 1) * ldq;
  const int cc = pi.coff + i + j * mx;
  const int fx = pi.xoff + i + j * (mx + 1);
  const int fy = pi.yoff + i + j * mx;
  const int src = mode & 3;
  // internal faces this cell lies on, and their records
  const bool fw = i == 0 && !pi.bc[0], fe = i == mx - 1 && !pi.bc[1];
  const bool fs = j == 0 && !pi.bc[2], fn = j == my - 1 && !pi.bc[3];
  const int iw = pi.gbase + j, ie = pi.gbase + my + j;
  const int is = pi.gbase + 2 * my + i, in_ = pi.gbase + 2 * my + mx + i;

  const T re = Q[qc + 1], rw = Q[qc - 1], rn = Q[qc + ldq], rs = Q[qc - ldq];
  const T f = L.F[cc], de = L.DeW[fx + 1], dw = L.DeW[fx], dn = L.DnS[fy + mx], ds = L.DnS[fy];
  const T id = L.ID[cc];
  const T self = (omega != T(1) || src == kInlArith) ? Q[qc] : T(0);
  // boundary data: every patch stores all four sides, so these are in bounds
  T gw = T(0), ge = T(0), gs = T(0), gn = T(0);
  if (!homog) {
    if (pi.bc[0] && (EARLY || i == 0)) gw = L.G[pi.goff + j];
    if (pi.bc[1] && (EARLY || i == mx - 1)) ge = L.G[pi.goff + my + j];
    if (pi.bc[2] && (EARLY || j == 0)) gs = L.G[pi.goff + 2 * my + i];
    if (pi.bc[3] && (EARLY || j == my - 1)) gn = L.G[pi.goff + 2 * my + mx + i];
  }
  // face records of the assembled and pushing passes, and the diagonal
  // interpolation sources
  FaceAux<T> aw{}, ae{}, as{}, an{};
  if (src == kInlArith || (mode & kPush)) {
    if (fw) aw = L.aux[iw];
    if (fe) ae = L.aux[ie];
    if (fs) as = L.aux[is];
    if (fn) an = L.aux[in_];
  }
  T xw = T(0), xe = T(0), xs = T(0), xn = T(0);
  if (src == kInlArith) {
    if (fw) xw = Q[qc + 1 + cf_diag_step(j + 1, my) * ldq];
    if (fe) xe = Q[qc - 1 + cf_diag_step(j + 1, my) * ldq];
    if (fs) xs = Q[qc + ldq + cf_diag_step(i + 1, mx)];
    if (fn) xn = Q[qc - ldq + cf_diag_step(i + 1, mx)];
  }

  T qe, qw, qn, qs;
  const T gv[4] = {gw, ge, gs, gn}, xv[4] = {xw, xe, xs, xn};
  const FaceAux<T> av[4] = {aw, ae, as, an};
  resolve_ghosts<T>(L, pi, i, j, re, rw, rn, rs, self, gv, xv, av, src == kInlArith, qe, qw, qn, qs);
  if (src == kInlRecords) {
    const int keep = mode & kKeepCf;
    if (fw) qw = record_ghost<T>(L, Q, Pq, iw, homog, keep);
    if (fe) qe = record_ghost<T>(L, Q, Pq, ie, homog, keep);
    if (fs) qs = record_ghost<T>(L, Q, Pq, is, homog, keep);
    if (fn) qn = record_ghost<T>(L, Q, Pq, in_, homog, keep);
  }
  const T tmp = relax_combine<T>(L, qe, qw, qn, qs, f, de, dw, dn, ds, id, self, omega);
  if (aux_out) {  // the push targets, for a caller that stores the value
    aux_out[0] = aw; aux_out[1] = ae; aux_out[2] = as; aux_out[3] = an;
  }
  return tmp;
}

// one cell of a colour pass in place on L.Q, with the sibling pushes
template <typename T, bool EARLY = true>
__device__ __forceinline__ void relax_cell(const LevelDev<T>& L, const T* __restrict__ Pq, const PatchInfo& pi,
                                           int i, int j, T omega, int homog, int mode) {
  T* __restrict__ Q = L.Q;
  FaceAux<T> a[4];
  const T y = relax_value<T, EARLY>(L, Q, Pq, pi, i, j, omega, homog, mode, (mode & kPush) ? a : nullptr);
  const int qc = pi.qoff + (i + 1) + (j + 1) * (pi.mx + 2);
  Q[qc] = y;
  if (mode & kPush) {
    // sibling ghosts that copy this cell; their readers are other-colour
    // cells, which this pass does not visit
    const bool fw = i == 0 && !pi.bc[0], fe = i == pi.mx - 1 && !pi.bc[1];
    const bool fs = j == 0 && !pi.bc[2], fn = j == pi.my - 1 && !pi.bc[3];
    if (fw && a[0].push >= 0) Q[a[0].push] = y;
    if (fe && a[1].push >= 0) Q[a[1].push] = y;
    if (fs && a[2].push >= 0) Q[a[2].push] = y;
    if (fn && a[3].push >= 0) Q[a[3].push] = y;
  }
}

// one lane of a colour pass over a 256-cell tile (tile: uniform across the
// wave); in a pass that keeps parent parts, the other colour's face cells
// store those of their records for the pass that follows
template <typename T, bool EARLY = true>
__device__ __forceinline__ void relax_item(const LevelDev<T>& L, const T* __restrict__ Pq, int tile,
                                           int lane, int colour, T omega, int homog, int mode) {
  const TileRec tr = load_tile(L.tiles, __builtin_amdgcn_readfirstlane(tile));
  tile_fields_early(tr);
  const PatchInfo& pi = tr.pi;
  const int mx = pi.mx, my = pi.my;
  const int c = tr.start + lane;
  if (c >= mx * my) return;
  const int j = cell_row(c, mx, pi.magic);
  const int i = c - j * mx;
  // (li + lj + par0) mod 2 == colour-1 with 1-based local indices
  if (((i + j + pi.par0) & 1) != colour - 1) {
    if (mode & kKeepCf) {
      if (i == 0 && !pi.bc[0]) keep_face_cf<T>(L, Pq, pi.gbase + j);
      if (i == mx - 1 && !pi.bc[1]) keep_face_cf<T>(L, Pq, pi.gbase + my + j);
      if (j == 0 && !pi.bc[2]) keep_face_cf<T>(L, Pq, pi.gbase + 2 * my + i);
      if (j == my - 1 && !pi.bc[3]) keep_face_cf<T>(L, Pq, pi.gbase + 2 * my + mx + i);
    }
    return;
  }
  relax_cell<T, EARLY>(L, Pq, pi, i, j, omega, homog, mode);
}

// Item t of a colour pass over colour-c cells only: the t-th cell of colour c
// in tile order.  A tile of 256 cells of a patch with even mx holds 128 cells
// of each colour (every row has mx/2 of them and tiles start on even cells),
// so half the lanes cover a pass.  The 64 items of a wave (t aligned to 64)
// lie in one tile, whose record is therefore read once per wave.
template <typename T, bool EARLY = true>
__device__ __forceinline__ void relax_item_colour(const LevelDev<T>& L, const T* __restrict__ Pq, int t,
                                                  int colour, T omega, int homog, int inl) {
  const TileRec tr = load_tile(L.tiles, __builtin_amdgcn_readfirstlane(t >> 7));
  tile_fields_early(tr);
  const PatchInfo& pi = tr.pi;
  const int mx = pi.mx;
  const int a = (tr.start >> 1) + (t & 127);  // index among the patch's colour-c cells
  if (2 * a >= mx * pi.my) return;
  const int j = cell_row(2 * a, mx, pi.magic);  // a / (mx/2)
  const int i = 2 * (a - j * (mx >> 1)) + ((colour - 1 + j + pi.par0) & 1);
  relax_cell<T, EARLY>(L, Pq, pi, i, j, omega, homog, inl);
}

// ---------------------------------------------------------------- fused sweep
// One red+black sweep of a tile per workgroup (levels with L.arith, every
// patch of even width).  The red half computes the tile's colour-1 cells and
// the tile's halo (colour-1 cells the colour-2 cells read: neighbours in
// other tiles or sibling patches, coarse-fine interpolation sources) with
// relax_value on Qin, exactly as a colour pass would, into an LDS region of
// the tile plus two rings.  After a barrier the black half updates the
// tile's colour-2 cells from the region (ghosts: sibling copies from the
// region, coarse-fine ghosts assembled from region cells and the parent part,
// partial physical ghosts from the region).  Results go to Qout, which no
// workgroup reads in this launch, and the tile's values are pushed into the
// sibling ghosts of Qout that copy them.  MODE is the red half's ghost source
// (frame / records / assembled); KEEP: the tile's face records' parent parts
// are computed from the parent and stored (their value for the black half
// too).
template <typename T>
__device__ __forceinline__ void sweep_cell_of(const SweepTile& st, int r, int colour, int& i, int& j) {
  // r-th cell of the colour in the tile (tile width even: tw/2 per row)
  const int half = st.tw >> 1;
  const int jj = r / half;
  const int ii = 2 * (r - jj * half) + ((colour - 1 + st.i0 + st.j0 + jj + st.pi.par0) & 1);
  i = st.i0 + ii;
  j = st.j0 + jj;
}

template <typename T>
__device__ __forceinline__ T keep_cf_value(const LevelDev<T>& L, const T* __restrict__ Pq, int idx) {
  const GhostRec r = L.ghost[idx];
  const T v = cf_part<T>(r, Pq);
  if (is_cf_kind(r.kind)) L.aux[idx].cf = v;
  return v;
}

template <typename T, int MODE, bool KEEP>
__global__ void __launch_bounds__(256) k_relax_sweep(LevelDev<T> L, const SweepTile* __restrict__ tiles,
                                                     const T* __restrict__ Qin, T* __restrict__ Qout,
                                                     const T* __restrict__ Pq, T omega, int homog) {
  __shared__ SweepTile st;
  __shared__ T R[kSweepRegion];
  {
    const uint4* __restrict__ src = reinterpret_cast<const uint4*>(tiles + blockIdx.x);
    uint4* dst = reinterpret_cast<uint4*>(&st);
    for (int x = threadIdx.x; x < (int)(sizeof(SweepTile) / 16); x += blockDim.x) dst[x] = src[x];
  }
  __syncthreads();
  const PatchInfo& P = st.pi;
  const int t = threadIdx.x;
  const int half = (st.tw * st.th) >> 1;
  const int rw = st.tw + 4;
  const int mx = P.mx, my = P.my, ldq = mx + 2;
  constexpr bool ASM = MODE == kInlArith;
  // ---- red half: two tile cells and one halo cell per lane, every load
  // issued before any value is used
  const bool has0 = t < half, has1 = t + 256 < half, hasH = t < st.nhalo;
  int ri[2], rj[2];
  sweep_cell_of<T>(st, has0 ? t : 0, 1, ri[0], rj[0]);
  sweep_cell_of<T>(st, has1 ? t + 256 : 0, 1, ri[1], rj[1]);
  const SweepHalo he = st.halo[hasH ? t : 0];
  const int hps = st.nhalo ? he.slot_ps >> 16 : 0;
  const PatchInfo& HP = hps ? st.sib[hps - 1] : P;
  const int hi = st.nhalo ? (he.ij & 0xffff) : 0, hj = st.nhalo ? (he.ij >> 16) : 0;
  T vr[3];
  FaceAux<T> rax[2], ray[2];
  if (MODE == kInlRecords) {  // ghosts from the records (the first sweep after a skipped fill)
    FaceAux<T> a4[4];
    if (has0) {
      vr[0] = relax_value<T, true>(L, Qin, Pq, P, ri[0], rj[0], omega, homog, MODE | kPush, a4);
      rax[0] = ri[0] == mx - 1 ? a4[1] : a4[0]; ray[0] = rj[0] == my - 1 ? a4[3] : a4[2];
    }
    if (has1) {
      vr[1] = relax_value<T, true>(L, Qin, Pq, P, ri[1], rj[1], omega, homog, MODE | kPush, a4);
      rax[1] = ri[1] == mx - 1 ? a4[1] : a4[0]; ray[1] = rj[1] == my - 1 ? a4[3] : a4[2];
    }
    if (hasH) vr[2] = relax_value<T, true>(L, Qin, Pq, HP, hi, hj, omega, homog, MODE);
  } else {
    // branch-free: a lane without an item loads a stand-in cell (the
    // tile's first), so that all three groups issue in one block
    CellIn<T> c0, c1, ch;
    cell_load<T>(L, Qin, P, ri[0], rj[0], ASM, c0);
    cell_load<T>(L, Qin, P, ri[1], rj[1], ASM, c1);
    cell_load<T>(L, Qin, HP, hi, hj, ASM, ch);
    if (has0) { vr[0] = cell_relax<T>(L, P, ri[0], rj[0], c0, homog, ASM, omega); rax[0] = c0.ax; ray[0] = c0.ay; }
    if (has1) { vr[1] = cell_relax<T>(L, P, ri[1], rj[1], c1, homog, ASM, omega); rax[1] = c1.ax; ray[1] = c1.ay; }
    if (hasH) vr[2] = cell_relax<T>(L, HP, hi, hj, ch, homog, ASM, omega);
  }
  if (KEEP) {  // the red tile cells' records, for the sweeps that follow
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (!(k ? has1 : has0)) continue;
      const int i = ri[k], j = rj[k];
      if (i == 0 && !P.bc[0]) keep_face_cf<T>(L, Pq, P.gbase + j);
      if (i == mx - 1 && !P.bc[1]) keep_face_cf<T>(L, Pq, P.gbase + my + j);
      if (j == 0 && !P.bc[2]) keep_face_cf<T>(L, Pq, P.gbase + 2 * my + i);
      if (j == my - 1 && !P.bc[3]) keep_face_cf<T>(L, Pq, P.gbase + 2 * my + mx + i);
    }
  }
  if (has0) R[(ri[0] - st.i0 + 2) + (rj[0] - st.j0 + 2) * rw] = vr[0];
  if (has1) R[(ri[1] - st.i0 + 2) + (rj[1] - st.j0 + 2) * rw] = vr[1];
  if (hasH) R[he.slot_ps & 0xffff] = vr[2];
  // ---- the black cells' own data (their neighbours come from the region)
  int bi[2], bj[2];
  sweep_cell_of<T>(st, has0 ? t : 0, 2, bi[0], bj[0]);
  sweep_cell_of<T>(st, has1 ? t + 256 : 0, 2, bi[1], bj[1]);
  CellIn<T> b0, b1;
  cell_load<T>(L, Qin, P, bi[0], bj[0], false, b0);
  cell_load<T>(L, Qin, P, bi[1], bj[1], false, b1);
  if (KEEP) {  // their coarse-fine records' parent parts, from the parent
    if (has0) {
      if ((bi[0] == 0 || bi[0] == mx - 1) && !P.bc[bi[0] == 0 ? 0 : 1])
        b0.ax.cf = keep_cf_value<T>(L, Pq, P.gbase + (bi[0] == mx - 1 ? my : 0) + bj[0]);
      if ((bj[0] == 0 || bj[0] == my - 1) && !P.bc[bj[0] == 0 ? 2 : 3])
        b0.ay.cf = keep_cf_value<T>(L, Pq, P.gbase + 2 * my + (bj[0] == my - 1 ? mx : 0) + bi[0]);
    }
    if (has1) {
      if ((bi[1] == 0 || bi[1] == mx - 1) && !P.bc[bi[1] == 0 ? 0 : 1])
        b1.ax.cf = keep_cf_value<T>(L, Pq, P.gbase + (bi[1] == mx - 1 ? my : 0) + bj[1]);
      if ((bj[1] == 0 || bj[1] == my - 1) && !P.bc[bj[1] == 0 ? 2 : 3])
        b1.ay.cf = keep_cf_value<T>(L, Pq, P.gbase + 2 * my + (bj[1] == my - 1 ? mx : 0) + bi[1]);
    }
  }
  __syncthreads();
  // ---- black half: neighbours and diagonal sources from the region
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (!(k ? has1 : has0)) continue;
    CellIn<T> c = k ? b1 : b0;
    const int i = bi[k], j = bj[k];
    const int sl = (i - st.i0 + 2) + (j - st.j0 + 2) * rw;
    c.re = R[sl + 1]; c.rw = R[sl - 1]; c.rn = R[sl + rw]; c.rs = R[sl - rw];
    const bool ee = i == mx - 1, nn = j == my - 1;
    const bool onx = i == 0 || ee, ony = j == 0 || nn;
    c.xx = R[sl + (ee ? -1 : 1) + (onx ? cf_diag_step(j + 1, my) : 0) * rw];
    c.xy = R[sl + (nn ? -rw : rw) + (ony ? cf_diag_step(i + 1, mx) : 0)];
    const T y = cell_relax<T>(L, P, i, j, c, homog, true, omega);
    Qout[P.qoff + (i + 1) + (j + 1) * ldq] = y;
    if (onx && !P.bc[ee ? 1 : 0] && c.ax.push >= 0) Qout[c.ax.push] = y;
    if (ony && !P.bc[nn ? 3 : 2] && c.ay.push >= 0) Qout[c.ay.push] = y;
  }
  // ---- the red tile cells to Qout, with their pushes
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (!(k ? has1 : has0)) continue;
    const int i = ri[k], j = rj[k];
    const T y = vr[k];
    Qout[P.qoff + (i + 1) + (j + 1) * ldq] = y;
    const bool ee = i == mx - 1, nn = j == my - 1;
    if ((i == 0 || ee) && !P.bc[ee ? 1 : 0] && rax[k].push >= 0) Qout[rax[k].push] = y;
    if ((j == 0 || nn) && !P.bc[nn ? 3 : 2] && ray[k].push >= 0) Qout[ray[k].push] = y;
  }
}

// ---------------------------------------------------------------- region relax
// A whole relax call (ebRelaxLevel.m:28-38: nsweeps x two colour passes, the
// ghosts refreshed after each) in one launch per level.  A workgroup stages
// its region tile in an LDS canvas in the level's index space: neighbours in
// the same patch and sibling cells across copy faces are then the adjacent
// canvas slots, and a coarse-fine ghost is assembled from the canvas and the
// parent part exactly as ebInterpCf's record would after ebFillGhosts.  Every
// cell the workgroup updates keeps its coefficients in registers for the
// whole call; pass p updates, in place, the cells of its colour whose depth
// is at least p (they read only cells of the other colour and themselves, so
// the in-place update is race free), with a barrier between passes.  The
// tile's own cells go to Qout, a second copy of Q that no workgroup reads in
// this launch; launch_fill_copy then moves them into Q while it refreshes the
// ghosts, as the fill after the last pass.
// Pass 1 of a call whose ghost frame is what the reference's first pass reads
// (frame = 1) takes the internal-face ghosts from the frame; otherwise, and in
// every later pass, they come from the canvas and the parent.
enum RegionFace : int { RF_NONE = 0, RF_DIR = 1, RF_NEU = 2, RF_QIF = 3, RF_QIF_END = 4, RF_LIF = 5, RF_COPY = 6 };

template <typename T>
struct RegionCell {
  T f, id, de, dw, dn, ds;
  T bx, by;   // x / y face: boundary data (physical side) or parent part (coarse-fine)
  int slot;   // canvas slot, -1: no cell
  int fl;     // x face: kind (bits 0-2), east (3), even position (4); y face: kind (5-7), north (8),
              // even (9); depth (bits 12-15)
  int dxy;    // canvas offsets of the diagonal interpolation sources: x face low 16 bits, y face high 16
};

__device__ __forceinline__ int region_cf_kind(int kind) {
  return kind == G_COPY ? RF_COPY : kind == G_QIF ? RF_QIF : kind == G_QIF_END ? RF_QIF_END : RF_LIF;
}
__device__ __forceinline__ int region_ghost_kind(int rf) {
  return rf == RF_QIF ? G_QIF : rf == RF_QIF_END ? G_QIF_END : G_LIF;
}

// Issues the loads of one updated cell straight into its state: the value
// (into q), the coefficients and, for a cell on a face, the boundary datum or
// the face record's kind and parent part and (with frame) the frame's ghost
// value.  Nothing loaded is used here, so the loads of all of a lane's cells
// are in flight together.  kx / ky: the internal faces' record kinds (-1: a
// physical side or no face), resolved by region_kinds once the loads land.
template <typename T>
__device__ __forceinline__ void region_load(const LevelDev<T>& L, const RegionPatch* sp, int cw, unsigned code,
                                            int homog, bool frame, RegionCell<T>& c, T& q, int& qc, T& frx,
                                            T& fry, int& kx, int& ky) {
  const unsigned cd = code == kRegionNone ? 0u : code;  // a stand-in cell for an empty entry
  const int ps = cd & 15, i = (cd >> 4) & 4095, j = (cd >> 16) & 4095;
  const PatchInfo& P = sp[ps].pi;
  const int mx = P.mx, my = P.my, ldq = mx + 2;
  qc = P.qoff + (i + 1) + (j + 1) * ldq;
  const int cc = P.coff + i + j * mx, fx = P.xoff + i + j * (mx + 1), fy = P.yoff + i + j * mx;
  c.slot = code == kRegionNone ? -1 : (sp[ps].ox + i) + (sp[ps].oy + j) * cw;
  q = L.Q[qc];
  c.f = L.F[cc];
  c.id = L.ID[cc];
  c.dw = L.DeW[fx];
  c.de = L.DeW[fx + 1];
  c.ds = L.DnS[fy];
  c.dn = L.DnS[fy + mx];
  const bool e = i == mx - 1, n = j == my - 1;
  int fl = (int)((cd >> 28) << 12), dx = 0, dy = 0;
  kx = ky = -1;
  c.bx = c.by = T(0);
  frx = fry = T(0);
  if (i == 0 || e) {
    const int side = e ? 1 : 0, bc = P.bc[side];
    if (bc) {
      fl |= bc == 1 ? RF_DIR : RF_NEU;
      if (!homog) c.bx = L.G[P.goff + (e ? my : 0) + j];
    } else {
      const FaceAux<T>* ax = L.aux + P.gbase + (e ? my : 0) + j;
      kx = ax->kind;
      c.bx = ax->cf;
      if (frame) frx = L.Q[qc + (e ? 1 : -1)];
      dx = (e ? -1 : 1) + cf_diag_step(j + 1, my) * cw;
    }
    fl |= (e ? 8 : 0) | (((j + 1) & 1) == 0 ? 16 : 0);
  }
  if (j == 0 || n) {
    const int side = n ? 3 : 2, bc = P.bc[side];
    if (bc) {
      fl |= (bc == 1 ? RF_DIR : RF_NEU) << 5;
      if (!homog) c.by = L.G[P.goff + 2 * my + (n ? mx : 0) + i];
    } else {
      const FaceAux<T>* ay = L.aux + P.gbase + 2 * my + (n ? mx : 0) + i;
      ky = ay->kind;
      c.by = ay->cf;
      if (frame) fry = L.Q[qc + (n ? ldq : -ldq)];
      dy = (n ? -cw : cw) + cf_diag_step(i + 1, mx);
    }
    fl |= (n ? 256 : 0) | (((i + 1) & 1) == 0 ? 512 : 0);
  }
  c.fl = fl;
  c.dxy = (dx & 0xffff) | (dy << 16);
}

// the internal faces' kinds, after the loads
template <typename T>
__device__ __forceinline__ void region_kinds(RegionCell<T>& c, int kx, int ky) {
  if (kx >= 0) c.fl |= region_cf_kind(kx);
  if (ky >= 0) c.fl |= region_cf_kind(ky) << 5;
}

__device__ __forceinline__ int region_depth(int fl) { return (fl >> 12) & 15; }
__device__ __forceinline__ bool region_face(int fl) { return (fl & 0x3ff) != 0; }

// one cell's update from the canvas: the expressions of resolve_ghosts and
// relax_combine, in the same operation order; fr* are the frame's ghost
// values (pass 1 of a call that reads the frame)
template <typename T, bool P2, bool W1>
__device__ __forceinline__ T region_update(const LevelDev<T>& L, const T* cv, int cw, const RegionCell<T>& c,
                                           bool frame, T frx, T fry, T omega) {
  const int s = c.slot;
  const T self = cv[s], re = cv[s + 1], rw = cv[s - 1], rn = cv[s + cw], rs = cv[s - cw];
  if (!region_face(c.fl))  // off every face (the host lists these cells first, so whole waves take this path)
    return relax_combine_s<T, P2, W1>(L, re, rw, rn, rs, c.f, c.de, c.dw, c.dn, c.ds, c.id, self, omega);
  T qe = re, qw = rw, qn = rn, qs = rs;
  const int xk = c.fl & 7, yk = (c.fl >> 5) & 7;
  if (xk) {
    const bool e = (c.fl >> 3) & 1;
    const T in = e ? rw : re;
    T v = e ? re : rw;
    if (xk == RF_DIR) {
      v = T(C83_) * c.bx + T(C13_) * in;
    } else if (xk == RF_NEU) {
      v = (T)L.h[0] * c.bx;
    } else if (frame) {
      v = frx;
    } else if (xk != RF_COPY) {
      const T diag = cv[s + (int)(short)(c.dxy & 0xffff)];
      v = cf_face_ghost<T>(region_ghost_kind(xk), ((c.fl >> 4) & 1) ? 2 : 1, self, in, rs, rn, diag, c.bx);
    }
    if (e) qe = v; else qw = v;
  }
  if (yk) {
    const bool n = (c.fl >> 8) & 1;
    const T in = n ? rs : rn;
    T v = n ? rn : rs;
    if (yk == RF_DIR) {
      v = T(C83_) * c.by + T(C13_) * in;
    } else if (yk == RF_NEU) {
      v = (T)L.h[1] * c.by;
    } else if (frame) {
      v = fry;
    } else if (yk != RF_COPY) {
      const T diag = cv[s + (int)(short)(c.dxy >> 16)];
      v = cf_face_ghost<T>(region_ghost_kind(yk), ((c.fl >> 9) & 1) ? 2 : 1, self, in, rw, re, diag, c.by);
    }
    if (n) qn = v; else qs = v;
  }
  return relax_combine_s<T, P2, W1>(L, qe, qw, qn, qs, c.f, c.de, c.dw, c.dn, c.ds, c.id, self, omega);
}

// npass = 2 x sweeps (even, <= L.rdepth).  Workgroup h takes tile
// (h mod 8) * per + h / 8 (workgroups are dealt to the 8 XCDs in turn, so
// each XCD runs a contiguous, spatially ordered range of tiles and shares
// their halo lines in its L2); the grid is padded to 8 * per.  The header,
// the patch records and a lane's entries sit at fixed offsets of the tile, so
// they arrive in one round trip, and the cells' data in a second: the parent
// parts of coarse-fine faces come from the face records' aux entries, which
// must hold those of the current parent (a fill, a preparing correction or
// launch_refresh_cf wrote them after the parent last changed).
// the epilogue's data of one ghost record, staged in LDS during the prologue
template <typename T>
struct GhostOp {
  T c, cf, ca, cfa, cb, cfb;
  int dst, kinds;
  unsigned s01, s23, s4;
};

static_assert(sizeof(GhostOp<double>) <= 80, "the host's LDS bound of a region launch assumes 80-byte ghost records");

// A ghost record's value from its kind, boundary value c, fine sources v0,
// v1 and parent part cf: the expressions of ghost_value, in the same order
template <typename T>
__device__ __forceinline__ T region_ghost(int kind, T c, T cf, T v0, T v1, int homog) {
  switch (kind) {
    case G_COPY:
      return v0;
    case G_QIF:
    case G_QIF_END:
      return (T(A1_) * v0 + T(A2_) * v1) + cf;
    case G_LIF:
      return (cf + T(C23_) * v0) - T(C13_) * v1;
    case G_LIF_CORNER:
      return cf + T(C13_) * v0;
    case G_DIR: {
      const T g = homog ? T(0) : c;
      return (T(C83_) * g - T(2) * v0) + T(C13_) * v1;
    }
    case G_NEU: {
      const T hg = homog ? T(0) : c;
      return v0 + hg;
    }
    default:
      return T(0);
  }
}

// The coarse-grid correction of cell (i, j) of a region patch from the staged
// parent window (Wc: parent values, Cc: QC values): ebTransfer.m:57-71 with
// prolongBilin's (9, 3, 3, 1) / 16 weights, in correct_item's operation order
template <typename T>
__device__ __forceinline__ T region_correction(const RegionPatch& R, int i, int j, const T* Wc, const T* Cc) {
  const int wx = (i >> 1) + 1, wy = (j >> 1) + 1;
  const int s = R.eoff + (wx - R.ex0) + (wy - R.ey0) * R.ew;
  const int sx = s + ((i & 1) ? 1 : -1), sy = s + ((j & 1) ? R.ew : -R.ew), sd = sy + ((i & 1) ? 1 : -1);
  const T ec = Wc[s] - Cc[s], ex = Wc[sx] - Cc[sx], ey = Wc[sy] - Cc[sy], ed = Wc[sd] - Cc[sd];
  return (((T(9) * ec + T(3) * ex) + T(3) * ey) + ed) / T(16);
}

// The parent part (cf_part) of a coarse-fine record from parent values at
// window slots s0..s2 of the staged window
template <typename T>
__device__ __forceinline__ T region_cf_parts(int kind, T w0, T w1, T w2) {
  switch (kind) {
    case G_QIF:
      return T(A3_) * w0;
    case G_QIF_END:
      return T(A3_) * ((T(B1_) * w0 + T(B2_) * w1) + T(B3_) * w2);
    case G_LIF:
    case G_LIF_CORNER:
      return T(C23_) * w0;
    default:
      return T(0);
  }
}

// The parent part of the coarse-fine record on side `side` (W E S N) at
// position t of a region patch, from the staged window: cf_rec's stencils
// (the host checked them against every record's parent indices)
template <typename T>
__device__ __forceinline__ T region_cf_window(int kind, int side, int t, const RegionPatch& R, const T* Wc) {
  const int mx = R.pi.mx, my = R.pi.my, cmx = mx >> 1, cmy = my >> 1;
  const int len = side < 2 ? my : mx, clen = side < 2 ? cmy : cmx, u = t + 1;
  const int fixed = side == 0 ? 0 : side == 1 ? cmx + 1 : side == 2 ? 0 : cmy + 1;
  auto at = [&](int along) {
    const int wx = side < 2 ? fixed : along, wy = side < 2 ? along : fixed;
    return Wc[R.eoff + (wx - R.ex0) + (wy - R.ey0) * R.ew];
  };
  if (kind == G_QIF_END) {
    const int a0 = u == len ? clen + 1 : 0, st = u == len ? -1 : 1;
    return region_cf_parts<T>(kind, at(a0), at(a0 + st), at(a0 + 2 * st));
  }
  const T w = at((u & 1) == 0 ? (u >> 1) : ((u + 1) >> 1));
  return region_cf_parts<T>(kind, w, w, w);
}

// dynamic LDS of a region launch: the canvas, then (ghosts) the tile's ghost
// records, (ring) its staged slice of the QC ring refresh, (win) the parent
// window of the folded correction: parent values, then QC values
__host__ __device__ inline size_t region_align16(size_t n) { return (n + 15) & ~(size_t)15; }
template <typename T>
__host__ __device__ inline size_t region_lds_bytes(int rslots, int rghmax, int rring, int ghosts, int ring,
                                                   int win = 0) {
  size_t b = region_align16((size_t)rslots * sizeof(T));
  if (ghosts) b += region_align16((size_t)rghmax * sizeof(GhostOp<T>));
  if (ring) b += region_align16((size_t)rring * sizeof(int)) + region_align16((size_t)rring * sizeof(T));
  b += 2 * region_align16((size_t)win * sizeof(T));
  return b;
}

// Qout: where the own cells go.  With ghosts set, the tile also writes the
// ghost records next to its own cells from the final canvas (what the fill
// after the call would write; parent parts from aux), so Qout is the level's
// complete new state.  Rq: when given, the restriction of the tile's coarse
// cells into the parent (and Rdup) and QC's interior (ebTransfer.m:10-33).
// ring: the QC ring refresh of the fill, from the parent Pq.  corr: the
// coarse-grid correction of the V-cycle's up leg (ebTransfer.m:57-71) is
// applied to every loaded cell first, and the parent parts of coarse-fine
// ghosts come from the parent Pq, staged per patch window (L.rwin), in place
// of the parent parts a preparing correction launch would have stored.
// fload (a level of one tile whose launch writes its ghosts and the
// restriction): then the FAS load into the parent Pd (ebCoarseLoad), by the
// same workgroup after a barrier, from the level's new state in Qout; bits: 1
// load, 2 mass correction (reflux), 4 the parent's physical ghosts evaluated
// from its interior, 8 reflux sums from FTrue (see load_item).
template <typename T, int NT, int K, int KL, bool P2, bool W1, bool FL = false>
__global__ void __launch_bounds__(NT) k_relax_region(LevelDev<T> L, const T* __restrict__ Pq, T* __restrict__ Qout,
                                                     int npass, int frame, T omega, int homog, int per, int ghosts,
                                                     T* __restrict__ Rq, T* __restrict__ Rdup, int ring, int corr,
                                                     const LevelDev<T>* __restrict__ Pd, int fload) {
  extern __shared__ __align__(16) unsigned char region_lds[];
  T* cv = reinterpret_cast<T*>(region_lds);
  __shared__ RegionPatch sp[kRegionMaxPatches];
  const int h = blockIdx.x;
  const int t = (h % kXcds) * per + h / kXcds;
  if (t >= L.nregion) return;
  const int tid = threadIdx.x;
  const RegionTile tr = L.region[t];
  const unsigned* __restrict__ cl = L.rcell + (size_t)t * kRegionStride;
  unsigned e1[K], e2[K], e0[KL];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    e1[k] = cl[tid + k * NT];
    e2[k] = cl[NT * K + tid + k * NT];
  }
#pragma unroll
  for (int k = 0; k < KL; ++k) e0[k] = cl[2 * NT * K + tid + k * NT];
  if (tid < kRegionMaxPatches) sp[tid] = L.rpatch[(size_t)t * kRegionMaxPatches + tid];
  const int cw = tr.cw;
  // the epilogue's records (at most one per lane), loaded beside the cells
  unsigned char* tail = region_lds + region_align16((size_t)L.rslots * sizeof(T));
  GhostOp<T>* gop = reinterpret_cast<GhostOp<T>*>(tail);
  if (ghosts) tail += region_align16((size_t)L.rghmax * sizeof(GhostOp<T>));
  const int rring = ring ? (L.nring_cf + L.nregion - 1) / L.nregion : 0;  // ring items per tile
  const int rstage = rring < NT ? rring : NT;  // staged ring items (one per lane)
  int* rdst = reinterpret_cast<int*>(tail);
  T* rval = reinterpret_cast<T*>(tail + region_align16((size_t)rstage * sizeof(int)));
  if (ring) tail += region_align16((size_t)rstage * sizeof(int)) + region_align16((size_t)rstage * sizeof(T));
  T* Wc = reinterpret_cast<T*>(tail);  // the folded correction's window (corr)
  T* Cc = reinterpret_cast<T*>(tail + region_align16((size_t)L.rwinmax * sizeof(T)));
  const int nwin = corr ? tr.nwin : 0;
  RegionWin wv[2];
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (tid + k * NT < nwin) wv[k] = L.rwin[tr.wofs + tid + k * NT];
  const bool my_gh = ghosts && tid < tr.ngh, my_rs = Rq && tid < tr.nrs;
  const int rx = t * rring + tid;
  const bool my_ring = ring && tid < rstage && rx < L.nring_cf;
  RegionGhost ge;
  RingRec rr;
  int rs_qc = 0, rs_pq = 0, rs_slot = 0;  // this lane's coarse cell (registers until the epilogue)
  RegionGhostW gw;
  if (my_gh) ge = L.rghost[tr.gofs + tid];
  if (my_gh && corr) gw = L.rgw[tr.gofs + tid];
  if (my_rs) {
    const RegionRestrict& re = L.rrs[tr.rofs + tid];
    rs_qc = re.qc;
    rs_pq = re.pq;
    rs_slot = re.slot;
  }
  if (my_ring) rr = L.ring_cf[rx];
  for (int x = tid; x < cw * tr.ch; x += NT) cv[x] = T(0);
  __syncthreads();
  T gcf = T(0), gcfa = T(0), gcfb = T(0), rv = T(0);
  T wval[2] = {T(0), T(0)}, cval[2] = {T(0), T(0)};
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (tid + k * NT < nwin) {
      wval[k] = Pq[wv[k].wm];
      cval[k] = L.QC[wv[k].w];
    }
  if (my_gh && !corr) {
    if (ge.rec >= 0) gcf = L.aux[ge.rec].cf;
    if (ge.reca >= 0) gcfa = L.aux[ge.reca].cf;
    if (ge.recb >= 0) gcfb = L.aux[ge.recb].cf;
  }
  if (my_ring) rv = Pq[rr.src];
  // every cell's loads, then their uses
  RegionCell<T> c1[K], c2[K];
  T q1[K], q2[K], fx1[K], fy1[K], fdx, fdy, q0[KL];
  int qc1[K], qc2[K], kx1[K], ky1[K], kx2[K], ky2[K], s0[KL];
#pragma unroll
  for (int k = 0; k < KL; ++k) {
    const unsigned cd = e0[k] == kRegionNone ? 0u : e0[k];
    const int ps = cd & 15, i = (cd >> 4) & 4095, j = (cd >> 16) & 4095;
    const PatchInfo& P = sp[ps].pi;
    s0[k] = e0[k] == kRegionNone ? -1 : (sp[ps].ox + i) + (sp[ps].oy + j) * cw;
    q0[k] = L.Q[P.qoff + (i + 1) + (j + 1) * (P.mx + 2)];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    region_load<T>(L, sp, cw, e1[k], homog, frame != 0, c1[k], q1[k], qc1[k], fx1[k], fy1[k], kx1[k], ky1[k]);
    region_load<T>(L, sp, cw, e2[k], homog, false, c2[k], q2[k], qc2[k], fdx, fdy, kx2[k], ky2[k]);
  }
  if (corr) {  // the staged window, then the corrections and parent parts from it
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (tid + k * NT < nwin) {
        Wc[tid + k * NT] = wval[k];
        Cc[tid + k * NT] = cval[k];
      }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KL; ++k)
      if (e0[k] != kRegionNone) {
        const unsigned cd = e0[k];
        q0[k] = q0[k] + region_correction<T>(sp[cd & 15], (cd >> 4) & 4095, (cd >> 16) & 4095, Wc, Cc);
      }
    auto fix = [&](unsigned cd, T& q, RegionCell<T>& c, int kx, int ky) {
      if (cd == kRegionNone) return;
      const RegionPatch& R = sp[cd & 15];
      const int i = (cd >> 4) & 4095, j = (cd >> 16) & 4095;
      q = q + region_correction<T>(R, i, j, Wc, Cc);
      if (kx == G_QIF || kx == G_QIF_END || kx == G_LIF) c.bx = region_cf_window<T>(kx, i == 0 ? 0 : 1, j, R, Wc);
      if (ky == G_QIF || ky == G_QIF_END || ky == G_LIF) c.by = region_cf_window<T>(ky, j == 0 ? 2 : 3, i, R, Wc);
    };
#pragma unroll
    for (int k = 0; k < K; ++k) {
      fix(e1[k], q1[k], c1[k], kx1[k], ky1[k]);
      fix(e2[k], q2[k], c2[k], kx2[k], ky2[k]);
    }
    if (my_gh) {
      auto wat = [&](unsigned short sl) { return sl != kRegionNoSlot ? Wc[sl] : T(0); };
      const int kind = ge.kinds & 15;
      if (kind == G_PCORNER) {
        gcfa = region_cf_parts<T>((ge.kinds >> 4) & 15, wat(gw.s[0]), wat(gw.s[1]), wat(gw.s[2]));
        gcfb = region_cf_parts<T>((ge.kinds >> 8) & 15, wat(gw.s[3]), wat(gw.s[4]), wat(gw.s[5]));
      } else {
        gcf = region_cf_parts<T>(kind, wat(gw.s[0]), wat(gw.s[1]), wat(gw.s[2]));
      }
    }
  }
#pragma unroll
  for (int k = 0; k < KL; ++k)
    if (s0[k] >= 0) cv[s0[k]] = q0[k];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    region_kinds<T>(c1[k], kx1[k], ky1[k]);
    region_kinds<T>(c2[k], kx2[k], ky2[k]);
    if (c1[k].slot >= 0) cv[c1[k].slot] = q1[k];
    if (c2[k].slot >= 0) cv[c2[k].slot] = q2[k];
  }
  if (my_gh) {
    GhostOp<T> o;
    o.c = T(ge.c); o.ca = T(ge.ca); o.cb = T(ge.cb);
    o.cf = gcf; o.cfa = gcfa; o.cfb = gcfb;
    o.dst = ge.dst;
    o.kinds = ge.kinds;
    o.s01 = ge.s01;
    o.s23 = ge.s23;
    o.s4 = ge.s45 & 0xffffu;
    gop[tid] = o;
  }
  if (my_ring) {
    rdst[tid] = rr.dst;
    rval[tid] = rv;
  }
  __syncthreads();
  const int dshift = L.rdepth - npass;
  for (int p = 1; p <= npass; p += 2) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (c1[k].slot >= 0 && region_depth(c1[k].fl) >= p + dshift)
        cv[c1[k].slot] = region_update<T, P2, W1>(L, cv, cw, c1[k], frame && p == 1, fx1[k], fy1[k], omega);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (c2[k].slot >= 0 && region_depth(c2[k].fl) >= p + 1 + dshift)
        cv[c2[k].slot] = region_update<T, P2, W1>(L, cv, cw, c2[k], false, T(0), T(0), omega);
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (c1[k].slot >= 0 && region_depth(c1[k].fl) == L.rdepth) Qout[qc1[k]] = cv[c1[k].slot];
    if (c2[k].slot >= 0 && region_depth(c2[k].fl) == L.rdepth) Qout[qc2[k]] = cv[c2[k].slot];
  }
  if (my_gh) {  // one ghost record per lane (at most NT per tile)
    const GhostOp<T> o = gop[tid];
    const int kind = o.kinds & 15;
    const unsigned a0 = o.s01 & 0xffffu, a1 = o.s01 >> 16;
    const T v0 = cv[a0], v1 = a1 != kRegionNoSlot ? cv[a1] : T(0);
    T v;
    if (kind == G_PCORNER) {
      const unsigned b0 = o.s23 & 0xffffu, b1 = o.s23 >> 16;
      const T a = region_ghost<T>((o.kinds >> 4) & 15, o.ca, o.cfa, v0, v1, homog);
      const T b = region_ghost<T>((o.kinds >> 8) & 15, o.cb, o.cfb, cv[b0], b1 != kRegionNoSlot ? cv[b1] : T(0),
                                  homog);
      v = (a + b) - cv[o.s4];
    } else {
      v = region_ghost<T>(kind, o.c, o.cf, v0, v1, homog);
    }
    Qout[o.dst] = v;
  }
  if (my_rs) {  // one coarse cell per lane (at most NT per tile)
    const int sl = rs_slot;
    const T R = T(0.25) * (((cv[sl] + cv[sl + 1]) + cv[sl + cw]) + cv[sl + cw + 1]);
    L.QC[rs_qc] = R;
    Rq[rs_pq] = R;
    if (Rdup) Rdup[rs_pq] = R;
  }
  if (my_ring) L.QC[rdst[tid]] = rval[tid];
  if (ring)  // a slice longer than the workgroup: the rest directly
    for (int x = NT + tid; x < rring && t * rring + x < L.nring_cf; x += NT) {
      const RingRec q = L.ring_cf[t * rring + x];
      L.QC[q.dst] = Pq[q.src];
    }
  if (FL && fload) {  // the workgroup's stores above are visible to it after the barrier
    __syncthreads();
    LevelDev<T> Lc = L;
    Lc.Q = Qout;
    const LevelDev<T>& P = *Pd;
    const int mc = (fload & 2) ? 1 : 0;
    const int n = load_items<T>(Lc, P, mc);
    for (int x = tid; x < n; x += NT) load_item<T>(Lc, P, x, mc, (fload & 4) ? 1 : 0, homog, (fload & 8) ? 1 : 0);
  }
}

// ebFillGhosts after a region relax: the ghosts from the new interior values
// in Qsrc (fill items, blocks [0, nfb)) and the interior copied from Qsrc into
// Q (one block per 256-cell tile); nothing in the launch reads Q
template <typename T>
__global__ void __launch_bounds__(kTile) k_fill_copy(LevelDev<T> L, const T* __restrict__ Pq, int homog,
                                                     const T* __restrict__ Qsrc, int nfb) {
  if ((int)blockIdx.x < nfb) {
    const int t = blockIdx.x * kTile + threadIdx.x;
    if (t < fill_items<T>(L, Pq)) fill_item<T>(L, Pq, t, homog, Qsrc);
    return;
  }
  const TileRec tr = load_tile(L.tiles, __builtin_amdgcn_readfirstlane((int)blockIdx.x - nfb));
  const PatchInfo& pi = tr.pi;
  const int c = tr.start + threadIdx.x;
  if (c >= pi.mx * pi.my) return;
  const int j = cell_row(c, pi.mx, pi.magic);
  const int i = c - j * pi.mx;
  const int qc = pi.qoff + (i + 1) + (j + 1) * (pi.mx + 2);
  L.Q[qc] = Qsrc[qc];
}

// the parent part of every coarse-fine record of a level from the current
// parent, into aux (what a fill stores, without touching Q)
template <typename T>
__global__ void __launch_bounds__(kTile) k_refresh_cf(LevelDev<T> L, const T* __restrict__ Pq) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t < L.nghost) keep_face_cf<T>(L, Pq, t);
}

// ---------------------------------------------------------------- streaming sweep
// one row of a lane's two columns A, A+1 (A even, the level's width even):
// coefficients, the faces (west faces of both, east face of A+1, south faces
// of both) and the old values; loaded at clamped, in-bounds indices
template <typename T>
struct StreamRow {
  T f0, f1, id0, id1, dw0, dw1, de1, ds0, ds1, q0, q1;
};

template <typename T>
__device__ __forceinline__ void stream_row(const LevelDev<T>& L, const PatchInfo& P, const T* __restrict__ Q, int A,
                                           int r, StreamRow<T>& w) {
  const int mx = P.mx, my = P.my;
  A = min(max(A, 0), mx - 2);
  const int rf = min(max(r, -1), my);  // south faces exist for rows 0..my
  r = min(max(r, 0), my - 1);
  const int cc = P.coff + A + r * mx, fx = P.xoff + A + r * (mx + 1), fy = P.yoff + A + max(rf, 0) * mx;
  const int qc = (A + 1) + (r + 1) * (mx + 2);
  w.f0 = L.F[cc]; w.f1 = L.F[cc + 1];
  w.id0 = L.ID[cc]; w.id1 = L.ID[cc + 1];
  w.dw0 = L.DeW[fx]; w.dw1 = L.DeW[fx + 1]; w.de1 = L.DeW[fx + 2];
  w.ds0 = L.DnS[fy]; w.ds1 = L.DnS[fy + 1];
  w.q0 = Q[qc]; w.q1 = Q[qc + 1];
}

// boundary value of side s at position t (0 when homogeneous)
template <typename T>
__device__ __forceinline__ T stream_g(const LevelDev<T>& L, const PatchInfo& P, int side, int t, int homog) {
  if (homog) return T(0);
  const int mx = P.mx, my = P.my;
  const int off = side == 0 ? 0 : side == 1 ? my : side == 2 ? 2 * my : 2 * my + mx;
  return L.G[P.goff + off + t];
}

// update of cell (i, j) from its row data (column k of the pair), the south
// face of the row above (dn) and its resolved-or-raw neighbours
template <typename T>
__device__ __forceinline__ T stream_cell(const LevelDev<T>& L, const PatchInfo& P, int i, int j, const StreamRow<T>& w,
                                         int k, T dn, T e, T wv, T n, T s, T self, T omega, int homog) {
  const int mx = P.mx, my = P.my;
  T g[4] = {T(0), T(0), T(0), T(0)};
  if (i == 0) g[0] = stream_g<T>(L, P, 0, j, homog);
  if (i == mx - 1) g[1] = stream_g<T>(L, P, 1, j, homog);
  if (j == 0) g[2] = stream_g<T>(L, P, 2, i, homog);
  if (j == my - 1) g[3] = stream_g<T>(L, P, 3, i, homog);
  const T x[4] = {T(0), T(0), T(0), T(0)};
  const FaceAux<T> a[4] = {};
  T qe, qw, qn, qs;
  resolve_ghosts<T>(L, P, i, j, e, wv, n, s, self, g, x, a, false, qe, qw, qn, qs);
  const T f = k ? w.f1 : w.f0, id = k ? w.id1 : w.id0;
  const T dw = k ? w.dw1 : w.dw0, de = k ? w.de1 : w.dw1, ds = k ? w.ds1 : w.ds0;
  return relax_combine<T>(L, qe, qw, qn, qs, f, de, dw, dn, ds, id, self, omega);
}

template <typename T>
__device__ __forceinline__ T lane_up(T v) {  // value of lane l-1
  return __shfl_up(v, 1, 64);
}
template <typename T>
__device__ __forceinline__ T lane_down(T v) {  // value of lane l+1
  return __shfl_down(v, 1, 64);
}

// One red+black sweep of a level that is one patch of even width with
// physical boundaries on every side (a uniform level), Qin -> Qout.  Each
// 64-lane wave owns a strip of columns, two per lane, and marches up a block
// of rows.  At row step j it loads row j of its columns and updates the red
// cell of row j-2 and the black cell of row j-3 (in the red column of row
// j-2), so each row's loads are in flight for a whole step before use.  Red
// updates read only old black values (the lane's own rows and, across
// columns, the neighbour lanes' through cross-lane moves); black updates read
// the red values of this sweep (own registers and the neighbour lanes'), so
// one pass over the rows does the whole sweep and every loaded line is used
// in the step that loads it; a row's two columns are stored together.  The
// strip's first and last lanes are a halo (their black values would need
// lanes outside the wave and are not stored) and the red rows just outside
// the block are recomputed.  Even row blocks march up and odd ones down, so
// that neighbouring blocks read their shared edge rows at about the same time.  Every update is the colour pass's expression
// (relax_combine) with the partial physical ghosts of ebKernels.m:51-70, so
// the sweep equals two colour passes bit for bit.
template <typename T>
__global__ void __launch_bounds__(64) k_sweep_stream(LevelDev<T> L, PatchInfo P, const T* __restrict__ Qin,
                                                     T* __restrict__ Qout, T omega, int homog, int kRows) {
  const int mx = P.mx, my = P.my, ldq = mx + 2;
  const int nstrip = (mx + kStreamCols - 1) / kStreamCols;
  // workgroups go to the 8 XCDs in turn: XCD x takes the x-th contiguous
  // range of (row block, strip) items, so the strips of a row block run side
  // by side on one XCD and share the cache lines that straddle their edges
  const int total = nstrip * ((my + kRows - 1) / kRows), per = (total + kXcds - 1) / kXcds;
  const int item = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  if (item >= total) return;
  const int strip = item % nstrip, blk = item / nstrip;
  const int j0 = blk * kRows, j1 = min(j0 + kRows, my);
  const int own0 = strip * kStreamCols, own1 = min(own0 + kStreamCols, mx);
  const int A = own0 - 2 + 2 * (int)threadIdx.x;  // the lane's columns A, A+1
  const bool owned = A >= own0 && A < own1;       // both columns or neither (own0, own1, A even)
  const T* __restrict__ Q = Qin + P.qoff;
  T* __restrict__ Qo = Qout + P.qoff;
  StreamRow<T> r0, r1, r2, r3;
  if (blk & 1) {
    // odd row blocks march down: they meet the even blocks above and below at
    // the start and the end of the launch, when both read the shared edge rows
    // rows jr+1, jr, jr-1 (loaded) around the red row jr = j+2, and row j (loading)
    stream_row<T>(L, P, Q, A, j1 + 1, r0);
    stream_row<T>(L, P, Q, A, j1, r1);
    stream_row<T>(L, P, Q, A, j1 - 1, r2);
    T rd1 = T(0), rd2 = T(0);     // red values of this sweep, one and two rows above the red row
    T pds0 = T(0), pds1 = T(0);   // south faces of the row two above the red row
    for (int j = j1 - 2; j >= j0 - 3; --j) {
      stream_row<T>(L, P, Q, A, j, r3);
      const int jr = j + 2;
      const int k = ((A + jr + P.par0) & 1) == 0 ? 0 : 1;  // red column of row jr (uniform)
      const int c = A + k;
      const bool col = c >= 0 && c < mx;
      const T qo = k ? r1.q0 : r1.q1;
      const T from_up = lane_up(r1.q1), from_down = lane_down(r1.q0);
      T red = T(0);
      if (col && jr >= max(j0 - 1, 0) && jr <= min(j1, my - 1)) {
        const T e = k ? from_down : qo, w = k ? qo : from_up;
        const T n = k ? r0.q1 : r0.q0, s = k ? r2.q1 : r2.q0, self = k ? r1.q1 : r1.q0;
        red = stream_cell<T>(L, P, c, jr, r1, k, k ? r0.ds1 : r0.ds0, e, w, n, s, self, omega, homog);
      }
      // black cell (c, jr+1): red of row jr+1 along the row, red of rows jr+2 and jr
      const T rup = lane_up(rd1), rdown = lane_down(rd1);
      const int jb = jr + 1;
      if (col && jb >= j0 && jb < j1) {
        const T e = k ? rdown : rd1, w = k ? rd1 : rup;
        const T self = k ? r0.q1 : r0.q0;
        const T v = stream_cell<T>(L, P, c, jb, r0, k, k ? pds1 : pds0, e, w, rd2, red, self, omega, homog);
        if (owned) {
          const int qa = (A + 1) + (jb + 1) * ldq;
          Qo[qa + k] = v;
          Qo[qa + 1 - k] = rd1;
        }
      }
      rd2 = rd1;
      rd1 = red;
      pds0 = r0.ds0;
      pds1 = r0.ds1;
      r0 = r1;
      r1 = r2;
      r2 = r3;
    }
    return;
  }
  // rows jr-1, jr, jr+1 (loaded) around the red row jr = j-2, and row j (loading)
  stream_row<T>(L, P, Q, A, j0 - 2, r0);
  stream_row<T>(L, P, Q, A, j0 - 1, r1);
  stream_row<T>(L, P, Q, A, j0, r2);
  T rd1 = T(0), rd2 = T(0);  // red values of this sweep, one and two rows below the red row
  for (int j = j0 + 1; j <= j1 + 2; ++j) {
    stream_row<T>(L, P, Q, A, j, r3);
    const int jr = j - 2;
    const int k = ((A + jr + P.par0) & 1) == 0 ? 0 : 1;  // red column of row jr (uniform)
    const int c = A + k;
    const bool col = c >= 0 && c < mx;
    const T qo = k ? r1.q0 : r1.q1;  // the other column's old (black) value in row jr
    const T from_up = lane_up(r1.q1), from_down = lane_down(r1.q0);
    T red = T(0);
    if (col && jr >= max(j0 - 1, 0) && jr <= min(j1, my - 1)) {
      const T e = k ? from_down : qo, w = k ? qo : from_up;
      const T n = k ? r2.q1 : r2.q0, s = k ? r0.q1 : r0.q0, self = k ? r1.q1 : r1.q0;
      red = stream_cell<T>(L, P, c, jr, r1, k, k ? r2.ds1 : r2.ds0, e, w, n, s, self, omega, homog);
    }
    // black cell (c, jr-1): red neighbours of row jr-1 along the row (the
    // lane's other column, rd1, and the neighbour lane's), red of rows jr and jr-2
    const T rup = lane_up(rd1), rdown = lane_down(rd1);
    const int jb = jr - 1;
    if (col && jb >= j0 && jb < j1) {
      const T e = k ? rdown : rd1, w = k ? rd1 : rup;
      const T self = k ? r0.q1 : r0.q0;
      const T v = stream_cell<T>(L, P, c, jb, r0, k, k ? r1.ds1 : r1.ds0, e, w, red, rd2, self, omega, homog);
      // row jb is complete in both columns (its red cell, in the other column,
      // was updated one step earlier): both are stored together
      if (owned) {
        const int qa = (A + 1) + (jb + 1) * ldq;
        Qo[qa + k] = v;
        Qo[qa + 1 - k] = rd1;
      }
    }
    rd2 = rd1;
    rd1 = red;
    r0 = r1;
    r1 = r2;
    r2 = r3;
  }
}

// ---------------------------------------------------------------- tiled smoother
// All 2*nsweeps colour passes of one relax call on one tile, with the tile
// and a 2-cell margin (clipped to the padded patch) staged in LDS: one HBM
// round trip per relax call instead of one per colour.  Inside the tile the
// passes are exact red-black Gauss-Seidel: physical ghosts are the partial
// ghosts of ebKernels.m:51-70 and coarse-fine ghosts are re-evaluated from the
// staged cells and the parent after every pass.  Same-level (sibling) ghosts
// and cells of neighbouring tiles keep the values they had when the call
// started (a block Gauss-Seidel coupling between tiles); the converged
// composite solution is unchanged because at the fixed point those values are
// stationary.
template <typename T>
__device__ __forceinline__ T lds_or_global(const T* S, const T* __restrict__ Q, int s, int qoff, int ldq, int a0,
                                           int a1, int b0, int b1, int lw) {
  const int d = s - qoff;
  const int pj = d / ldq;
  const int pi_ = d - pj * ldq;
  if (pi_ >= a0 && pi_ <= a1 && pj >= b0 && pj <= b1) return S[(pi_ - a0) + (pj - b0) * lw];
  return Q[s];
}

template <typename T>
__device__ __forceinline__ T ghost_value_tile(const GhostRec& r, const T* S, const T* __restrict__ Q,
                                              const T* __restrict__ Pq, int qoff, int ldq, int a0, int a1, int b0,
                                              int b1, int lw, T frozen) {
  auto q = [&](int s) { return lds_or_global<T>(S, Q, s, qoff, ldq, a0, a1, b0, b1, lw); };
  switch (r.kind) {
    case G_QIF:
      return (T(A1_) * q(r.s[0]) + T(A2_) * q(r.s[1])) + T(A3_) * Pq[r.s[2]];
    case G_QIF_END:
      return (T(A1_) * q(r.s[0]) + T(A2_) * q(r.s[1])) +
             T(A3_) * ((T(B1_) * Pq[r.s[2]] + T(B2_) * Pq[r.s[3]]) + T(B3_) * Pq[r.s[4]]);
    case G_LIF:
      return (T(C23_) * Pq[r.s[2]] + T(C23_) * q(r.s[0])) - T(C13_) * q(r.s[1]);
    default:  // same-level copy: frozen for the call
      return frozen;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_relax_tiled(LevelDev<T> L, const T* __restrict__ Pq, int nsweeps,
                                                     T omega, int homog) {
  extern __shared__ unsigned char smem_raw[];
  T* S = reinterpret_cast<T*>(smem_raw);
  const STile tl = L.stiles[blockIdx.x];
  const PatchInfo& pi = L.P[tl.p];
  const int mx = pi.mx, my = pi.my, ldq = mx + 2;
  const int a0 = max(tl.i0 - 1, 0), a1 = min(tl.i1 + 3, mx + 1);  // padded patch coordinates
  const int b0 = max(tl.j0 - 1, 0), b1 = min(tl.j1 + 3, my + 1);
  const int lw = a1 - a0 + 1, lh = b1 - b0 + 1;
  T* __restrict__ Q = L.Q;
  for (int x = threadIdx.x; x < lw * lh; x += blockDim.x) {
    const int lj = x / lw;
    S[x] = Q[pi.qoff + (a0 + x - lj * lw) + (b0 + lj) * ldq];
  }
  __syncthreads();
  const int tw = tl.i1 - tl.i0 + 1, th = tl.j1 - tl.j0 + 1;
  const T hx = (T)L.h[0], hy = (T)L.h[1];
  const T hx2 = hx * hx, hy2 = hy * hy;
  for (int it = 0; it < nsweeps; ++it)
    for (int colour = 1; colour <= 2; ++colour) {
      const bool inl = !(it == 0 && colour == 1);
      for (int x = threadIdx.x; x < tw * th; x += blockDim.x) {
        const int jj = x / tw;
        const int i = tl.i0 + x - jj * tw, j = tl.j0 + jj;
        if (((i + j + pi.par0) & 1) != colour - 1) continue;
        const int s = (i + 1 - a0) + (j + 1 - b0) * lw;
        T qe = S[s + 1], qw = S[s - 1], qn = S[s + lw], qs = S[s - lw];
        if (i == 0) {
          if (pi.bc[0]) {
            const T g = homog ? T(0) : L.G[pi.goff + j];
            qw = pi.bc[0] == 1 ? T(C83_) * g + T(C13_) * S[s + 1] : hx * g;
          } else if (inl) {
            qw = ghost_value_tile<T>(L.ghost[pi.gbase + j], S, Q, Pq, pi.qoff, ldq, a0, a1, b0, b1, lw, qw);
          }
        }
        if (i == mx - 1) {
          if (pi.bc[1]) {
            const T g = homog ? T(0) : L.G[pi.goff + my + j];
            qe = pi.bc[1] == 1 ? T(C83_) * g + T(C13_) * S[s - 1] : hx * g;
          } else if (inl) {
            qe = ghost_value_tile<T>(L.ghost[pi.gbase + my + j], S, Q, Pq, pi.qoff, ldq, a0, a1, b0, b1, lw, qe);
          }
        }
        if (j == 0) {
          if (pi.bc[2]) {
            const T g = homog ? T(0) : L.G[pi.goff + 2 * my + i];
            qs = pi.bc[2] == 1 ? T(C83_) * g + T(C13_) * S[s + lw] : hy * g;
          } else if (inl) {
            qs = ghost_value_tile<T>(L.ghost[pi.gbase + 2 * my + i], S, Q, Pq, pi.qoff, ldq, a0, a1, b0, b1, lw, qs);
          }
        }
        if (j == my - 1) {
          if (pi.bc[3]) {
            const T g = homog ? T(0) : L.G[pi.goff + 2 * my + mx + i];
            qn = pi.bc[3] == 1 ? T(C83_) * g + T(C13_) * S[s - lw] : hy * g;
          } else if (inl) {
            qn = ghost_value_tile<T>(L.ghost[pi.gbase + 2 * my + mx + i], S, Q, Pq, pi.qoff, ldq, a0, a1, b0, b1,
                                     lw, qn);
          }
        }
        const int cc = pi.coff + i + j * mx;
        const int fx = pi.xoff + i + j * (mx + 1);
        const int fy = pi.yoff + i + j * mx;
        T tmp = L.F[cc] + div_h2<T>(qe * L.DeW[fx + 1], hx2, (T)L.ih2[0], L.pow2);
        tmp = tmp + div_h2<T>(qw * L.DeW[fx], hx2, (T)L.ih2[0], L.pow2);
        tmp = tmp + div_h2<T>(qn * L.DnS[fy + mx], hy2, (T)L.ih2[1], L.pow2);
        tmp = tmp + div_h2<T>(qs * L.DnS[fy], hy2, (T)L.ih2[1], L.pow2);
        tmp = tmp * L.ID[cc];
        S[s] = omega == T(1) ? tmp : omega * tmp + (T(1) - omega) * S[s];
      }
      __syncthreads();
    }
  for (int x = threadIdx.x; x < tw * th; x += blockDim.x) {
    const int jj = x / tw;
    const int i = tl.i0 + x - jj * tw, j = tl.j0 + jj;
    Q[pi.qoff + (i + 1) + (j + 1) * ldq] = S[(i + 1 - a0) + (j + 1 - b0) * lw];
  }
}

// ---------------------------------------------------------------- transfers
template <typename T>
__device__ __forceinline__ void ring_item(const LevelDev<T>& L, const T* __restrict__ Pq, int t) {
  const RingRec rr = L.ring_all[t];
  L.QC[rr.dst] = Pq[rr.src];
}

template <typename T>
__device__ __forceinline__ void restrict_item(const LevelDev<T>& L, T* __restrict__ Pq, int t,
                                              int qc_only, T* __restrict__ Pdup = nullptr,
                                              const T* __restrict__ Qsrc = nullptr) {
  const FootRec fr = L.foot[t];
  const FootRec& pi = fr;  // patch offsets carried inline
  const int ldq = pi.mx + 2;
  const T* __restrict__ Q = Qsrc ? Qsrc : L.Q;
  const int q00 = pi.qoff + (2 * fr.ci + 1) + (2 * fr.cj + 1) * ldq;
  const T R = T(0.25) * (((Q[q00] + Q[q00 + 1]) + Q[q00 + ldq]) + Q[q00 + ldq + 1]);
  L.QC[pi.qcoff + (fr.ci + 1) + (fr.cj + 1) * (pi.mx / 2 + 2)] = R;
  if (!qc_only) Pq[fr.pq] = R;
  if (Pdup) Pdup[fr.pq] = R;  // the parent's second array (streaming sweeps)
}

// The fill that follows a region relax and the restriction that follows the
// pre-smoothing (ebVcycle.m:21-22) as one launch.  Nothing in it reads what
// it writes: the ghosts read the new interior from Qsrc and the parent parts
// of coarse-fine records from aux (current: the region launch read them and
// the parent has not changed since), the restriction reads Qsrc and writes
// the parent and QC's interior, and the interior copy writes Q.  The fill's
// QC ring refresh is left out: the load that follows the restriction gathers
// the ring again from the restricted parent before anything reads it.
// Blocks: [0, nfb) ghosts, then one per 256-cell tile (copy), then the
// restriction's footprint cells.
template <typename T>
__global__ void __launch_bounds__(kTile) k_fill_copy_restrict(LevelDev<T> L, int homog, const T* __restrict__ Qsrc,
                                                              int nfb, T* __restrict__ Pq, T* __restrict__ Pdup) {
  const int b = blockIdx.x;
  if (b < nfb) {
    const int t = b * kTile + threadIdx.x;
    if (t >= L.nghost) return;
    const GhostRec r = L.ghost[t];
    T v;
    if (r.kind == G_PCORNER) {
      const T a = ghost_value_cf<T>(L.ghost[r.s[0]], Qsrc, L.aux[r.s[0]].cf, homog);
      const T c = ghost_value_cf<T>(L.ghost[r.s[1]], Qsrc, L.aux[r.s[1]].cf, homog);
      v = (a + c) - Qsrc[r.s[2]];
    } else {
      v = ghost_value_cf<T>(r, Qsrc, is_cf_kind(r.kind) ? L.aux[t].cf : T(0), homog);
    }
    L.Q[r.dst] = v;
    return;
  }
  if (b < nfb + L.ntile) {
    const TileRec tr = load_tile(L.tiles, __builtin_amdgcn_readfirstlane(b - nfb));
    const PatchInfo& pi = tr.pi;
    const int c = tr.start + threadIdx.x;
    if (c >= pi.mx * pi.my) return;
    const int j = cell_row(c, pi.mx, pi.magic);
    const int i = c - j * pi.mx;
    const int qc = pi.qoff + (i + 1) + (j + 1) * (pi.mx + 2);
    L.Q[qc] = Qsrc[qc];
    return;
  }
  const int t = (b - nfb - L.ntile) * kTile + threadIdx.x;
  if (t < L.nfoot) restrict_item<T>(L, Pq, t, 0, Pdup, Qsrc);
}


// L(u) at one cell of a padded array (ebKernels.m:13-25)
// the level's h^2 and reciprocals for div_h2
template <typename T>
struct H2 {
  T hx2, hy2, ihx2, ihy2;
  int pow2;
};

template <typename T>
__device__ __forceinline__ H2<T> level_h2(const LevelDev<T>& L) {
  const T hx = (T)L.h[0], hy = (T)L.h[1];
  return H2<T>{hx * hx, hy * hy, (T)L.ih2[0], (T)L.ih2[1], L.pow2};
}

template <typename T>
__device__ __forceinline__ T op5(T u, T w, T e, T s, T n, T dw, T de, T ds, T dn, T cc, const H2<T>& g) {
  const T fw = dw * (u - w);
  const T fe = de * (e - u);
  const T fs = ds * (u - s);
  const T fn = dn * (n - u);
  return (div_h2<T>(-(fe - fw), g.hx2, g.ihx2, g.pow2) - div_h2<T>(fn - fs, g.hy2, g.ihy2, g.pow2)) + cc * u;
}

template <typename T>
__device__ __forceinline__ T op_at(const T* __restrict__ Q, int qc, int ldq, T dw, T de, T ds, T dn,
                                   T cc, const H2<T>& g) {
  return op5<T>(Q[qc], Q[qc - 1], Q[qc + 1], Q[qc - ldq], Q[qc + ldq], dw, de, ds, dn, cc, g);
}

// restricted fine residual R(F - L u) over one coarse footprint cell
template <typename T>
__device__ __forceinline__ T restricted_residual(const LevelDev<T>& L, const FootRec& pi, int ci, int cj) {
  const int mx = pi.mx;
  const int ldq = mx + 2;
  const H2<T> g = level_h2<T>(L);
  T rf[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int i = 2 * ci + (s & 1);
    const int j = 2 * cj + (s >> 1);
    const int qc = pi.qoff + (i + 1) + (j + 1) * ldq;
    const int cc = pi.coff + i + j * mx;
    const int fx = pi.xoff + i + j * (mx + 1);
    const int fy = pi.yoff + i + j * mx;
    rf[s] = L.F[cc] - op_at<T>(L.Q, qc, ldq, L.DeW[fx], L.DeW[fx + 1], L.DnS[fy], L.DnS[fy + mx], L.CC[cc], g);
  }
  return T(0.25) * (((rf[0] + rf[1]) + rf[2]) + rf[3]);
}

// one flux-balance contribution (ebFaceCorrections.m:29-55), in FP64;
// `ring` is the QC ring value just outside the coarse face
// hc: the coarse (parent) level's h^2 in FP64
template <typename T>
__device__ __forceinline__ double corr_value(const CorrRec& e, const T* __restrict__ Q, double ring,
                                             const T* __restrict__ QC, const T* __restrict__ DeW,
                                             const T* __restrict__ DnS, const H2<double>& hc) {
  const T* D = e.face < 2 ? DeW : DnS;
  const double t1 = (double)D[e.d1] * ((double)Q[e.in1] - (double)Q[e.g1]);
  const double t2 = (double)D[e.d2] * ((double)Q[e.in2] - (double)Q[e.g2]);
  const double t3 = e.cf * (ring - (double)QC[e.qcin]);
  return e.face < 2 ? div_h2<double>((t1 + t2) + t3, hc.hx2, hc.ihx2, hc.pow2)
                    : div_h2<double>((t1 + t2) + t3, hc.hy2, hc.ihy2, hc.pow2);
}

// FP64 h^2 of a level whatever the field type (reflux sums are FP64)
template <typename T>
__device__ __forceinline__ H2<double> level_h2_64(const LevelDev<T>& L) {
  return H2<double>{L.h[0] * L.h[0], L.h[1] * L.h[1], L.ih2[0], L.ih2[1], L.pow2};
}

// ebCoarseLoad for one footprint cell, QC ring read from QC (standalone form)
template <typename T>
__device__ __forceinline__ void coarse_load_item(const LevelDev<T>& L, const LevelDev<T>& P, int t) {
  const FootRec fr = L.foot[t];
  const FootRec& pi = fr;  // patch offsets carried inline
  const T R = restricted_residual<T>(L, pi, fr.ci, fr.cj);
  const int ldc = pi.mx / 2 + 2;
  const int qcc = pi.qcoff + (fr.ci + 1) + (fr.cj + 1) * ldc;
  const T Lc = op_at<T>(L.QC, qcc, ldc, P.DeW[fr.dew_w], P.DeW[fr.dew_e], P.DnS[fr.dns_s], P.DnS[fr.dns_n],
                        P.CC[fr.pc], level_h2<T>(P));
  P.F[fr.pc] = R + Lc;
}

// from_ft: the sum starts from FTrue instead of F.  In a loop whose cycles
// each begin with F = FTrue (ebSolveCycles.m:30-34) the two are equal at a
// target cell, which only this item writes; the loop then resets F once
// before its first cycle instead of before each one (the FAS load overwrites
// covered cells without reading them, and no other cell's F changes).
template <typename T>
__device__ __forceinline__ void reflux_item(const LevelDev<T>& L, const LevelDev<T>& P, int t, int ring_from_parent,
                                            int from_ft = 0) {
  const int cell = P.tgt_cell[t];
  double acc = from_ft ? (double)P.FT[cell] : (double)P.F[cell];
  const H2<double> hc = level_h2_64<T>(P);
  for (int e = P.tgt_ptr[t]; e < P.tgt_ptr[t + 1]; ++e) {
    const CorrRec cr = L.corr[P.tgt_ent[e]];
    const double ring = ring_from_parent ? (double)P.Q[L.wmap[cr.ring]] : (double)L.QC[cr.ring];
    acc = acc + corr_value<T>(cr, L.Q, ring, L.QC, L.DeW, L.DnS, hc);
  }
  P.F[cell] = (T)acc;
}

// ring gather + FAS load + reflux of level k into level k-1 in one sweep
// (ebTransfer 'ring' followed by ebCoarseLoad): ring values are read from the
// parent through the window map, which is exactly what the gather writes.
template <typename T>
__device__ __forceinline__ int load_items(const LevelDev<T>& L, const LevelDev<T>& P, int mass_corr) {
  return L.nfoot + L.nring_all + (mass_corr ? P.ntgt : 0);
}

// A parent value the FAS load reads: Pq[q], or, when the parent's frame was
// not filled (rec >= 0: a physical ghost record of the parent), the value
// the fill would have written, from the parent's interior
template <typename T>
__device__ __forceinline__ T parent_at(const LevelDev<T>& P, const T* __restrict__ Pq, int q, int rec, int homog) {
  if (rec < 0) return Pq[q];
  const GhostRec r = P.ghost[rec];
  if (r.kind == G_PCORNER) {
    const T a = ghost_value_cf<T>(P.ghost[r.s[0]], Pq, T(0), homog);
    const T b = ghost_value_cf<T>(P.ghost[r.s[1]], Pq, T(0), homog);
    return (a + b) - Pq[r.s[2]];
  }
  return ghost_value_cf<T>(r, Pq, T(0), homog);
}

// pghost: the parent's frame was not filled after the restriction, so its
// physical ghosts are evaluated here (homog: homogeneous boundary data)
template <typename T>
__device__ __forceinline__ void load_item(const LevelDev<T>& L, const LevelDev<T>& P, int t, int mass_corr,
                                          int pghost = 0, int homog = 0, int from_ft = 0) {
  if (t < L.nfoot) {
    const FootRec fr = L.foot[t];
    const FootRec& pi = fr;  // patch offsets carried inline
    const T R = restricted_residual<T>(L, pi, fr.ci, fr.cj);
    const int ldc = pi.mx / 2 + 2;
    const int w = pi.qcoff + (fr.ci + 1) + (fr.cj + 1) * ldc;
    const int cmx = pi.mx / 2, cmy = pi.my / 2;
    const T* __restrict__ QC = L.QC;
    const T* __restrict__ Pq = P.Q;
    const T u = QC[w];
    T vw, ve, vs, vn;
    if (pghost) {
      vw = fr.ci == 0 ? parent_at<T>(P, Pq, L.wmap[w - 1], L.wgr[w - 1], homog) : QC[w - 1];
      ve = fr.ci == cmx - 1 ? parent_at<T>(P, Pq, L.wmap[w + 1], L.wgr[w + 1], homog) : QC[w + 1];
      vs = fr.cj == 0 ? parent_at<T>(P, Pq, L.wmap[w - ldc], L.wgr[w - ldc], homog) : QC[w - ldc];
      vn = fr.cj == cmy - 1 ? parent_at<T>(P, Pq, L.wmap[w + ldc], L.wgr[w + ldc], homog) : QC[w + ldc];
    } else {
      vw = fr.ci == 0 ? Pq[L.wmap[w - 1]] : QC[w - 1];
      ve = fr.ci == cmx - 1 ? Pq[L.wmap[w + 1]] : QC[w + 1];
      vs = fr.cj == 0 ? Pq[L.wmap[w - ldc]] : QC[w - ldc];
      vn = fr.cj == cmy - 1 ? Pq[L.wmap[w + ldc]] : QC[w + ldc];
    }
    const T Lc = op5<T>(u, vw, ve, vs, vn, P.DeW[fr.dew_w], P.DeW[fr.dew_e], P.DnS[fr.dns_s], P.DnS[fr.dns_n],
                        P.CC[fr.pc], level_h2<T>(P));
    P.F[fr.pc] = R + Lc;
  } else if (t < L.nfoot + L.nring_all) {
    if (pghost) {
      const int x = t - L.nfoot;
      const RingRec rr = L.ring_all[x];
      L.QC[rr.dst] = parent_at<T>(P, P.Q, rr.src, L.ring_rec[x], homog);
    } else {
      ring_item<T>(L, P.Q, t - L.nfoot);
    }
  } else {
    reflux_item<T>(L, P, t - L.nfoot - L.nring_all, 1, from_ft);
  }
}

template <typename T>
__device__ __forceinline__ void correct_item(const LevelDev<T>& L, const T* __restrict__ Pq, int tile,
                                             int lane, int init, int prep = 0, T* __restrict__ Qdup = nullptr) {
  const TileRec& tr = L.tiles[tile];
  const PatchInfo& pi = tr.pi;
  const int mx = pi.mx, my = pi.my;
  const int c = tr.start + lane;
  if (c >= mx * my) return;
  const int j = cell_row(c, mx, pi.magic);
  const int i = c - j * mx;
  // prep: the first post-smoothing pass assembles its ghosts, so store the
  // parent parts of this cell's coarse-fine face records (the parent is
  // final for the level's post-smoothing) and, below, push the corrected
  // value into the sibling ghosts that copy it
  const bool fw = prep && i == 0 && !pi.bc[0], fe = prep && i == mx - 1 && !pi.bc[1];
  const bool fs = prep && j == 0 && !pi.bc[2], fn = prep && j == my - 1 && !pi.bc[3];
  const int iw = pi.gbase + j, ie = pi.gbase + my + j;
  const int is = pi.gbase + 2 * my + i, in_ = pi.gbase + 2 * my + mx + i;
  int pw = -1, pe = -1, ps = -1, pn = -1;
  if (fw) { keep_face_cf<T>(L, Pq, iw); pw = L.aux[iw].push; }
  if (fe) { keep_face_cf<T>(L, Pq, ie); pe = L.aux[ie].push; }
  if (fs) { keep_face_cf<T>(L, Pq, is); ps = L.aux[is].push; }
  if (fn) { keep_face_cf<T>(L, Pq, in_); pn = L.aux[in_].push; }
  if (prep) {  // and those of the corner records, from the corner cells (a region launch's ghosts read them)
    const int c0 = pi.gbase + 2 * my + 2 * mx;
    if (i == 0 && j == 0) keep_face_cf<T>(L, Pq, c0);
    if (i == mx - 1 && j == 0) keep_face_cf<T>(L, Pq, c0 + 1);
    if (i == 0 && j == my - 1) keep_face_cf<T>(L, Pq, c0 + 2);
    if (i == mx - 1 && j == my - 1) keep_face_cf<T>(L, Pq, c0 + 3);
  }
  const int ldc = mx / 2 + 2;
  const int w0 = pi.qcoff + ((i >> 1) + 1) + ((j >> 1) + 1) * ldc;
  const int dx = (i & 1) ? 1 : -1;
  const int dy = (j & 1) ? ldc : -ldc;
  const int* __restrict__ wmap = L.wmap;
  const T* __restrict__ QC = L.QC;
  T e_c, e_x, e_y, e_d;
  if (init) {
    e_c = Pq[wmap[w0]];
    e_x = Pq[wmap[w0 + dx]];
    e_y = Pq[wmap[w0 + dy]];
    e_d = Pq[wmap[w0 + dx + dy]];
  } else {
    e_c = Pq[wmap[w0]] - QC[w0];
    e_x = Pq[wmap[w0 + dx]] - QC[w0 + dx];
    e_y = Pq[wmap[w0 + dy]] - QC[w0 + dy];
    e_d = Pq[wmap[w0 + dx + dy]] - QC[w0 + dx + dy];
  }
  const T p = (((T(9) * e_c + T(3) * e_x) + T(3) * e_y) + e_d) / T(16);
  const int qc = pi.qoff + (i + 1) + (j + 1) * (mx + 2);
  const T y = init ? p : L.Q[qc] + p;
  L.Q[qc] = y;
  if (Qdup) Qdup[qc] = y;  // the level's second array (streaming sweeps)
  if (pw >= 0) L.Q[pw] = y;
  if (pe >= 0) L.Q[pe] = y;
  if (ps >= 0) L.Q[ps] = y;
  if (pn >= 0) L.Q[pn] = y;
}

// ---------------------------------------------------------------- kernels
// The level descriptor's fields a pass reads, requested together at kernel
// entry: an empty asm statement that takes them as scalar operands makes the
// compiler issue their kernel-argument loads at once instead of each after
// the branch that first needs it (one round trip instead of several).
template <typename T>
__device__ __forceinline__ void relax_args_early(const LevelDev<T>& L) {
  asm volatile("" ::"s"(L.tiles), "s"(L.Q), "s"(L.F), "s"(L.ID), "s"(L.DeW), "s"(L.DnS), "s"(L.G), "s"(L.aux));
  asm volatile("" ::"s"(L.h[0]), "s"(L.h[1]), "s"(L.ih2[0]), "s"(L.ih2[1]), "s"(L.pow2), "s"(L.ghost));
}

// one colour pass, the mode fixed at compile time so that every variant is
// straight-line code with all of a cell's loads issued together
template <typename T, int MODE>
__global__ void __launch_bounds__(kTile) k_relax(LevelDev<T> L, const T* __restrict__ Pq, int colour,
                                                 T omega, int homog) {
  relax_args_early<T>(L);
  asm volatile("" ::"s"(Pq), "s"(colour), "s"(omega), "s"(homog));
  relax_item<T>(L, Pq, blockIdx.x, threadIdx.x, colour, omega, homog, MODE);
}

// the same over the colour's cells only (every patch of even width): one
// 128-lane workgroup per tile, for passes in which the other colour has
// nothing to do
template <typename T, int MODE>
__global__ void __launch_bounds__(kTile / 2) k_relax_c(LevelDev<T> L, const T* __restrict__ Pq, int colour,
                                                       T omega, int homog) {
  relax_args_early<T>(L);
  asm volatile("" ::"s"(Pq), "s"(colour), "s"(omega), "s"(homog));
  relax_item_colour<T>(L, Pq, blockIdx.x * (kTile / 2) + threadIdx.x, colour, omega, homog, MODE);
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_fill(LevelDev<T> L, const T* __restrict__ Pq, int homog) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t < fill_items<T>(L, Pq)) fill_item<T>(L, Pq, t, homog);
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_ring(LevelDev<T> L, const T* __restrict__ Pq) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t < L.nring_all) ring_item<T>(L, Pq, t);
}

// the QC ring refresh of ebFillGhosts (patches with coarse-fine faces) alone
template <typename T>
__global__ void __launch_bounds__(kTile) k_ring_cf(LevelDev<T> L, const T* __restrict__ Pq) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t < L.nring_cf) {
    const RingRec rr = L.ring_cf[t];
    L.QC[rr.dst] = Pq[rr.src];
  }
}

// A filldown step (ebTransfer.m:73-79: restrict level k, fill level k-1) as
// one launch: items [0, nfoot) restrict F into P and F's QC interior; the
// rest fill P's ghosts (and QC ring), reading P's covered cells as the
// restriction of F's values (CoveredQ) rather than P's cells the first items
// write, so nothing in the launch reads what another item writes
// C, nring, ncf: the composite residual's QC refresh (ebCompResidual.m:23-39)
// of F's child C, whose parent F is final after the previous step: its ring
// (nring items) and, for the top level, its coarse-fine parent parts (ncf
// items).  Nothing else in the launch writes C or reads what these write.
template <typename T>
__global__ void __launch_bounds__(kTile) k_restrict_fill(LevelDev<T> F, LevelDev<T> P, const T* __restrict__ PPq,
                                                         int homog, LevelDev<T> C, int nring, int ncf) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t < F.nfoot) {
    restrict_item<T>(F, P.Q, t, 0);
    return;
  }
  const int nf = fill_items<T>(P, PPq), x = t - F.nfoot;
  if (x < nf) {
    fill_item_via<T>(P, PPq, x, homog, CoveredQ<T>{P.Q, F.Q, F.pcov});
  } else if (x < nf + nring) {
    ring_item<T>(C, F.Q, x - nf);
  } else if (x < nf + nring + ncf) {
    keep_face_cf<T>(C, F.Q, x - nf - nring);
  }
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_restrict(LevelDev<T> L, T* __restrict__ Pq, int qc_only,
                                                    T* __restrict__ Pdup) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t < L.nfoot) restrict_item<T>(L, Pq, t, qc_only, Pdup);
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_coarse_load(LevelDev<T> L, LevelDev<T> P) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t < L.nfoot) coarse_load_item<T>(L, P, t);
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_reflux(LevelDev<T> L, LevelDev<T> P) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t < P.ntgt) reflux_item<T>(L, P, t, 0);
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_load(LevelDev<T> L, LevelDev<T> P, int mass_corr, int pghost, int homog,
                                                const T* __restrict__ PPq, int from_ft) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  const int n = load_items<T>(L, P, mass_corr);
  if (t < n) {
    load_item<T>(L, P, t, mass_corr, pghost, homog, from_ft);
  } else if (PPq && t < n + P.nghost) {
    // the parent's parent parts from its own parent (what the left-out fill
    // of the parent stores): the parent's pre-smoothing assembles with them
    keep_face_cf<T>(P, PPq, t - n);
  }
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_correct(LevelDev<T> L, const T* __restrict__ Pq, int init, int prep,
                                                   T* __restrict__ Qdup) {
  correct_item<T>(L, Pq, blockIdx.x, threadIdx.x, init, prep, Qdup);
}

// ---------------------------------------------------------------- coarsest
// u = Ainv (F + bcRhs), rows accumulated in ascending column order; the
// caller has staged rhs in LDS.  Returns the value stored.
template <typename T>
__device__ __forceinline__ T coarsest_row(const CoarsestDev& C, const double* rhs, int i, int n1,
                                          T* __restrict__ Q) {
  // ascending-j accumulation as in the oracle; the loads of 16 columns are
  // issued before their products are added so that they overlap
  const double* __restrict__ a = C.Ainv + i;
  const size_t M = (size_t)C.M;
  double acc = 0.0;
  int j = 0;
  for (; j + 16 <= C.M; j += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = a[(size_t)(j + u) * M];
#pragma unroll
    for (int u = 0; u < 16; ++u) acc = acc + v[u] * rhs[j + u];
  }
  for (; j < C.M; ++j) acc = acc + a[(size_t)j * M] * rhs[j];
  const int ci = i % n1, cj = i / n1;
  Q[(ci + 1) + (cj + 1) * (n1 + 2)] = (T)acc;
  return (T)acc;
}

// one wave per block of the coarsest apply
constexpr int kCoarsestThreads = 64;

// With C.gtab, the block (one wave) also writes the level-1 ghost records
// whose sources are its rows, as the fill after the solve would
// (ebFillGhosts): the rows' values come from the lanes that computed them
// (shuffles every lane takes part in), the expressions are region_ghost's and
// fill_item's corner rule.
template <typename T>
__global__ void __launch_bounds__(kCoarsestThreads) k_coarsest(CoarsestDev C, LevelDev<T> L1, int homog) {
  extern __shared__ double rhs[];
  const int g0 = C.gtab ? C.gofs[blockIdx.x] : 0, g1 = C.gtab ? C.gofs[blockIdx.x + 1] : 0;
  for (int j = threadIdx.x; j < C.M; j += blockDim.x)
    rhs[j] = j < C.N ? (homog ? (double)L1.F[j] : (double)L1.F[j] + C.bcRhs[j]) : 0.0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  T qv = T(0);
  if (i < C.N) qv = coarsest_row<T>(C, rhs, i, L1.n[0], L1.Q);
  const int base = blockIdx.x * blockDim.x;
  for (int e0 = g0; e0 < g1; e0 += blockDim.x) {  // uniform in the block
    const int lane = threadIdx.x, e = min(e0 + lane, g1 - 1);
    const CoarsestGhost G = C.gtab[e];
    T v[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] = __shfl(qv, G.rows[k] - base, 64);
    const GhostRec r = L1.ghost[G.rec];
    T val;
    if (r.kind == G_PCORNER) {
      const GhostRec A = L1.ghost[r.s[0]], B = L1.ghost[r.s[1]];
      const T a = region_ghost<T>(A.kind, T(A.c), T(0), v[0], v[1], homog);
      const T b = region_ghost<T>(B.kind, T(B.c), T(0), v[2], v[3], homog);
      val = (a + b) - v[4];
    } else {
      val = region_ghost<T>(r.kind, T(r.c), T(0), v[0], v[1], homog);
    }
    if (e0 + lane < g1) L1.Q[r.dst] = val;
  }
}

// ---------------------------------------------------------------- groups
// A group of workgroups that runs one hierarchy's steps together: the items
// of a step are strided over the group's threads and dependent steps are
// separated by sync().  With one workgroup sync() is __syncthreads().  With
// several, thread 0 of each workgroup releases the workgroup's writes (an
// agent-scope fence after the workgroup barrier), counts in on the group's
// counter and polls it until every workgroup of the group has arrived, then
// acquires the others' writes.  The wait is bounded: after kGrpSpinLimit
// polls it sets *err and stops waiting (as does every later wait once *err is
// set), so a launch whose workgroups were not all resident ends with an error
// instead of hanging.  The counter is zero when the launch starts; target
// grows by nwg per barrier.
constexpr unsigned kGrpSpinLimit = 1u << 22;
struct Grp {
  int rank, size;  // thread index in the group, threads in the group
  int wg, nwg;     // workgroup index in the group, workgroups in the group
  unsigned* ctr;
  int* err;
  unsigned target;
  __device__ static Grp block() { return Grp{(int)threadIdx.x, (int)blockDim.x, 0, 1, nullptr, nullptr, 0u}; }
  __device__ void sync() {
    __syncthreads();
    if (nwg > 1) {
      if (threadIdx.x == 0) {
        target += (unsigned)nwg;
        __threadfence();
        __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned polls = 0;
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
          if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
          if (++polls > kGrpSpinLimit) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        __threadfence();
      }
      __syncthreads();
    }
  }
};

// ---------------------------------------------------------------- coarse cycle
// The V-cycle of levels 1..kc inside one workgroup (ebVcycle.m:15-35 for
// k <= kc): every step is a workgroup-strided loop over the step's items,
// separated by barriers.  The descent's smoother reads the materialised
// ghosts in its first pass; after the correction the fill is folded into the
// first post-smoothing pass (inline ghosts), as in the multi-kernel form.
template <typename T>
__device__ void blk_relax(Grp& g, const LevelDev<T>& L, const T* Pq, int nsweeps, T omega, int homog, int first_inl) {
  const int ncell = L.ntile * kTile;
  for (int it = 0; it < nsweeps; ++it)
    for (int colour = 1; colour <= 2; ++colour) {
      // one specialised loop per ghost source (frame / records)
      const int inl = (it == 0 && colour == 1) ? first_inl : 1;
      if (L.even_mx) {
        if (inl) {
          for (int x = g.rank; x < (ncell >> 1); x += g.size)
            relax_item_colour<T, false>(L, Pq, x, colour, omega, homog, kInlRecords);
        } else {
          for (int x = g.rank; x < (ncell >> 1); x += g.size)
            relax_item_colour<T, false>(L, Pq, x, colour, omega, homog, 0);
        }
      } else {
        if (inl) {
          for (int x = g.rank; x < ncell; x += g.size)
            relax_item<T, false>(L, Pq, x / kTile, x % kTile, colour, omega, homog, kInlRecords);
        } else {
          for (int x = g.rank; x < ncell; x += g.size)
            relax_item<T, false>(L, Pq, x / kTile, x % kTile, colour, omega, homog, 0);
        }
      }
      g.sync();
    }
  if (nsweeps > 0) {
    const int n = fill_items<T>(L, Pq);
    for (int x = g.rank; x < n; x += g.size) fill_item<T>(L, Pq, x, homog);
    g.sync();
  }
}

template <typename T>
__device__ void blk_fill(Grp& g, const LevelDev<T>& L, const T* Pq, int homog) {
  const int n = fill_items<T>(L, Pq);
  for (int x = g.rank; x < n; x += g.size) fill_item<T>(L, Pq, x, homog);
  g.sync();
}

// ebVcycle.m:15-35 for levels kc..1 inside one workgroup; rhs is LDS scratch
// of at least C.M doubles for the coarsest solve
template <typename T>
__device__ void blk_vcycle(Grp& g, const LevelDev<T>* __restrict__ LV, int kc, const CoarsestDev& C, int npre,
                           int npost, T omega, int homog, int mass_corr, double* rhs, int from_ft = 0) {
  for (int k = kc; k >= 2; --k) {
    const LevelDev<T>& L = LV[k - 1];
    const LevelDev<T>& P = LV[k - 2];
    blk_relax<T>(g, L, P.Q, npre, omega, homog, 0);
    for (int x = g.rank; x < L.nfoot; x += g.size) restrict_item<T>(L, P.Q, x, 0);
    g.sync();
    blk_fill<T>(g, P, k >= 3 ? (const T*)LV[k - 3].Q : (const T*)nullptr, homog);
    const int n = load_items<T>(L, P, mass_corr);
    for (int x = g.rank; x < n; x += g.size) load_item<T>(L, P, x, mass_corr, 0, 0, from_ft);
    g.sync();
  }
  {
    // the coarsest solve in the group's first workgroup (rhs is its LDS)
    const LevelDev<T>& L1 = LV[0];
    if (g.wg == 0) {
      for (int j = threadIdx.x; j < C.M; j += blockDim.x)
        rhs[j] = j < C.N ? (homog ? (double)L1.F[j] : (double)L1.F[j] + C.bcRhs[j]) : 0.0;
      __syncthreads();
      for (int i = threadIdx.x; i < C.N; i += blockDim.x) coarsest_row<T>(C, rhs, i, L1.n[0], L1.Q);
    }
    g.sync();
    blk_fill<T>(g, L1, (const T*)nullptr, homog);
  }
  for (int k = 2; k <= kc; ++k) {
    const LevelDev<T>& L = LV[k - 1];
    const T* Pq = LV[k - 2].Q;
    const int ncell = L.ntile * kTile;
    for (int x = g.rank; x < ncell; x += g.size) correct_item<T>(L, Pq, x / kTile, x % kTile, 0);
    g.sync();
    if (npost > 0) {
      blk_relax<T>(g, L, Pq, npost, omega, homog, 1);
    } else {
      blk_fill<T>(g, L, Pq, homog);
    }
  }
}

// The bottom of the filldown (ebTransfer.m:73-79) in one workgroup:
// restrict k -> k-1, fill k-1, for k = kfrom..kto (kto >= 2), with barriers
// between the dependent steps (the small levels' steps would otherwise each
// cost a launch).
// kqc..ktop: levels whose QC ring (and, for ktop == cftop, parent parts) the
// composite residual's refresh needs; step k gathers level k+1's from level
// k, final after the step before (kqc > ktop: none).
template <typename T>
__global__ void __launch_bounds__(1024) k_filldown_small(const LevelDev<T>* __restrict__ LV, int kfrom, int kto,
                                                          int homog, int kqc, int ktop, int cftop) {
  Grp g = Grp::block();
  for (int k = kfrom; k >= kto; --k) {
    const LevelDev<T>& L = LV[k - 1];
    const LevelDev<T>& P = LV[k - 2];
    for (int x = threadIdx.x; x < L.nfoot; x += blockDim.x) restrict_item<T>(L, P.Q, x, 0);
    if (k + 1 >= kqc && k + 1 <= ktop) {
      const LevelDev<T>& C = LV[k];
      for (int x = threadIdx.x; x < C.nring_all; x += blockDim.x) ring_item<T>(C, L.Q, x);
      if (k + 1 == cftop)
        for (int x = threadIdx.x; x < C.nghost; x += blockDim.x) keep_face_cf<T>(C, L.Q, x);
    }
    __syncthreads();
    blk_fill<T>(g, P, k >= 3 ? (const T*)LV[k - 3].Q : (const T*)nullptr, homog);
  }
}

__device__ __forceinline__ void blk_copy_words(void* dst, const void* src, long long nbytes) {
  const uint32_t* __restrict__ a = (const uint32_t*)src;
  uint32_t* __restrict__ b = (uint32_t*)dst;
  const long long n = nbytes >> 2;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) b[i] = a[i];
}

// Copy segments (src, dst, 4-byte words) as one workgroup-wide loop with
// several loads in flight per thread, so that many small arrays cost about
// one memory latency rather than one each.  Each segment starts on a chunk of
// 8 words (first is a multiple of 8) and both its ends are 16-byte aligned,
// so a chunk inside a segment moves as two 16-byte loads and stores.
struct CopySeg {
  const uint32_t* src;
  uint32_t* dst;
  int words;
  int first;  // chunk-aligned start in the concatenated word space
};
constexpr int kMaxCopySegs = 24 * kStageMaxLevels;

__device__ __forceinline__ int copy_seg_words(int words) { return (words + 7) & ~7; }

// last segment with first <= w: a fixed number of halving steps, so that the
// searches of several chunks interleave
__device__ __forceinline__ int copy_seg_find(const CopySeg* seg, int nseg, int w) {
  int lo = 0;
#pragma unroll
  for (int step = 128; step > 0; step >>= 1) {
    const int m = lo + step;
    if (m < nseg && seg[m].first <= w) lo = m;
  }
  return lo;
}

__device__ void blk_copy_segments(const CopySeg* seg, int nseg, int total) {
  static_assert(kMaxCopySegs <= 256, "copy_seg_find covers 256 segments");
  constexpr int U = 4;
  const int nchunk = total >> 3;
  for (int base = threadIdx.x; base < nchunk; base += U * (int)blockDim.x) {
    uint4 v0[U], v1[U];
    int where[U], off[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = base + u * (int)blockDim.x;
      where[u] = -1;
      off[u] = 0;
      if (c < nchunk) {
        const int k = copy_seg_find(seg, nseg, c << 3);
        where[u] = k;
        off[u] = (c << 3) - seg[k].first;
        if (off[u] + 8 <= seg[k].words) {
          const uint4* q = reinterpret_cast<const uint4*>(seg[k].src + off[u]);
          v0[u] = q[0];
          v1[u] = q[1];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (where[u] < 0) continue;
      const CopySeg& S = seg[where[u]];
      if (off[u] + 8 <= S.words) {
        uint4* d = reinterpret_cast<uint4*>(S.dst + off[u]);
        d[0] = v0[u];
        d[1] = v1[u];
      } else {
        for (int w = off[u]; w < S.words; ++w) S.dst[w] = S.src[w];  // the segment's last partial chunk
      }
    }
  }
}

// Copy the arrays of levels 1..kc into LDS (layout of stage_level_bytes) and
// publish descriptors pointing at the copies in LS.
template <typename T>
__device__ void blk_stage_levels(const LevelDev<T>* __restrict__ LV, int kc, LevelDev<T>* LS, char* base,
                                 CopySeg* seg, int* nseg_total) {
  for (int k = 0; k < kc; ++k) blk_copy_words(&LS[k], &LV[k], (long long)sizeof(LevelDev<T>));
  __syncthreads();
  if (threadIdx.x == 0) {  // layout: one segment per staged array
    long long off = 0;
    int ns = 0, words = 0;
    for (int k = 0; k < kc; ++k) {
      const LevelDev<T>& L = LV[k];
      LS[k].aux = nullptr;  // the workgroup passes evaluate ghosts from records
#define BSAM_STAGE_SEG(field, type, count)                             \
  {                                                                    \
    const long long nb = (long long)(count) * (long long)sizeof(type); \
    if (nb > 0) {                                                      \
      seg[ns] = CopySeg{(const uint32_t*)L.field, (uint32_t*)(base + off), (int)(nb >> 2), words}; \
      words += copy_seg_words((int)(nb >> 2));                         \
      LS[k].field = (type*)(base + off);                               \
      off += stage_align(nb);                                          \
      ++ns;                                                            \
    }                                                                  \
  }
      BSAM_STAGE_FIELDS(BSAM_STAGE_SEG, L, T)
      if (k + 1 < kc) {  // level k+1's reflux targets on level k
        const int fine_ncorr = LV[k + 1].ncorr;
        BSAM_STAGE_SEG(tgt_cell, int, L.ntgt)
        BSAM_STAGE_SEG(tgt_ptr, int, L.ntgt + 1)
        BSAM_STAGE_SEG(tgt_ent, int, fine_ncorr)
      }
#undef BSAM_STAGE_SEG
    }
    nseg_total[0] = ns;
    nseg_total[1] = words;
  }
  __syncthreads();
  blk_copy_segments(seg, nseg_total[0], nseg_total[1]);
  __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(1024) k_coarse_cycle(const LevelDev<T>* __restrict__ LV, int kc,
                                                       CoarsestDev C, int npre, int npost, T omega,
                                                       int homog, int mass_corr, int staged, int from_ft) {
  extern __shared__ __attribute__((aligned(16))) double smem[];
  double* rhs = smem;
  Grp g = Grp::block();
  if (!staged) {
    blk_vcycle<T>(g, LV, kc, C, npre, npost, omega, homog, mass_corr, rhs, from_ft);
    return;
  }
  // levels 1..kc live in LDS for the whole cycle; the cycle writes Q, F and QC
  __shared__ LevelDev<T> LS[kStageMaxLevels];
  __shared__ CopySeg seg[kMaxCopySegs];
  __shared__ int nseg_total[2];
  char* base = reinterpret_cast<char*>(smem) + stage_align((long long)C.M * (long long)sizeof(double));
  blk_stage_levels<T>(LV, kc, LS, base, seg, nseg_total);
  blk_vcycle<T>(g, LS, kc, C, npre, npost, omega, homog, mass_corr, rhs, from_ft);
  __syncthreads();
  if (threadIdx.x == 0) {  // write back what the cycle changes: Q, F, QC
    int ns = 0, words = 0;
    for (int k = 0; k < kc; ++k) {
      const LevelDev<T>& G = LV[k];
      const LevelDev<T>& S = LS[k];
      const int wq = (int)((long long)G.nq * (long long)sizeof(T) >> 2);
      const int wf = (int)((long long)G.ncell * (long long)sizeof(T) >> 2);
      const int wc = (int)((long long)G.nqc * (long long)sizeof(T) >> 2);
      seg[ns++] = CopySeg{(const uint32_t*)S.Q, (uint32_t*)G.Q, wq, words};
      words += copy_seg_words(wq);
      seg[ns++] = CopySeg{(const uint32_t*)S.F, (uint32_t*)G.F, wf, words};
      words += copy_seg_words(wf);
      if (wc) {
        seg[ns++] = CopySeg{(const uint32_t*)S.QC, (uint32_t*)G.QC, wc, words};
        words += copy_seg_words(wc);
      }
    }
    nseg_total[0] = ns;
    nseg_total[1] = words;
  }
  __syncthreads();
  blk_copy_segments(seg, nseg_total[0], nseg_total[1]);
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}

// block-level reduction of (max, sum, max, count) into out (thread 0 writes)
__device__ __forceinline__ void block_reduce4(double a, double b, double c, long long n, ResidualOut* out);

// block-level reduction of (max, sum, max, count) into out (thread 0 writes)
__device__ __forceinline__ void block_reduce4(double a, double b, double c, long long n,
                                              ResidualOut* out) {
  __shared__ double sa[kTile / 64], sb[kTile / 64], sc[kTile / 64];
  __shared__ long long sn[kTile / 64];
  a = wave_max(a);
  b = wave_sum(b);
  c = wave_max(c);
  n = wave_sum_ll(n);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sa[w] = a; sb[w] = b; sc[w] = c; sn[w] = n;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ra = sa[0], rb = sb[0], rc = sc[0];
    long long rn = sn[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
      ra = fmax(ra, sa[k]); rb = rb + sb[k]; rc = fmax(rc, sc[k]); rn += sn[k];
    }
    out->linf = ra; out->l2sq = rb; out->fnorm = rc; out->ncell = rn;
  }
}

// QC refresh of every level k >= kmin in one launch (ebCompResidual.m:23-39):
// ring gather and QC interior = R(u) (no parent scatter).  Items of the
// combined table: (level, item) with item < nring_all + nfoot.
template <typename T>
__global__ void __launch_bounds__(kTile) k_qc_refresh_all(const LevelDev<T>* __restrict__ LV,
                                                          const int2* __restrict__ items, int nitems) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t >= nitems) return;
  const int2 it = items[t];
  const LevelDev<T>& L = LV[it.x];
  if (it.y < L.nring_all) {
    ring_item<T>(L, LV[it.x - 1].Q, it.y);
  } else if (it.y < L.nring_all + L.nfoot) {
    restrict_item<T>(L, (T*)nullptr, it.y - L.nring_all, 1);
  } else {  // a coarse-fine record's parent part (the top level's, see qc_items)
    keep_face_cf<T>(L, LV[it.x - 1].Q, it.y - L.nring_all - L.nfoot);
  }
}

// one cell of the composite residual: r = FTrue - L(Q) + reflux chunks
// (ebCompResidual.m:54-78).  Accumulates max|r|, sum r^2 da, max|FTrue| and
// the visible count; optionally writes r to rout[level].
template <typename T, typename TO>
__device__ __forceinline__ void residual_item(const LevelDev<T>* __restrict__ LV, int nlev, int lvl, int tile,
                                              int lane, int mass_corr, TO* const* __restrict__ rout,
                                              int k0, double& a, double& b, double& f, long long& n) {
  const LevelDev<T>& L = LV[lvl];
  const TileRec& tr = L.tiles[tile];
  const PatchInfo& pi = tr.pi;
  const int mx = pi.mx, my = pi.my;
  const int c = tr.start + lane;
  if (c >= mx * my) return;
  const int cc = pi.coff + c;
  if (lvl + 1 < k0) return;
  const int j = cell_row(c, mx, pi.magic);
  const int i = c - j * mx;
  const int ldq = mx + 2;
  const int qc = pi.qoff + (i + 1) + (j + 1) * ldq;
  const int fx = pi.xoff + i + j * (mx + 1);
  const int fy = pi.yoff + i + j * mx;
  T r = L.FT[cc] - op_at<T>(L.Q, qc, ldq, L.DeW[fx], L.DeW[fx + 1], L.DnS[fy], L.DnS[fy + mx], L.CC[cc],
                            level_h2<T>(L));
  double rd = (double)r;
  if (mass_corr && lvl + 1 < nlev) {
    const int row = L.tgt_of_cell[cc];
    if (row >= 0) {
      const LevelDev<T>& Fn = LV[lvl + 1];
      const H2<double> hc = level_h2_64<T>(L);
      for (int e = L.tgt_ptr[row]; e < L.tgt_ptr[row + 1]; ++e) {
        const CorrRec cr = Fn.corr[L.tgt_ent[e]];
        rd = rd + corr_value<T>(cr, Fn.Q, (double)Fn.QC[cr.ring], Fn.QC, Fn.DeW, Fn.DnS, hc);
      }
    }
  }
  if (rout && rout[lvl]) rout[lvl][cc] = (TO)rd;
  if (L.vis[cc]) {
    a = fmax(a, fabs(rd));
    b = b + rd * rd * (L.h[0] * L.h[1]);
    f = fmax(f, fabs((double)L.FT[cc]));
    n += 1;
  }
}

// composite residual of every level in [k0, nlev] (ebCompResidual.m:54-78);
// optionally writes r (the defect-correction right side).  The tile table is
// (level, tile) pairs; a workgroup takes every gridDim.x-th tile and writes one
// part, so the final reduction reads at most kResidualParts parts.
template <typename T, typename TO>
__global__ void __launch_bounds__(kTile) k_residual_all(const LevelDev<T>* __restrict__ LV, int nlev,
                                                        const int2* __restrict__ tl, int ntl, int mass_corr,
                                                        ResidualOut* __restrict__ part,
                                                        TO* const* __restrict__ rout, int k0) {
  double a = 0.0, b = 0.0, f = 0.0;
  long long n = 0;
  for (int t = blockIdx.x; t < ntl; t += gridDim.x) {
    const int2 lt = tl[t];
    residual_item<T, TO>(LV, nlev, lt.x, lt.y, (int)threadIdx.x, mass_corr, rout, k0, a, b, f, n);
  }
  block_reduce4(a, b, f, n, &part[blockIdx.x]);
}

// The parts reduced in a fixed order by one workgroup (a separate launch: the
// kernel boundary orders the parts' writes more cheaply than device-scope
// fences in every workgroup would); out_host (pinned host memory) also
// receives the result when given, so no copy follows.
// 1024 lanes, each summing its parts (every 1024th, in order) eight loads at a
// time, then waves and the workgroup in a fixed order: the result is the
// same on every run.
constexpr int kReduceThreads = 1024;
// seq (with out_host): a device counter, incremented per launch and written
// into out_host->pad2[0] after the scalars (system-scope fence between), so
// that the host can wait for the result by polling the pinned buffer
__global__ void __launch_bounds__(kReduceThreads) k_reduce_parts(const ResidualOut* __restrict__ parts,
                                                                 int nparts, ResidualOut* __restrict__ out,
                                                                 ResidualOut* __restrict__ out_host,
                                                                 unsigned long long* __restrict__ seq) {
  double a = 0.0, b = 0.0, f = 0.0;
  long long n = 0;
  for (int k0 = threadIdx.x; k0 < nparts; k0 += 8 * kReduceThreads) {
    ResidualOut p[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = k0 + u * kReduceThreads;
      if (k < nparts) {
        p[u] = parts[k];
      } else {
        p[u].linf = p[u].l2sq = p[u].fnorm = 0.0;
        p[u].ncell = 0;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a = fmax(a, p[u].linf);
      b = b + p[u].l2sq;
      f = fmax(f, p[u].fnorm);
      n += p[u].ncell;
    }
  }
  __shared__ double sa[kReduceThreads / 64], sb[kReduceThreads / 64], sc[kReduceThreads / 64];
  __shared__ long long sn[kReduceThreads / 64];
  a = wave_max(a);
  b = wave_sum(b);
  f = wave_max(f);
  n = wave_sum_ll(n);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sa[w] = a; sb[w] = b; sc[w] = f; sn[w] = n;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ResidualOut r;
    r.linf = sa[0]; r.l2sq = sb[0]; r.fnorm = sc[0]; r.ncell = sn[0];
    for (int k = 1; k < kReduceThreads / 64; ++k) {
      r.linf = fmax(r.linf, sa[k]); r.l2sq = r.l2sq + sb[k]; r.fnorm = fmax(r.fnorm, sc[k]); r.ncell += sn[k];
    }
    r.pad = 0.0;
    r.pad2[0] = r.pad2[1] = r.pad2[2] = 0;
    *out = r;
    if (out_host) {  // the scalars, then (seq) the flag; pad2 is the flag's word
      out_host->linf = r.linf;
      out_host->l2sq = r.l2sq;
      out_host->fnorm = r.fnorm;
      out_host->ncell = r.ncell;
      if (seq) {
        const unsigned long long v = *seq + 1ULL;
        *seq = v;
        __threadfence_system();
        *(volatile long long*)&out_host->pad2[0] = (long long)v;
      }
    }
  }
}

// F = FTrue on every level (ebSolveCycles.m:30-34); blockIdx.y is the level
template <typename T>
__global__ void __launch_bounds__(kTile) k_reset_force(const LevelDev<T>* __restrict__ LV) {
  const LevelDev<T>& L = LV[blockIdx.y];
  const int n = L.ncell;
  T* __restrict__ F = L.F;
  const T* __restrict__ FT = L.FT;
  for (int x = blockIdx.x * kTile + threadIdx.x; x < n; x += gridDim.x * kTile) F[x] = FT[x];
}

// every level back to its initial state in one launch: Q = Q0, F = FTrue,
// QC = 0 (blockIdx.y: the level)
template <typename T>
__global__ void __launch_bounds__(kTile) k_reset_levels(const LevelDev<T>* __restrict__ LV,
                                                        const T* const* __restrict__ Q0) {
  const LevelDev<T>& L = LV[blockIdx.y];
  const T* __restrict__ q0 = Q0[blockIdx.y];
  const int n = max(L.nq, max(L.ncell, L.nqc));
  for (int x = blockIdx.x * kTile + threadIdx.x; x < n; x += gridDim.x * kTile) {
    if (x < L.nq) L.Q[x] = q0[x];
    if (x < L.ncell) L.F[x] = L.FT[x];
    if (x < L.nqc) L.QC[x] = T(0);
  }
}

// composite mean: sum of visible u * da (ebCompositeMean.m:4-23)
template <typename T>
__global__ void __launch_bounds__(kTile) k_mass(LevelDev<T> L, ResidualOut* __restrict__ part) {
  const TileRec& tr = L.tiles[blockIdx.x];
  const PatchInfo& pi = tr.pi;
  const int mx = pi.mx, my = pi.my;
  const int c = tr.start + (int)threadIdx.x;
  double s = 0.0;
  long long n = 0;
  if (c < mx * my && L.vis[pi.coff + c]) {
    const int j = cell_row(c, mx, pi.magic);
    const int i = c - j * mx;
    s = (double)L.Q[pi.qoff + (i + 1) + (j + 1) * (mx + 2)];
    n = 1;
  }
  block_reduce4(0.0, s * (L.h[0] * L.h[1]), 0.0, n, &part[blockIdx.x]);
}

// ---------------------------------------------------------------- batched cycles
// Workgroup-wide forms of the solve-stage steps, used by the batched kernel:
// one workgroup carries one hierarchy through a whole stage step, so any
// number of independent hierarchies advance in one launch and each one's
// arithmetic is exactly that of its solo run.
__device__ __forceinline__ void block_reduce4_wg(double a, double b, double c, long long n, ResidualOut* out) {
  __shared__ double sa[16], sb[16], sc[16];
  __shared__ long long sn[16];
  a = wave_max(a);
  b = wave_sum(b);
  c = wave_max(c);
  n = wave_sum_ll(n);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sa[w] = a; sb[w] = b; sc[w] = c; sn[w] = n;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ra = sa[0], rb = sb[0], rc = sc[0];
    long long rn = sn[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
      ra = fmax(ra, sa[k]); rb = rb + sb[k]; rc = fmax(rc, sc[k]); rn += sn[k];
    }
    out->linf = ra; out->l2sq = rb; out->fnorm = rc; out->ncell = rn;
  }
  __syncthreads();
}

template <typename T>
__device__ void blk_reset_force(Grp& g, const LevelDev<T>* __restrict__ LV, int nlev) {
  for (int k = 0; k < nlev; ++k) {
    const LevelDev<T>& L = LV[k];
    T* __restrict__ F = L.F;
    const T* __restrict__ FT = L.FT;
    const int n = L.ncell;
    for (int x = g.rank; x < n; x += g.size) F[x] = FT[x];
  }
  g.sync();
}

template <typename T>
__device__ void blk_filldown(Grp& g, const LevelDev<T>* __restrict__ LV, int nlev, int homog) {
  for (int k = nlev; k >= 2; --k) {  // ebTransfer.m:73-79
    const LevelDev<T>& L = LV[k - 1];
    const LevelDev<T>& P = LV[k - 2];
    for (int x = g.rank; x < L.nfoot; x += g.size) restrict_item<T>(L, P.Q, x, 0);
    g.sync();
    blk_fill<T>(g, P, k >= 3 ? (const T*)LV[k - 3].Q : (const T*)nullptr, homog);
  }
}

// composite residual (ebCompResidual.m:23-84): QC refresh of the AMR levels,
// then every visible cell of levels >= root
__device__ void blk_residual(Grp& g, const LevelDev<double>* __restrict__ LV, int nlev, int root, int mass_corr,
                             float* const* rout, ResidualOut* out, ResidualOut* parts) {
  if (mass_corr) {
    for (int k = max(2, root + 1); k <= nlev; ++k) {
      const LevelDev<double>& L = LV[k - 1];
      const double* Pq = LV[k - 2].Q;
      const int n = L.nring_all + L.nfoot;
      for (int x = g.rank; x < n; x += g.size) {
        if (x < L.nring_all) ring_item<double>(L, Pq, x);
        else restrict_item<double>(L, (double*)nullptr, x - L.nring_all, 1);
      }
    }
    g.sync();
  }
  double a = 0.0, b = 0.0, f = 0.0;
  long long n = 0;
  for (int k = root; k <= nlev; ++k) {
    const int ncell = LV[k - 1].ntile * kTile;
    for (int x = g.rank; x < ncell; x += g.size)
      residual_item<double, float>(LV, nlev, k - 1, x / kTile, x % kTile, mass_corr, rout, root, a, b, f, n);
  }
  if (g.nwg == 1) {
    block_reduce4_wg(a, b, f, n, out);
    return;
  }
  // per-workgroup partials, combined by the group's first workgroup in
  // workgroup order
  block_reduce4_wg(a, b, f, n, parts + g.wg);
  g.sync();
  if (g.wg == 0 && threadIdx.x == 0) {
    ResidualOut r = parts[0];
    for (int w = 1; w < g.nwg; ++w) {
      r.linf = fmax(r.linf, parts[w].linf);
      r.l2sq = r.l2sq + parts[w].l2sq;
      r.fnorm = fmax(r.fnorm, parts[w].fnorm);
      r.ncell += parts[w].ncell;
    }
    *out = r;
  }
}

// One cycle (or the stage preamble) of every batch item; item i runs on the
// group of workgroups [wg0, wg0 + nwg) of the launch (grp_item maps a
// workgroup to its item) and counts its barriers on ctr[i].  Every
// workgroup of the launch must be resident at once (a cooperative launch).
__global__ void __launch_bounds__(1024) k_batch_cycle(const BatchItem* __restrict__ items,
                                                      const int* __restrict__ grp_item, unsigned* ctr, int* err,
                                                      int npre, int npost, double omega, int mass_corr, int mixed,
                                                      int inner) {
  extern __shared__ double rhs[];
  const int ii = grp_item[blockIdx.x];
  const BatchItem& it = items[ii];
  const int wg = (int)blockIdx.x - it.wg0;
  Grp g{wg * (int)blockDim.x + (int)threadIdx.x, it.nwg * (int)blockDim.x, wg, it.nwg, ctr + ii, err, 0u};
  const LevelDev<double>* LV = it.LV64;
  const int nlev = it.nlev;
  if (it.phase == 0) {  // ebSolveCycles.m:11-19: fill every level, baseline residual
    for (int k = 1; k <= nlev; ++k)
      blk_fill<double>(g, LV[k - 1], k > 1 ? (const double*)LV[k - 2].Q : nullptr, 0);
    blk_residual(g, LV, nlev, it.root, mass_corr, mixed ? it.rout32 : nullptr, it.out, it.parts);
    return;
  }
  if (!mixed) {  // ebSolveCycles.m:28-41
    blk_reset_force<double>(g, LV, nlev);
    blk_vcycle<double>(g, LV, nlev, it.cst, npre, npost, omega, 0, mass_corr, rhs);
    blk_filldown<double>(g, LV, nlev, 0);
    blk_residual(g, LV, nlev, it.root, mass_corr, nullptr, it.out, it.parts);
    return;
  }
  // FP32 V-cycles on the correction problem, FP64 update, filldown, ghosts,
  // residual (which also writes the next FP32 right side)
  const LevelDev<float>* L32 = it.LV32;
  for (int k = 0; k < nlev; ++k) {
    const LevelDev<float>& L = L32[k];
    for (int x = g.rank; x < L.nq; x += g.size) L.Q[x] = 0.0f;
    for (int x = g.rank; x < L.nqc; x += g.size) L.QC[x] = 0.0f;
  }
  g.sync();
  for (int i = 0; i < inner; ++i) {
    blk_reset_force<float>(g, L32, nlev);
    blk_vcycle<float>(g, L32, nlev, it.cst, npre, npost, (float)omega, 1, mass_corr, rhs);
    if (i + 1 < inner) blk_filldown<float>(g, L32, nlev, 1);
  }
  for (int k = 0; k < nlev; ++k) {
    const LevelDev<double>& L = LV[k];
    const float* E = L32[k].Q;
    const int ncell = L.ntile * kTile;
    for (int x = g.rank; x < ncell; x += g.size) {
      const TileRec& tr = L.tiles[x / kTile];
      const int c = tr.start + x % kTile;
      if (c >= tr.pi.mx * tr.pi.my) continue;
      const int j = cell_row(c, tr.pi.mx, tr.pi.magic);
      const int i2 = c - j * tr.pi.mx;
      const int qc = tr.pi.qoff + (i2 + 1) + (j + 1) * (tr.pi.mx + 2);
      L.Q[qc] = L.Q[qc] + (double)E[qc];
    }
  }
  g.sync();
  blk_filldown<double>(g, LV, nlev, 0);
  for (int k = 1; k <= nlev; ++k) blk_fill<double>(g, LV[k - 1], k > 1 ? (const double*)LV[k - 2].Q : nullptr, 0);
  blk_residual(g, LV, nlev, it.root, mass_corr, it.rout32, it.out, it.parts);
}

// ---------------------------------------------------------------- elementwise
// Q64 interior += Q32 interior (defect-correction update)
__global__ void __launch_bounds__(kTile) k_add_correction(LevelDev<double> L, const float* __restrict__ E) {
  const TileRec& tr = L.tiles[blockIdx.x];
  const PatchInfo& pi = tr.pi;
  const int mx = pi.mx, my = pi.my;
  const int c = tr.start + (int)threadIdx.x;
  if (c >= mx * my) return;
  const int j = cell_row(c, mx, pi.magic);
  const int i = c - j * mx;
  const int qc = pi.qoff + (i + 1) + (j + 1) * (mx + 2);
  L.Q[qc] = L.Q[qc] + (double)E[qc];
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_fill_value(T* __restrict__ a, long long n, T v) {
  const long long t = (long long)blockIdx.x * kTile + threadIdx.x;
  if (t < n) a[t] = v;
}

template <typename TD, typename TS>
__global__ void __launch_bounds__(kTile) k_convert(TD* __restrict__ d, const TS* __restrict__ s,
                                                   long long n) {
  const long long t = (long long)blockIdx.x * kTile + threadIdx.x;
  if (t < n) d[t] = (TD)s[t];
}

// BDF2 / BE right-hand side (ParabolicBSAM/src/ebpRhs.m:31-34)
__global__ void __launch_bounds__(kTile) k_bdf2_rhs(double* __restrict__ FT, const double* __restrict__ f,
                                                    const double* __restrict__ Uo,
                                                    const double* __restrict__ Uoo, double dt,
                                                    int first, long long n) {
  const long long t = (long long)blockIdx.x * kTile + threadIdx.x;
  if (t >= n) return;
  const double fv = f ? f[t] : 0.0;
  FT[t] = first ? fv + Uo[t] / dt : fv + (2.0 * Uo[t] - 0.5 * Uoo[t]) / dt;
}

// the pointwise part of the Crank-Nicolson right-hand side, 2 f + c Uo with
// c = 2/dt + aDt (ParabolicBSAM/src/ebpRhs.m:35-38); the operator part follows
__global__ void __launch_bounds__(kTile) k_cn_rhs(double* __restrict__ FT, const double* __restrict__ f,
                                                  const double* __restrict__ Uo, double c, long long n) {
  const long long t = (long long)blockIdx.x * kTile + threadIdx.x;
  if (t >= n) return;
  const double fv = f ? f[t] : 0.0;
  FT[t] = 2.0 * fv + c * Uo[t];
}

// history: rotate 0: Uoo = Uo = interior(Q); 1: Uoo = Uo, Uo = interior(Q);
// 2: Uo = interior(Q), Uoo kept
__global__ void __launch_bounds__(kTile) k_history(LevelDev<double> L, double* __restrict__ Uo,
                                                   double* __restrict__ Uoo, int rotate) {
  const TileRec& tr = L.tiles[blockIdx.x];
  const PatchInfo& pi = tr.pi;
  const int mx = pi.mx, my = pi.my;
  const int c = tr.start + (int)threadIdx.x;
  if (c >= mx * my) return;
  const int j = cell_row(c, mx, pi.magic);
  const int i = c - j * mx;
  const double u = L.Q[pi.qoff + (i + 1) + (j + 1) * (mx + 2)];
  const int cc = pi.coff + c;
  if (rotate != 2) Uoo[cc] = rotate ? Uo[cc] : u;
  Uo[cc] = u;
}

// ---------------------------------------------------------------- regrid transfer
// Conservative-linear child value (ebpRegrid.m:95-108): the parent value with
// a quarter of each slope subtracted (first child along the axis) or added.
__device__ __forceinline__ double conslin_child(double c0, double sx, double sy, int odd_i, int odd_j) {
  const double a = odd_i ? c0 + sx / 4 : c0 - sx / 4;
  return odd_j ? a + sy / 4 : a - sy / 4;
}

// transferLevel (ebpRegrid.m:62-92) for one cell of a new level: Q from the
// conservative-linear prolongation of the parent window (centred slopes), the
// history from that of the parent footprint (one-sided slopes at its edges,
// :111-134), both overwritten by the old hierarchy's same-level values where
// an old patch covers the cell.  oq / oc: the old Q (padded) and history
// index of the cell or -1; fc: per footprint cell, the parent's history index.
__global__ void __launch_bounds__(kTile) k_regrid_level(LevelDev<double> L, const double* __restrict__ Pq,
                                                        int conslin, int hist, double* __restrict__ Uoo_k,
                                                        const double* __restrict__ Uoo_all,
                                                        const int* __restrict__ fc, const int* __restrict__ oq,
                                                        const int* __restrict__ oc, const double* __restrict__ Qold,
                                                        const double* __restrict__ Uoo_old) {
  const TileRec& tr = L.tiles[blockIdx.x];
  const PatchInfo& pi = tr.pi;
  const int mx = pi.mx, my = pi.my;
  const int c = tr.start + (int)threadIdx.x;
  if (c >= mx * my) return;
  const int j = cell_row(c, mx, pi.magic);
  const int i = c - j * mx;
  const int cc = pi.coff + c;
  double q = 0.0, u = 0.0;
  if (conslin) {
    const int ci = i >> 1, cj = j >> 1, ldc = mx / 2 + 2;
    auto w = [&](int a, int b) { return Pq[L.wmap[pi.qcoff + a + b * ldc]]; };
    const double c0 = w(ci + 1, cj + 1);
    const double sx = (w(ci + 2, cj + 1) - w(ci, cj + 1)) / 2;
    const double sy = (w(ci + 1, cj + 2) - w(ci + 1, cj)) / 2;
    q = conslin_child(c0, sx, sy, i & 1, j & 1);
    if (hist) {
      const int cmx = mx / 2, cmy = my / 2, base = pi.coff / 4;
      auto pv = [&](int a, int b) { return Uoo_all[fc[base + a + b * cmx]]; };
      const double p0 = pv(ci, cj);
      double px = 0.0, py = 0.0;
      if (cmx > 1)
        px = ci == 0 ? pv(1, cj) - pv(0, cj)
             : ci == cmx - 1 ? pv(cmx - 1, cj) - pv(cmx - 2, cj)
                             : (pv(ci + 1, cj) - pv(ci - 1, cj)) / 2;
      if (cmy > 1)
        py = cj == 0 ? pv(ci, 1) - pv(ci, 0)
             : cj == cmy - 1 ? pv(ci, cmy - 1) - pv(ci, cmy - 2)
                             : (pv(ci, cj + 1) - pv(ci, cj - 1)) / 2;
      u = conslin_child(p0, px, py, i & 1, j & 1);
    }
  }
  const int o = oq[cc];
  if (o >= 0) {
    q = Qold[o];
    if (hist) u = Uoo_old[oc[cc]];
  }
  L.Q[pi.qoff + (i + 1) + (j + 1) * (mx + 2)] = q;
  if (hist) Uoo_k[cc] = u;
}

// ---------------------------------------------------------------- launchers
static inline unsigned nblk(long long n) { return (unsigned)((n + kTile - 1) / kTile); }

template <typename T, int MODE>
static void launch_relax_m(const LevelDev<T>& L, const T* Pq, int colour, T omega, int homog, hipStream_t s) {
  if (L.even_mx && !(MODE & kKeepCf))
    hipLaunchKernelGGL((k_relax_c<T, MODE>), dim3(L.ntile), dim3(kTile / 2), 0, s, L, Pq, colour, omega,
                       homog);
  else
    hipLaunchKernelGGL((k_relax<T, MODE>), dim3(L.ntile), dim3(kTile), 0, s, L, Pq, colour, omega, homog);
}

template <typename T>
void launch_relax(const LevelDev<T>& L, const LevelDev<T>* parent, int colour, T omega, int homog, int mode,
                  hipStream_t s) {
  if (!L.ntile) return;
  const T* Pq = parent ? (const T*)parent->Q : (const T*)nullptr;
  switch (mode) {
#define BSAM_RELAX_MODE(m) \
  case m: launch_relax_m<T, m>(L, Pq, colour, omega, homog, s); break;
    BSAM_RELAX_MODE(0) BSAM_RELAX_MODE(kPush) BSAM_RELAX_MODE(kKeepCf) BSAM_RELAX_MODE(kKeepCf | kPush)
    BSAM_RELAX_MODE(kInlRecords) BSAM_RELAX_MODE(kInlRecords | kPush)
    BSAM_RELAX_MODE(kInlRecords | kKeepCf) BSAM_RELAX_MODE(kInlRecords | kKeepCf | kPush)
    BSAM_RELAX_MODE(kInlArith) BSAM_RELAX_MODE(kInlArith | kPush)
    BSAM_RELAX_MODE(kInlArith | kKeepCf) BSAM_RELAX_MODE(kInlArith | kKeepCf | kPush)
#undef BSAM_RELAX_MODE
    default: break;
  }
}

template <typename T>
void launch_relax_sweep(const LevelDev<T>& L, const LevelDev<T>* parent, const T* Qin, T* Qout, int mode, int keep,
                        T omega, int homog, hipStream_t s) {
  if (!L.nsweep) return;
  const T* Pq = parent ? (const T*)parent->Q : (const T*)nullptr;
  const dim3 g(L.nsweep), b(256);
#define BSAM_SWEEP(m, k) hipLaunchKernelGGL((k_relax_sweep<T, m, k>), g, b, 0, s, L, L.sweep, Qin, Qout, Pq, omega, homog)
  if (keep) {
    if (mode == 0) BSAM_SWEEP(0, true);
    else if (mode == kInlRecords) BSAM_SWEEP(kInlRecords, true);
    else BSAM_SWEEP(kInlArith, true);
  } else {
    if (mode == 0) BSAM_SWEEP(0, false);
    else if (mode == kInlRecords) BSAM_SWEEP(kInlRecords, false);
    else BSAM_SWEEP(kInlArith, false);
  }
#undef BSAM_SWEEP
}

template <typename T>
bool tiled_fits(const LevelDev<T>& L) {
  return L.nstile > 0 && (long long)L.stile_lds * (long long)sizeof(T) <= kTiledLdsBytes;
}

template <typename T>
void launch_relax_tiled(const LevelDev<T>& L, const LevelDev<T>* parent, int nsweeps, T omega, int homog,
                        hipStream_t s) {
  if (!L.nstile || nsweeps <= 0) return;
  hipLaunchKernelGGL(k_relax_tiled<T>, dim3(L.nstile), dim3(256), (size_t)L.stile_lds * sizeof(T), s, L,
                     parent ? (const T*)parent->Q : (const T*)nullptr, nsweeps, omega, homog);
}

template <typename T>
void launch_fill(const LevelDev<T>& L, const LevelDev<T>* parent, int homog, hipStream_t s) {
  const long long n = (long long)L.nghost + (parent ? L.nring_cf : 0);
  if (!n) return;
  hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(kTile), 0, s, L,
                     parent ? (const T*)parent->Q : (const T*)nullptr, homog);
}

template <typename T>
void launch_relax_region(const LevelDev<T>& L, const LevelDev<T>* parent, int nsweeps, int frame, T omega,
                         int homog, hipStream_t s, T* Qout, int ghosts, T* Rq, T* Rdup, int ring, int corr,
                         const LevelDev<T>* parent_dev, int fload) {
  if (!L.nregion || nsweeps <= 0) return;
  const int per = (L.nregion + kXcds - 1) / kXcds;
  const int rg = (ring && parent) ? 1 : 0;
  const int rring = rg ? (L.nring_cf + L.nregion - 1) / L.nregion : 0;
  const int cr = (corr && parent && L.rcorr) ? 1 : 0;
  const size_t lds = region_lds_bytes<T>(L.rslots, L.rghmax, std::min(rring, kRegionThreads), ghosts, rg,
                                         cr ? L.rwinmax : 0);
  // the kernel for the level's spacing and omega (compile-time forms of the
  // update, the same numbers as the general one)
  static bool raised = false;  // the canvas plus staged records can pass 64 KB (gfx950: 160 KB per CU)
  if (!raised) {
    const void* forms[8] = {(const void*)k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, true, true>,
                            (const void*)k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, true, false>,
                            (const void*)k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, false, true>,
                            (const void*)k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, false, false>,
                            (const void*)k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, true, true, true>,
                            (const void*)k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, true, false, true>,
                            (const void*)k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, false, true, true>,
                            (const void*)k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, false, false, true>};
    for (const void* f : forms)
      (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kStageLdsLimit);
    raised = true;
  }
  const int fl = (parent_dev && L.nregion == 1) ? fload : 0;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(per * kXcds), dim3(kRegionThreads), lds, s, L,
                       parent ? (const T*)parent->Q : (const T*)nullptr, Qout ? Qout : L.Qs, 2 * nsweeps, frame,
                       omega, homog, per, ghosts, Rq, Rdup, (ring && parent) ? 1 : 0, cr, parent_dev, fl);
  };
  // the fused load is a form of its own, so the other launches run the
  // kernel without its code
  const bool w1 = omega == T(1);
  if (fl) {
    if (L.pow2 && w1) go(k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, true, true, true>);
    else if (L.pow2) go(k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, true, false, true>);
    else if (w1) go(k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, false, true, true>);
    else go(k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, false, false, true>);
  } else if (L.pow2 && w1) go(k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, true, true>);
  else if (L.pow2) go(k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, true, false>);
  else if (w1) go(k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, false, true>);
  else go(k_relax_region<T, kRegionThreads, kRegionK, kRegionKL, false, false>);
}

template <typename T>
void launch_refresh_cf(const LevelDev<T>& L, const LevelDev<T>* parent, hipStream_t s) {
  if (!L.nghost || !parent || !L.aux) return;
  hipLaunchKernelGGL(k_refresh_cf<T>, dim3(nblk(L.nghost)), dim3(kTile), 0, s, L, (const T*)parent->Q);
}

template <typename T>
void launch_fill_copy(const LevelDev<T>& L, const LevelDev<T>* parent, int homog, const T* Qsrc, hipStream_t s) {
  const long long n = (long long)L.nghost + (parent ? L.nring_cf : 0);
  const int nfb = (int)nblk(n);
  if (nfb + L.ntile == 0) return;
  hipLaunchKernelGGL(k_fill_copy<T>, dim3(nfb + L.ntile), dim3(kTile), 0, s, L,
                     parent ? (const T*)parent->Q : (const T*)nullptr, homog, Qsrc, nfb);
}

template <typename T>
void launch_fill_copy_restrict(const LevelDev<T>& L, const LevelDev<T>& parent, int homog, const T* Qsrc, T* dup,
                               hipStream_t s) {
  const int nfb = (int)nblk(L.nghost);
  const int nb = nfb + L.ntile + (int)nblk(L.nfoot);
  if (!nb) return;
  hipLaunchKernelGGL(k_fill_copy_restrict<T>, dim3(nb), dim3(kTile), 0, s, L, homog, Qsrc, nfb, parent.Q, dup);
}

template <typename T>
void launch_ring(const LevelDev<T>& L, const LevelDev<T>& parent, hipStream_t s) {
  if (!L.nring_all) return;
  hipLaunchKernelGGL(k_ring<T>, dim3(nblk(L.nring_all)), dim3(kTile), 0, s, L, (const T*)parent.Q);
}

template <typename T>
void launch_restrict_fill(const LevelDev<T>& F, const LevelDev<T>& P, const LevelDev<T>* PP, int homog,
                          hipStream_t s, const LevelDev<T>* C, int cf) {
  const int nring = C ? C->nring_all : 0, ncf = (C && cf && C->aux) ? C->nghost : 0;
  const long long n = (long long)F.nfoot + P.nghost + (PP ? P.nring_cf : 0) + nring + ncf;
  if (!n) return;
  hipLaunchKernelGGL(k_restrict_fill<T>, dim3(nblk(n)), dim3(kTile), 0, s, F, P, PP ? (const T*)PP->Q : nullptr,
                     homog, C ? *C : F, nring, ncf);
}

template <typename T>
void launch_ring_cf(const LevelDev<T>& L, const LevelDev<T>& parent, hipStream_t s) {
  if (!L.nring_cf) return;
  hipLaunchKernelGGL(k_ring_cf<T>, dim3(nblk(L.nring_cf)), dim3(kTile), 0, s, L, (const T*)parent.Q);
}

template <typename T>
void launch_filldown_small(const LevelDev<T>* LV_dev, int kfrom, int kto, int homog, hipStream_t s, int kqc, int ktop,
                           int cftop) {
  kto = std::max(kto, 2);
  if (kfrom < kto) return;
  hipLaunchKernelGGL(k_filldown_small<T>, dim3(1), dim3(1024), 0, s, LV_dev, kfrom, kto, homog, kqc, ktop, cftop);
}

template <typename T>
void launch_restrict(const LevelDev<T>& L, const LevelDev<T>& parent, int qc_only, hipStream_t s, T* dup) {
  if (!L.nfoot) return;
  hipLaunchKernelGGL(k_restrict<T>, dim3(nblk(L.nfoot)), dim3(kTile), 0, s, L, parent.Q, qc_only, dup);
}

template <typename T>
void launch_coarse_load(const LevelDev<T>& L, const LevelDev<T>& parent, int mass_corr, hipStream_t s) {
  if (L.nfoot) hipLaunchKernelGGL(k_coarse_load<T>, dim3(nblk(L.nfoot)), dim3(kTile), 0, s, L, parent);
  if (mass_corr && parent.ntgt)
    hipLaunchKernelGGL(k_reflux<T>, dim3(nblk(parent.ntgt)), dim3(kTile), 0, s, L, parent);
}

template <typename T>
void launch_load(const LevelDev<T>& L, const LevelDev<T>& parent, int mass_corr, hipStream_t s, int pghost,
                 int homog, const T* PPq, int from_ft) {
  const long long n = (long long)L.nfoot + L.nring_all + (mass_corr ? parent.ntgt : 0);
  if (!n) return;
  const long long nr = (PPq && parent.aux) ? parent.nghost : 0;
  hipLaunchKernelGGL(k_load<T>, dim3(nblk(n + nr)), dim3(kTile), 0, s, L, parent, mass_corr, pghost, homog,
                     nr ? PPq : (const T*)nullptr, from_ft);
}

template <typename T>
void launch_correct(const LevelDev<T>& L, const LevelDev<T>& parent, int init, hipStream_t s, int prep, T* dup) {
  if (!L.ntile) return;
  hipLaunchKernelGGL(k_correct<T>, dim3(L.ntile), dim3(kTile), 0, s, L, (const T*)parent.Q, init,
                     prep && L.arith && L.aux ? 1 : 0, dup);
}

template <typename T>
void launch_coarsest(const CoarsestDev& C, const LevelDev<T>& L1, int homog, hipStream_t s) {
  const unsigned nb = (unsigned)((C.N + kCoarsestThreads - 1) / kCoarsestThreads);
  hipLaunchKernelGGL(k_coarsest<T>, dim3(nb), dim3(kCoarsestThreads), C.M * sizeof(double), s, C, L1, homog);
}

template <typename T>
void launch_coarse_cycle(const LevelDev<T>* LV_dev, const LevelDev<T>* LV_host, int kc, const CoarsestDev& C,
                         int npre, int npost, T omega, int homog, int mass_corr, hipStream_t s, int from_ft) {
  const long long rhs_b = stage_align((long long)C.M * (long long)sizeof(double));
  long long stage_b = 0;
  for (int k = 0; k < kc; ++k) stage_b += stage_level_bytes<T>(LV_host[k], k + 1 < kc ? LV_host[k + 1].ncorr : -1);
  // the dynamic budget: gfx950's 160 KB of LDS per workgroup less the
  // kernel's static arrays (the staged level descriptors among them), at
  // most kStageLdsLimit
  static long long cap = -1;
  if (cap < 0) {
    hipFuncAttributes fa{};
    long long st = 8192;
    if (hipFuncGetAttributes(&fa, (const void*)k_coarse_cycle<T>) == hipSuccess) st = (long long)fa.sharedSizeBytes;
    cap = std::min<long long>(kStageLdsLimit, 160LL * 1024 - st);
    (void)hipFuncSetAttribute((const void*)k_coarse_cycle<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)cap);
  }
  const int staged = (kc <= kStageMaxLevels && rhs_b + stage_b <= cap) ? 1 : 0;
  const size_t lds = staged ? (size_t)(rhs_b + stage_b) : (size_t)C.M * sizeof(double);
  hipLaunchKernelGGL(k_coarse_cycle<T>, dim3(1), dim3(1024), lds, s, LV_dev, kc, C, npre, npost, omega, homog,
                     mass_corr, staged, from_ft);
}

template <typename T>
void launch_qc_refresh_all(const LevelDev<T>* LV_dev, const int2* items, int nitems, hipStream_t s) {
  if (!nitems) return;
  hipLaunchKernelGGL(k_qc_refresh_all<T>, dim3(nblk(nitems)), dim3(kTile), 0, s, LV_dev, items, nitems);
}

template <typename T, typename TO>
void launch_residual_all(const LevelDev<T>* LV_dev, int nlev, const int2* tl, int ntl, int mass_corr,
                         ResidualOut* parts, TO* const* rout, int k0, hipStream_t s) {
  if (!ntl) return;
  hipLaunchKernelGGL((k_residual_all<T, TO>), dim3(std::min(ntl, kResidualParts)), dim3(kTile), 0, s, LV_dev, nlev,
                     tl, ntl, mass_corr, parts, rout, k0);
}

template <typename T>
void launch_reset_force(const LevelDev<T>* LV_dev, int nlev, int max_cells, hipStream_t s) {
  if (nlev <= 0 || max_cells <= 0) return;
  const unsigned gx = std::min(nblk(max_cells), 4096u);
  hipLaunchKernelGGL(k_reset_force<T>, dim3(gx, nlev), dim3(kTile), 0, s, LV_dev);
}

template <typename T>
void launch_reset_levels(const LevelDev<T>* LV_dev, const T* const* Q0_dev, int nlev, int max_elems, hipStream_t s) {
  if (nlev <= 0 || max_elems <= 0) return;
  const unsigned gx = std::min(nblk(max_elems), 4096u);
  hipLaunchKernelGGL(k_reset_levels<T>, dim3(gx, nlev), dim3(kTile), 0, s, LV_dev, Q0_dev);
}

void launch_reduce(const ResidualOut* parts, int nparts, ResidualOut* out, hipStream_t s, ResidualOut* out_host,
                   unsigned long long* seq) {
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(kReduceThreads), 0, s, parts, nparts, out, out_host, seq);
}

template <typename T>
void launch_mass(const LevelDev<T>& L, ResidualOut* parts, hipStream_t s) {
  if (!L.ntile) return;
  hipLaunchKernelGGL(k_mass<T>, dim3(L.ntile), dim3(kTile), 0, s, L, parts);
}

void launch_add_correction(const LevelDev<double>& L, const LevelDev<float>& E, hipStream_t s) {
  if (!L.ntile) return;
  hipLaunchKernelGGL(k_add_correction, dim3(L.ntile), dim3(kTile), 0, s, L, (const float*)E.Q);
}

template <typename T>
void launch_fill_value(T* a, long long n, T v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_value<T>, dim3(nblk(n)), dim3(kTile), 0, s, a, n, v);
}

template <typename TD, typename TS>
void launch_convert(TD* d, const TS* src, long long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_convert<TD, TS>), dim3(nblk(n)), dim3(kTile), 0, s, d, src, n);
}

void launch_bdf2_rhs(double* FT, const double* f, const double* Uo, const double* Uoo, double dt,
                     int first, long long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_bdf2_rhs, dim3(nblk(n)), dim3(kTile), 0, s, FT, f, Uo, Uoo, dt, first, n);
}

void launch_cn_rhs(double* FT, const double* f, const double* Uo, double c, long long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cn_rhs, dim3(nblk(n)), dim3(kTile), 0, s, FT, f, Uo, c, n);
}

void launch_history(const LevelDev<double>& L, double* Uo, double* Uoo, int rotate, hipStream_t s) {
  if (!L.ntile) return;
  hipLaunchKernelGGL(k_history, dim3(L.ntile), dim3(kTile), 0, s, L, Uo, Uoo, rotate);
}

// about 32 waves per CU in FP64 and 8 in FP32 (the fastest heights measured
// on a 4096 x 4096 level once neighbouring row blocks share their edge rows)
template <typename T>
int stream_rows(const PatchInfo& P, int ncu) {
  const int nstrip = (P.mx + kStreamCols - 1) / kStreamCols;
  const long long waves = sizeof(T) == 8 ? 32 : 8;
  const long long want = (long long)P.my * nstrip / (waves * std::max(ncu, 1));
  int rows = kStreamRowsMin;
  while (rows * 2 <= want && rows * 2 <= StreamRows<T>::value) rows *= 2;
  return rows;
}

template <typename T>
void launch_sweep_stream(const LevelDev<T>& L, const PatchInfo& P, const T* Qin, T* Qout, T omega, int homog,
                         int ncu, int rows_req, hipStream_t s) {
  const int rows = rows_req > 0 ? rows_req : stream_rows<T>(P, ncu);
  const int nstrip = (P.mx + kStreamCols - 1) / kStreamCols;
  const int nblk_rows = (P.my + rows - 1) / rows;
  const int per = (nstrip * nblk_rows + kXcds - 1) / kXcds;
  hipLaunchKernelGGL(k_sweep_stream<T>, dim3(per * kXcds), dim3(64), 0, s, L, P, Qin, Qout, omega, homog,
                     rows);
}

void launch_regrid_level(const LevelDev<double>& L, const double* Pq, int conslin, int hist, double* Uoo_k,
                         const double* Uoo_all, const int* fc, const int* oq, const int* oc, const double* Qold,
                         const double* Uoo_old, hipStream_t s) {
  if (!L.ntile) return;
  hipLaunchKernelGGL(k_regrid_level, dim3(L.ntile), dim3(kTile), 0, s, L, Pq, conslin, hist, Uoo_k, Uoo_all, fc, oq,
                     oc, Qold, Uoo_old);
}

int batch_cycle_blocks_per_cu(int max_m) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch_cycle, 1024, (size_t)max_m * sizeof(double)) !=
      hipSuccess)
    return 0;
  return nb;
}

int launch_batch_cycle(const BatchItem* items, const int* grp_item, int nwg, unsigned* ctr, int* err, int max_m,
                       int npre, int npost, double omega, int mass_corr, int mixed, int inner, hipStream_t s) {
  if (nwg <= 0) return 0;
  void* args[] = {(void*)&items, (void*)&grp_item, (void*)&ctr, (void*)&err, (void*)&npre, (void*)&npost,
                  (void*)&omega, (void*)&mass_corr, (void*)&mixed, (void*)&inner};
  return (int)hipLaunchCooperativeKernel((const void*)k_batch_cycle, dim3(nwg), dim3(1024), args,
                                         (unsigned)((size_t)max_m * sizeof(double)), s);
}

// ---------------------------------------------------------------- tagging
// Relative truncation error of coarse cell (ci, cj) of a patch
// (EllipticBSAM/src/ebRte.m:14-42): tau = L_2h(R u) - R(L_h u), where R u has
// a ghost ring averaged from the fine ghost pairs, and the 2h operator uses
// face-averaged D and the 4-cell mean of C.
__device__ double rte_tau(const LevelDev<double>& L, const PatchInfo& pi, int ci, int cj) {
  const int mx = pi.mx, my = pi.my, ldq = mx + 2, cmx = mx / 2, cmy = my / 2;
  const double* __restrict__ Q = L.Q + pi.qoff;
  const double* __restrict__ DeW = L.DeW + pi.xoff;  // (mx+1) x my
  const double* __restrict__ DnS = L.DnS + pi.yoff;  // mx x (my+1)
  const double* __restrict__ CC = L.CC + pi.coff;
  const H2<double> g = level_h2<double>(L);
  double lh[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int i = 2 * ci + (s & 1), j = 2 * cj + (s >> 1);
    lh[s] = op_at<double>(Q, (i + 1) + (j + 1) * ldq, ldq, DeW[i + j * (mx + 1)], DeW[i + 1 + j * (mx + 1)],
                          DnS[i + j * mx], DnS[i + (j + 1) * mx], CC[i + j * mx], g);
  }
  const double rlh = 0.25 * (((lh[0] + lh[1]) + lh[2]) + lh[3]);
  // R u at coarse (I, J), I in [-1, cmx], J in [-1, cmy] (not both outside)
  auto ru = [&](int I, int J) -> double {
    if (I < 0) return 0.5 * (Q[(2 * J + 1) * ldq] + Q[(2 * J + 2) * ldq]);
    if (I >= cmx) return 0.5 * (Q[(mx + 1) + (2 * J + 1) * ldq] + Q[(mx + 1) + (2 * J + 2) * ldq]);
    if (J < 0) return 0.5 * (Q[2 * I + 1] + Q[2 * I + 2]);
    if (J >= cmy) return 0.5 * (Q[(2 * I + 1) + (my + 1) * ldq] + Q[(2 * I + 2) + (my + 1) * ldq]);
    const int q = (2 * I + 1) + (2 * J + 1) * ldq;
    return 0.25 * (((Q[q] + Q[q + 1]) + Q[q + ldq]) + Q[q + 1 + ldq]);
  };
  const int i0 = 2 * ci, j0 = 2 * cj;
  const double dw = 0.5 * (DeW[i0 + j0 * (mx + 1)] + DeW[i0 + (j0 + 1) * (mx + 1)]);
  const double de = 0.5 * (DeW[i0 + 2 + j0 * (mx + 1)] + DeW[i0 + 2 + (j0 + 1) * (mx + 1)]);
  const double ds = 0.5 * (DnS[i0 + j0 * mx] + DnS[i0 + 1 + j0 * mx]);
  const double dn = 0.5 * (DnS[i0 + (j0 + 2) * mx] + DnS[i0 + 1 + (j0 + 2) * mx]);
  const double cc = 0.25 * (((CC[i0 + j0 * mx] + CC[i0 + 1 + j0 * mx]) + CC[i0 + (j0 + 1) * mx]) +
                            CC[i0 + 1 + (j0 + 1) * mx]);
  const double hx = 2.0 * L.h[0], hy = 2.0 * L.h[1];
  const H2<double> gc{hx * hx, hy * hy, 0.25 * L.ih2[0], 0.25 * L.ih2[1], L.pow2};
  const double lc = op5<double>(ru(ci, cj), ru(ci - 1, cj), ru(ci + 1, cj), ru(ci, cj - 1), ru(ci, cj + 1), dw, de, ds,
                                dn, cc, gc);
  return lc - rlh;
}

// One item per cell of the level (256-cell tiles): indicator > thr into the
// level-index raw tag map and the coverage map (EllipticBSAM/src/ebTagCluster.m:27-75).
// ulap: sqrt(sdx^2 + sdy^2) of the undivided second differences of Q
// (ebTagCluster.m:50-55); rte: |tau| of the cell's 2x2 block (ebRte.m:40-41).
__global__ void __launch_bounds__(kTile) k_tag_cells(LevelDev<double> L, const int2* __restrict__ lo, int indicator,
                                                     double thr, uint8_t* __restrict__ raw,
                                                     uint8_t* __restrict__ cover) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t >= L.ntile * kTile) return;
  const TileRec& tr = L.tiles[t / kTile];
  const PatchInfo& pi = tr.pi;
  const int c = tr.start + t % kTile;
  if (c >= pi.mx * pi.my) return;
  const int j = c / pi.mx, i = c - j * pi.mx;
  double s;
  if (indicator == 0) {
    const int ldq = pi.mx + 2;
    const double* __restrict__ Q = L.Q + pi.qoff;
    const int q = (i + 1) + (j + 1) * ldq;
    const double u = Q[q];
    const double sdx = (Q[q + 1] - 2.0 * u) + Q[q - 1];
    const double sdy = (Q[q + ldq] - 2.0 * u) + Q[q - ldq];
    s = sqrt(sdx * sdx + sdy * sdy);
  } else {
    s = fabs(rte_tau(L, pi, i >> 1, j >> 1));
  }
  const int2 b = lo[tr.p];
  const long long at = (long long)(b.x + i) + (long long)(b.y + j) * L.n[0];
  raw[at] = s > thr ? 1 : 0;
  cover[at] = 1;
}

// Buffer and nesting clip per level-index cell (ebTagCluster.m:78-90,
// dilate :113-127, erode8 :130-141): a cell is allowed when it and every
// in-domain 8-neighbour are covered; it is tagged when allowed and a raw tag
// lies within b cells in both directions.
__global__ void __launch_bounds__(kTile) k_tag_finish(const uint8_t* __restrict__ raw,
                                                      const uint8_t* __restrict__ cover, int n1, int n2, int b,
                                                      uint8_t* __restrict__ out) {
  const long long t = (long long)blockIdx.x * kTile + threadIdx.x;
  if (t >= (long long)n1 * n2) return;
  const int I = (int)(t % n1), J = (int)(t / n1);
  bool allowed = true;
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      const int x = I + dx, y = J + dy;
      if (x >= 0 && x < n1 && y >= 0 && y < n2) allowed = allowed && cover[x + (long long)y * n1];
    }
  bool tag = false;
  if (allowed) {
    const int x0 = max(I - b, 0), x1 = min(I + b, n1 - 1), y0 = max(J - b, 0), y1 = min(J + b, n2 - 1);
    for (int y = y0; y <= y1 && !tag; ++y)
      for (int x = x0; x <= x1; ++x)
        if (raw[x + (long long)y * n1]) {
          tag = true;
          break;
        }
  }
  out[t] = tag ? 1 : 0;
}

void launch_tag_cells(const LevelDev<double>& L, const int2* lo, int indicator, double thr, uint8_t* raw,
                      uint8_t* cover, hipStream_t s) {
  if (!L.ntile) return;
  hipLaunchKernelGGL(k_tag_cells, dim3(L.ntile), dim3(kTile), 0, s, L, lo, indicator, thr, raw, cover);
}

void launch_tag_finish(const uint8_t* raw, const uint8_t* cover, int n1, int n2, int buffer, uint8_t* out,
                       hipStream_t s) {
  const long long n = (long long)n1 * n2;
  if (!n) return;
  hipLaunchKernelGGL(k_tag_finish, dim3(nblk(n)), dim3(kTile), 0, s, raw, cover, n1, n2, buffer < 0 ? 0 : buffer, out);
}

// ---------------------------------------------------------------- diffuse-domain geometry
// Signed distance (positive inside) from each point to a closed, radially
// ordered polygon, and the diffuse-domain fields of the transmission problem
// (DiffuseDomainBSAM/src/ddStarfishGeometry.m:35-72, ddTransmissionProblem.m:14-15).
// The squared distance is the minimum of the point-to-segment distances over
// every segment (the reference prunes with a box tree; the minimum is the
// same), each evaluated in the reference's expression order; segments are
// staged through LDS in chunks shared by the workgroup.
constexpr int kSegChunk = 256;

__global__ void __launch_bounds__(kTile) k_polygon_fields(const double* __restrict__ ax, const double* __restrict__ ay,
                                                          const double* __restrict__ ex, const double* __restrict__ ey,
                                                          const double* __restrict__ len2, int ns,
                                                          const double* __restrict__ x, const double* __restrict__ y,
                                                          long long npts, double eps, double* __restrict__ d,
                                                          double* __restrict__ phi, double* __restrict__ delta) {
  __shared__ double sa[kSegChunk], sb[kSegChunk], se[kSegChunk], sf[kSegChunk], sl[kSegChunk];
  const long long i = (long long)blockIdx.x * kTile + threadIdx.x;
  const bool live = i < npts;
  const double xi = live ? x[i] : 0.0, yi = live ? y[i] : 0.0;
  const double twopi = 2.0 * M_PI;
  double ang = atan2(yi, xi);
  ang = ang - floor(ang / twopi) * twopi;  // mod(atan2(y, x), 2 pi)
  int seg = (int)floor(ang * ns / twopi);
  seg = min(max(seg, 0), ns - 1);
  const double dx0 = xi - ax[seg], dy0 = yi - ay[seg];
  double best = dx0 * dx0 + dy0 * dy0;  // a boundary vertex bounds the distance
  for (int c0 = 0; c0 < ns; c0 += kSegChunk) {
    const int nc = min(kSegChunk, ns - c0);
    __syncthreads();
    for (int j = threadIdx.x; j < nc; j += kTile) {
      sa[j] = ax[c0 + j];
      sb[j] = ay[c0 + j];
      se[j] = ex[c0 + j];
      sf[j] = ey[c0 + j];
      sl[j] = len2[c0 + j];
    }
    __syncthreads();
    for (int j = 0; j < nc; ++j) {
      double t = ((xi - sa[j]) * se[j] + (yi - sb[j]) * sf[j]) / sl[j];
      t = fmax(0.0, fmin(1.0, t));
      const double px = xi - sa[j] - t * se[j];
      const double py = yi - sb[j] - t * sf[j];
      const double dd = px * px + py * py;
      best = fmin(best, dd);
    }
  }
  if (!live) return;
  // the ray at the point's angle meets its enclosing segment at `radius`
  const double ux = cos(ang), uy = sin(ang);
  const double radius = (ax[seg] * ey[seg] - ay[seg] * ex[seg]) / (ux * ey[seg] - uy * ex[seg]);
  const double dist = (hypot(xi, yi) > radius ? -1.0 : 1.0) * sqrt(best);
  d[i] = dist;
  if (phi) phi[i] = (1.0 + tanh(dist / eps)) / 2.0;
  if (delta) {
    const double sh = 1.0 / cosh(dist / eps);
    delta[i] = sh * sh / (2.0 * eps);
  }
}

void launch_polygon_fields(const double* seg5, int ns, const double* x, const double* y, long long npts, double eps,
                           double* d, double* phi, double* delta, hipStream_t s) {
  if (npts <= 0) return;
  hipLaunchKernelGGL(k_polygon_fields, dim3(nblk(npts)), dim3(kTile), 0, s, seg5, seg5 + ns, seg5 + 2 * ns,
                     seg5 + 3 * ns, seg5 + 4 * ns, ns, x, y, npts, eps, d, phi, delta);
}

#define BSAM_INST(T)                                                                                    \
  template void launch_relax<T>(const LevelDev<T>&, const LevelDev<T>*, int, T, int, int, hipStream_t); \
  template bool tiled_fits<T>(const LevelDev<T>&);                                                      \
  template void launch_relax_tiled<T>(const LevelDev<T>&, const LevelDev<T>*, int, T, int, hipStream_t); \
  template void launch_fill<T>(const LevelDev<T>&, const LevelDev<T>*, int, hipStream_t);               \
  template void launch_relax_region<T>(const LevelDev<T>&, const LevelDev<T>*, int, int, T, int, hipStream_t, T*, int, \
                                       T*, T*, int, int, const LevelDev<T>*, int);                          \
  template void launch_fill_copy<T>(const LevelDev<T>&, const LevelDev<T>*, int, const T*, hipStream_t); \
  template void launch_fill_copy_restrict<T>(const LevelDev<T>&, const LevelDev<T>&, int, const T*, T*,        \
                                             hipStream_t);                                                  \
  template void launch_refresh_cf<T>(const LevelDev<T>&, const LevelDev<T>*, hipStream_t);              \
  template void launch_ring<T>(const LevelDev<T>&, const LevelDev<T>&, hipStream_t);                    \
  template void launch_ring_cf<T>(const LevelDev<T>&, const LevelDev<T>&, hipStream_t);                 \
  template void launch_restrict_fill<T>(const LevelDev<T>&, const LevelDev<T>&, const LevelDev<T>*, int,   \
                                        hipStream_t, const LevelDev<T>*, int);                            \
  template void launch_restrict<T>(const LevelDev<T>&, const LevelDev<T>&, int, hipStream_t, T*);       \
  template void launch_coarse_load<T>(const LevelDev<T>&, const LevelDev<T>&, int, hipStream_t);        \
  template void launch_load<T>(const LevelDev<T>&, const LevelDev<T>&, int, hipStream_t, int, int, const T*, int); \
  template void launch_correct<T>(const LevelDev<T>&, const LevelDev<T>&, int, hipStream_t, int, T*);        \
  template void launch_filldown_small<T>(const LevelDev<T>*, int, int, int, hipStream_t, int, int, int);   \
  template void launch_reset_force<T>(const LevelDev<T>*, int, int, hipStream_t);                         \
  template void launch_reset_levels<T>(const LevelDev<T>*, const T* const*, int, int, hipStream_t);       \
  template void launch_sweep_stream<T>(const LevelDev<T>&, const PatchInfo&, const T*, T*, T, int, int, int,     \
                                       hipStream_t);                                                         \
  template void launch_relax_sweep<T>(const LevelDev<T>&, const LevelDev<T>*, const T*, T*, int, int, T, int,   \
                                      hipStream_t);                                                          \
  template void launch_coarsest<T>(const CoarsestDev&, const LevelDev<T>&, int, hipStream_t);           \
  template void launch_coarse_cycle<T>(const LevelDev<T>*, const LevelDev<T>*, int, const CoarsestDev&, int, int, T, \
                                       int, int, \
                                       hipStream_t, int);                                                    \
  template void launch_qc_refresh_all<T>(const LevelDev<T>*, const int2*, int, hipStream_t);            \
  template void launch_mass<T>(const LevelDev<T>&, ResidualOut*, hipStream_t);                          \
  template void launch_fill_value<T>(T*, long long, T, hipStream_t);
BSAM_INST(double)
BSAM_INST(float)
template void launch_residual_all<double, float>(const LevelDev<double>*, int, const int2*, int, int, ResidualOut*,
                                                 float* const*, int, hipStream_t);
template void launch_residual_all<double, double>(const LevelDev<double>*, int, const int2*, int, int, ResidualOut*,
                                                 double* const*, int, hipStream_t);
template void launch_residual_all<float, float>(const LevelDev<float>*, int, const int2*, int, int, ResidualOut*,
                                                float* const*, int, hipStream_t);
template void launch_fill_value<uint8_t>(uint8_t*, long long, uint8_t, hipStream_t);
template void launch_convert<float, double>(float*, const double*, long long, hipStream_t);
template void launch_convert<double, float>(double*, const float*, long long, hipStream_t);

}  // namespace bsam
