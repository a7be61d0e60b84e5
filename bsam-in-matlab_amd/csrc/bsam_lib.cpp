// Host side of libbsamgpu: hierarchy upload, descriptor tables, the solve
// loop and the C ABI declared in include/bsamgpu.h.
//
// The descriptor tables are derived from the patch boxes the way the reference
// derives its gather/scatter descriptors (EllipticBSAM/src/ebBuildOps.m:17-199,
// ebGatherParentCoeffs.m:12-63), but flattened into per-level arrays of
// absolute element indices so that every kernel sweeps all patches of a level
// in one launch.  The solve loop replays the reference's call sequence
// (ebSolveCycles.m / ebVcycle.m / ebRelaxLevel.m / ebTransfer.m) as a stream
// of level-wide launches, optionally captured once into a HIP graph and
// replayed per cycle.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/bsamgpu.h"
#include "bsam_kernels.h"

using namespace bsam;

// explicit inverse of the level-1 matrix (bsam_lu.cpp, host code)
bool bsam_lu_inverse(double* A, int M, double* Ainv);
// tag clustering (bsam_cluster.cpp, host code): returns the box count and
// writes at most cap boxes
int bsam_cluster_impl(const unsigned char* tags, const unsigned char* allowed, int n1, int n2, int min_width,
                      double fill_ratio, int* boxes, int cap, int* ndropped);

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(BSAM_E_HIP, "bsamgpu:hip: %s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// two HIP events released on every return path
struct EventPair {
  hipEvent_t e[2] = {nullptr, nullptr};
  hipError_t create() {
    for (hipEvent_t& x : e) {
      const hipError_t r = hipEventCreate(&x);
      if (r != hipSuccess) return r;
    }
    return hipSuccess;
  }
  ~EventPair() {
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

struct Box {
  int ilo, ihi, jlo, jhi;
};

// host copy of one level: inputs, offsets and tables
struct HLevel {
  int n[2] = {0, 0};
  double h[2] = {0, 0};
  int np = 0;
  std::vector<Box> box;
  std::vector<PatchInfo> pi;
  int ncell = 0, nq = 0, nx = 0, ny = 0, nqc = 0, ng = 0;
  std::vector<double> Q, FT, DeW, DnS, CC, ID, G;
  std::vector<TileRec> tiles;
  std::vector<STile> stiles;
  int stile_lds = 0;
  std::vector<GhostRec> ghost;
  std::vector<RingRec> ring_all, ring_cf;
  std::vector<FootRec> foot;
  std::vector<CorrRec> corr;
  std::vector<int> corr_target;   // parent cell of each corr entry (export)
  std::vector<int> tgt_cell, tgt_ptr, tgt_ent, tgt_of_cell, wmap, wmap_patch;
  std::vector<uint8_t> vis;
  std::vector<int8_t> klass;      // ghost classes per patch (export)
  std::vector<int> klass_off;
  // parent physical ghosts the FAS load reads: per QC window cell and per
  // ring item, the parent ghost record at the source (-1: an interior cell);
  // pgh_ok when every such record is physical (the load can evaluate them
  // from the parent's interior instead of reading a filled frame)
  std::vector<int> wgr, ring_rec;
  int pgh_ok = 0;
  // the region tiles' ghost records as uploaded: their boundary values are
  // copies of the ghost records' (refreshed with them by bsam_update_coeffs)
  std::vector<RegionGhost> rghost;
};

template <typename T>
struct DevBufs {
  std::vector<void*> allocs;
  ~DevBufs() {
    for (void* p : allocs) (void)hipFree(p);
  }
  template <typename U>
  int alloc(U** p, size_t n) {
    *p = nullptr;
    if (n == 0) return 0;
    if (hipMalloc((void**)p, n * sizeof(U)) != hipSuccess)
      return fail(BSAM_E_NOMEM, "bsamgpu:nomem: device allocation of %zu bytes failed", n * sizeof(U));
    allocs.push_back(*p);
    return 0;
  }
};

}  // namespace

struct bsam_ctx {
  int device = 0;
};

struct bsam_hier {
  bsam_ctx* ctx = nullptr;
  int nlev = 0, root = 0;
  int ncu = 0;  // compute units of the device (sizes the fused-sweep choice)
  double dom[4] = {0, 0, 0, 0};
  bsam_solve_opts opts{};
  std::vector<HLevel> H;                   // index 0 = level 1
  std::vector<double> ainv_host;           // host-only contexts: the coarsest inverse (column-major)
  std::vector<LevelDev<double>> d64;
  std::vector<LevelDev<float>> d32;
  CoarsestDev cst{};
  bool singular = false;
  bool have32 = false;
  hipStream_t stream = nullptr;
  DevBufs<char> mem;
  ResidualOut* parts = nullptr;            // residual partials, all levels
  std::vector<int> part_off;               // per level
  int nparts = 0;
  LevelDev<double>* lv64 = nullptr;        // device copies of the level descriptors
  LevelDev<float>* lv32 = nullptr;
  int2* res_tiles = nullptr;               // (level, tile) over all levels
  int nres_tiles = 0;
  int2* qc_items = nullptr;                // (level, item): ring + QC interior, k >= max(2, root+1)
  int nqc_items = 0;
  float** rout32 = nullptr;                // per level FP32 FTrue (defect right side)
  float** rout32f = nullptr;               // per level FP32 F (the same, one inner cycle)
  double** rout64 = nullptr;               // per level FP64 FTrue (cn right side)
  float* q32all = nullptr;                 // FP32 Q of all levels (contiguous)
  float* qc32all = nullptr;                // FP32 QC of all levels (contiguous)
  long long nq32all = 0, nqc32all = 0;
  ResidualOut* dres = nullptr;
  ResidualOut* hres = nullptr;             // pinned; the reduction writes it
  // the reductions' counter on the device and the count the host expects
  // in hres->pad2[0] after the last one it enqueued (wait_residual)
  unsigned long long* dseq = nullptr;
  long long seq_expect = 0;
  std::vector<double*> Q0;                 // uploaded initial guess per level
  double** q0dev = nullptr;                // the same pointers, on the device (bsam_hier_reset)
  double* Uo = nullptr;                    // parabolic history (packed cells, all levels)
  double* Uoo = nullptr;
  bool have_uoo = false;                   // Uoo holds u^{n-1} (after one step or an upload)
  double* src = nullptr;                   // resident source f(t1), packed cells
  bool have_src = false;
  double par_shift = std::numeric_limits<double>::quiet_NaN();  // H.par.aDt (cn right side)
  std::vector<long long> cell_base;        // per level offset in the history arrays
  long long ncell_all = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<double> graph_key;          // the options the captured cycle was built with
  ~bsam_hier() {
    // nothing the stream still runs may outlive the graph, pinned buffer or memory it reads
    if (stream) (void)hipStreamSynchronize(stream);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (hres) (void)hipHostFree(hres);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// ------------------------------------------------------------------ tables
inline int qidx(const PatchInfo& p, int li, int lj) { return p.qoff + li + lj * (p.mx + 2); }

// multiplier m with umulhi(c, m) == c / mx for every cell index c of an
// mx-by-my patch (ceil(2^32/mx), verified), or 0 when none is exact
unsigned row_magic(int mx, int my) {
  const unsigned long long m = ((1ULL << 32) + (unsigned long long)mx - 1) / (unsigned long long)mx;
  if (m > 0xffffffffULL) return 0;
  for (long long c = 0; c < (long long)mx * my; ++c)
    if ((long long)(((unsigned long long)c * m) >> 32) != c / mx) return 0;
  return (unsigned)m;
}

// h is a power of two: then so is h^2 and q / h^2 == q * (1 / h^2) exactly
bool is_pow2(double h) {
  int e = 0;
  return h > 0 && std::frexp(h, &e) == 0.5 && e > -60 && e < 60;
}
inline int cidx(const PatchInfo& p, int i, int j) { return p.coff + i + j * p.mx; }

void make_tiles(HLevel& L) {
  L.tiles.clear();
  L.stiles.clear();
  L.stile_lds = 0;
  for (int p = 0; p < L.np; ++p) {
    const int nc = L.pi[p].mx * L.pi[p].my;
    for (int s = 0; s < nc; s += kTile) {
      TileRec tr{};
      tr.p = p;
      tr.start = s;  // tr.pi is filled once the tables are built (finalize_tiles)
      L.tiles.push_back(tr);
    }
    // smoother tile = the whole patch; its LDS region is the padded patch
    L.stiles.push_back(STile{p, 0, L.pi[p].mx - 1, 0, L.pi[p].my - 1});
    L.stile_lds = std::max(L.stile_lds, (L.pi[p].mx + 2) * (L.pi[p].my + 2));
  }
}

// Physical-face and corner records shared by every level; internal-face
// records need the owner map and parent window (build_level_tables).
GhostRec phys_rec(const HLevel& L, const PatchInfo& P, int side, int t) {
  GhostRec r{};
  const int mx = P.mx, my = P.my;
  int dst, s0, s1;
  double h;
  switch (side) {
    case 0: dst = qidx(P, 0, t + 1); s0 = qidx(P, 1, t + 1); s1 = qidx(P, 2, t + 1); h = L.h[0]; break;
    case 1: dst = qidx(P, mx + 1, t + 1); s0 = qidx(P, mx, t + 1); s1 = qidx(P, mx - 1, t + 1); h = L.h[0]; break;
    case 2: dst = qidx(P, t + 1, 0); s0 = qidx(P, t + 1, 1); s1 = qidx(P, t + 1, 2); h = L.h[1]; break;
    default: dst = qidx(P, t + 1, my + 1); s0 = qidx(P, t + 1, my); s1 = qidx(P, t + 1, my - 1); h = L.h[1]; break;
  }
  const int goff = side == 0 ? 0 : side == 1 ? my : side == 2 ? 2 * my : 2 * my + mx;
  r.gi = P.goff + goff + t;
  const double g = L.G[r.gi];
  r.dst = dst;
  if (P.bc[side] == 1) {
    r.kind = G_DIR; r.s[0] = s0; r.s[1] = s1; r.c = g;
  } else {
    r.kind = G_NEU; r.s[0] = s0; r.c = h * g;
  }
  return r;
}

// interpolation record of a coarse-fine ghost (ebInterpCf.m:45-153).
// win(wi,wj) returns the parent Q index of QC window cell (wi,wj).
template <typename Win>
GhostRec cf_rec(const PatchInfo& P, int side, int t, bool qif, Win win) {
  GhostRec r{};
  const int mx = P.mx, my = P.my, cmx = mx / 2, cmy = my / 2;
  // side-local frame: position u along the face (1..len), depth 1/2 cells
  const int len = (side < 2) ? my : mx;
  const int u = t + 1;
  auto q = [&](int depth, int along) {  // padded Q index `depth` cells inside
    switch (side) {
      case 0: return qidx(P, depth, along);
      case 1: return qidx(P, mx + 1 - depth, along);
      case 2: return qidx(P, along, depth);
      default: return qidx(P, along, my + 1 - depth);
    }
  };
  auto w = [&](int along) {  // ring window cell on this side
    switch (side) {
      case 0: return win(0, along);
      case 1: return win(cmx + 1, along);
      case 2: return win(along, 0);
      default: return win(along, cmy + 1);
    }
  };
  const int clen = (side < 2) ? cmy : cmx;
  switch (side) {
    case 0: r.dst = qidx(P, 0, u); break;
    case 1: r.dst = qidx(P, mx + 1, u); break;
    case 2: r.dst = qidx(P, u, 0); break;
    default: r.dst = qidx(P, u, my + 1); break;
  }
  if (qif) {
    if (u == len) {
      r.kind = G_QIF_END; r.s[0] = q(2, len); r.s[1] = q(1, len);
      r.s[2] = w(clen + 1); r.s[3] = w(clen); r.s[4] = w(clen - 1);
    } else if (u == 1) {
      r.kind = G_QIF_END; r.s[0] = q(2, 1); r.s[1] = q(1, 1);
      r.s[2] = w(0); r.s[3] = w(1); r.s[4] = w(2);
    } else if ((u & 1) == 0) {
      r.kind = G_QIF; r.s[0] = q(2, u + 2); r.s[1] = q(1, u + 1); r.s[2] = w(u / 2);
    } else {
      r.kind = G_QIF; r.s[0] = q(2, u - 2); r.s[1] = q(1, u - 1); r.s[2] = w((u + 1) / 2);
    }
  } else {
    if ((u & 1) == 0) {
      r.kind = G_LIF; r.s[0] = q(1, u); r.s[1] = q(1, u - 1); r.s[2] = w(u / 2);
    } else {
      r.kind = G_LIF; r.s[0] = q(1, u); r.s[1] = q(1, u + 1); r.s[2] = w((u + 1) / 2);
    }
  }
  return r;
}

// Whether every face record of a level has the layout that the smoother's
// neighbour-assembled passes assume: coarse-fine records name the cells of
// cf_rec's stencils, and each sibling record's push target is a sibling ghost
// that copies this patch's boundary cell.  Levels that fail keep evaluating
// ghosts from the records in every pass.
bool face_records_arith(const HLevel& L) {
  std::unordered_map<int, int> by_dst;
  for (int t = 0; t < (int)L.ghost.size(); ++t) by_dst[L.ghost[t].dst] = t;
  std::vector<int> pushed(L.ghost.size(), 0), copy_face(L.ghost.size(), 0);
  for (int p = 0; p < L.np; ++p) {
    const PatchInfo& P = L.pi[p];
    const int mx = P.mx, my = P.my;
    if (mx < 2 || my < 2) return false;
    int off = P.gbase;
    for (int side = 0; side < 4; ++side) {
      const int len = side < 2 ? my : mx;
      auto q = [&](int depth, int along) {
        switch (side) {
          case 0: return qidx(P, depth, along);
          case 1: return qidx(P, mx + 1 - depth, along);
          case 2: return qidx(P, along, depth);
          default: return qidx(P, along, my + 1 - depth);
        }
      };
      for (int t = 0; t < len; ++t) {
        if (off + t >= (int)L.ghost.size()) return false;
        const GhostRec& r = L.ghost[off + t];
        const int u = t + 1;
        const bool even = (u & 1) == 0;
        if (r.dst != q(0, u)) return false;
        if (P.bc[side]) {
          if (r.kind != G_DIR && r.kind != G_NEU) return false;
          continue;
        }
        switch (r.kind) {
          case G_COPY: {
            const auto it = by_dst.find(r.gi);
            if (it == by_dst.end()) return false;
            const GhostRec& m = L.ghost[it->second];
            if (m.kind != G_COPY || m.s[0] != q(1, u)) return false;
            ++pushed[it->second];
            copy_face[off + t] = 1;
            break;
          }
          case G_QIF:
            if (u == 1 || u == len) return false;
            if (even ? (r.s[0] != q(2, u + 2) || r.s[1] != q(1, u + 1))
                     : (r.s[0] != q(2, u - 2) || r.s[1] != q(1, u - 1)))
              return false;
            break;
          case G_QIF_END:
            if ((u != 1 && u != len) || r.s[0] != q(2, u) || r.s[1] != q(1, u)) return false;
            break;
          case G_LIF:
            if (r.s[0] != q(1, u) || r.s[1] != (even ? q(1, u - 1) : q(1, u + 1))) return false;
            break;
          default:
            return false;
        }
      }
      off += len;
    }
  }
  // every sibling face ghost has exactly one cell that pushes to it
  for (size_t t = 0; t < L.ghost.size(); ++t)
    if (pushed[t] != copy_face[t]) return false;
  return true;
}

struct OwnerMap;
bool build_sweep_tiles(const HLevel& L, std::vector<SweepTile>& out);
constexpr int kSweepWGsPerCU = 3;  // resident fused-sweep workgroups per CU
constexpr long long kStreamMinCells = 1LL << 20;

struct OwnerMap {
  int n0 = 0, n1 = 0;
  std::vector<int> o;  // p+1, 0 = none
  void build(const HLevel& L) {
    n0 = L.n[0]; n1 = L.n[1];
    o.assign((size_t)n0 * n1, 0);
    for (int p = 0; p < L.np; ++p) {
      const Box& b = L.box[p];
      for (int j = b.jlo; j <= b.jhi; ++j)
        for (int i = b.ilo; i <= b.ihi; ++i) o[(size_t)(i - 1) + (size_t)(j - 1) * n0] = p + 1;
    }
  }
  int at(int gi, int gj) const {  // 1-based, 0 outside
    if (gi < 1 || gi > n0 || gj < 1 || gj > n1) return 0;
    return o[(size_t)(gi - 1) + (size_t)(gj - 1) * n0];
  }
};

// Fused-sweep tiles of a level: each patch split into tiles of at most
// kSweepTW x kSweepTH cells, each with the halo its colour-2 cells read in the
// second half of a sweep (the reads of k_relax_sweep's black phase, mirrored
// here).  Returns false, and the level keeps colour passes, when a needed cell
// is not of colour 1, a halo or its patch table overflows, or a position
// falls outside the tile's region.
bool build_sweep_tiles(const HLevel& L, std::vector<SweepTile>& out) {
  out.clear();
  OwnerMap own;
  own.build(L);
  auto colour1 = [&](int p, int i, int j) { return ((i + j + L.pi[p].par0) & 1) == 0; };
  for (int p = 0; p < L.np; ++p) {
    const PatchInfo& P = L.pi[p];
    const Box& b = L.box[p];
    const int mx = P.mx, my = P.my;
    if (mx % 2 || mx < 2 || my < 2) return false;
    const int ntx = (mx + kSweepTW - 1) / kSweepTW, nty = (my + kSweepTH - 1) / kSweepTH;
    for (int ty = 0; ty < nty; ++ty)
      for (int tx = 0; tx < ntx; ++tx) {
        // column boundaries on even cells: every tile row then holds tw/2
        // cells of each colour (the kernel's cell mapping)
        const int hm = mx / 2;
        const int i0 = 2 * (int)((long long)hm * tx / ntx), i1 = 2 * (int)((long long)hm * (tx + 1) / ntx);
        const int j0 = (int)((long long)my * ty / nty), j1 = (int)((long long)my * (ty + 1) / nty);
        const int tw = i1 - i0, th = j1 - j0;
        if (tw < 2 || tw > kSweepTW || tw % 2 || th < 1 || th > kSweepTH) return false;
        SweepTile st;
        std::memset(&st, 0, sizeof st);
        st.p = p; st.i0 = i0; st.j0 = j0; st.tw = tw; st.th = th;
        st.pi = P;
        const int rw = tw + 4;
        std::map<int, int> seen;  // region slot -> halo index
        std::vector<int> sibs;
        bool ok = true;
        // a colour-1 cell at P-local position (ni, nj) (possibly in a sibling)
        auto need = [&](int ni, int nj) {
          const int rx = ni - i0 + 2, ry = nj - j0 + 2;
          if (rx < 0 || rx >= rw || ry < 0 || ry >= th + 4) { ok = false; return; }
          const int slot = rx + ry * rw;
          if (ni >= 0 && ni < mx && nj >= 0 && nj < my) {
            if (!colour1(p, ni, nj)) { ok = false; return; }
            if (ni >= i0 && ni < i1 && nj >= j0 && nj < j1) return;  // a tile red cell
            if (seen.count(slot)) return;
            seen[slot] = 1;
            if (st.nhalo >= kSweepMaxHalo) { ok = false; return; }
            st.halo[st.nhalo++] = SweepHalo{slot, ni | (nj << 16)};
            return;
          }
          const int gi = b.ilo + ni, gj = b.jlo + nj;
          const int q = own.at(gi, gj) - 1;
          if (q < 0 || q == p) { ok = false; return; }
          const Box& qb = L.box[q];
          const int si = gi - qb.ilo, sj = gj - qb.jlo;
          if (!colour1(q, si, sj)) { ok = false; return; }
          if (seen.count(slot)) return;
          seen[slot] = 1;
          int ps = -1;
          for (size_t k = 0; k < sibs.size(); ++k)
            if (sibs[k] == q) ps = (int)k;
          if (ps < 0) {
            if ((int)sibs.size() >= kSweepMaxSib) { ok = false; return; }
            sibs.push_back(q);
            ps = (int)sibs.size() - 1;
          }
          if (st.nhalo >= kSweepMaxHalo) { ok = false; return; }
          st.halo[st.nhalo++] = SweepHalo{slot | ((ps + 1) << 16), si | (sj << 16)};
        };
        for (int j = j0; j < j1 && ok; ++j)
          for (int i = i0; i < i1 && ok; ++i) {
            if (colour1(p, i, j)) continue;
            const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
            for (int d = 0; d < 4 && ok; ++d) {
              const int ni = i + di[d], nj = j + dj[d];
              if (ni >= 0 && ni < mx && nj >= 0 && nj < my) { need(ni, nj); continue; }
              // the cell lies on face d: its ghost (relax_cell's forms)
              const int ii = i - di[d], ij = j - dj[d];  // inward neighbour
              if (P.bc[d]) { need(ii, ij); continue; }
              const int t = d < 2 ? j : i, len = d < 2 ? my : mx, u = t + 1;
              const int off = d == 0 ? 0 : d == 1 ? my : d == 2 ? 2 * my : 2 * my + mx;
              const GhostRec& r = L.ghost[P.gbase + off + t];
              const int ai = d < 2 ? 0 : 1, aj = d < 2 ? 1 : 0;  // along-face step
              switch (r.kind) {
                case G_COPY: {
                  need(ni, nj);
                  // the record copies the cell just needed
                  const int q = own.at(b.ilo + ni, b.jlo + nj) - 1;
                  if (q < 0 || r.s[0] != qidx(L.pi[q], b.ilo + ni - L.box[q].ilo + 1, b.jlo + nj - L.box[q].jlo + 1))
                    ok = false;
                  break;
                }
                case G_QIF:
                  if (u == 1 || u == len) { ok = false; break; }
                  if ((u & 1) == 0) { need(i + ai, j + aj); need(ii + 2 * ai, ij + 2 * aj); }
                  else { need(i - ai, j - aj); need(ii - 2 * ai, ij - 2 * aj); }
                  break;
                case G_QIF_END:
                  need(ii, ij);
                  break;
                case G_LIF:
                  if ((u & 1) == 0) need(i - ai, j - aj); else need(i + ai, j + aj);
                  break;
                default:
                  ok = false;
              }
            }
          }
        if (!ok) return false;
        st.nsib = (int)sibs.size();
        for (size_t k = 0; k < sibs.size(); ++k) st.sib[k] = L.pi[sibs[k]];
        out.push_back(st);
      }
  }
  return true;
}

// Ghost records of a level without a parent (level 1): all faces physical,
// corners extrapolated (ebFillGhosts.m:53-89).
// ------------------------------------------------------------ region tiles
// Tables of the whole-relax-call kernel (k_relax_region).  For every cell of
// the level, the cells its colour-pass update reads (ebKernels.m:37-95 with
// the ghosts of ebFillGhosts / ebInterpCf): the same-patch neighbours, the
// sibling cell a copy face ghost copies, the inward neighbour of a Dirichlet
// side, and the interpolation sources of a coarse-fine face (the record's
// cells, which face_records_arith has checked to follow the fixed stencil),
// plus the cell itself.  A tile owns a set of cells (pieces of one or more
// patches); walking the passes of a call backwards from its own cells gives
// each needed cell its depth (region_code), and the canvas is the bounding
// box of those cells with one slot of margin, so that every read of a pass
// stays in it.
struct RegionBuild {
  std::vector<RegionTile> tiles;
  std::vector<RegionPatch> patches;  // kRegionMaxPatches per tile
  std::vector<int> pid;              // their patch indices (-1: unused slot)
  std::vector<unsigned> cells;       // kRegionStride per tile
  std::vector<RegionGhost> ghosts;   // per tile, RegionTile::gofs / ngh
  std::vector<RegionRestrict> rs;    // per tile, RegionTile::rofs / nrs
  std::vector<RegionWin> win;        // per tile, RegionTile::wofs / nwin (build_region_windows)
  std::vector<RegionGhostW> gw;      // parallel to ghosts
  int max_slots = 0, max_gh = 0, max_rs = 0, max_win = 0;
};

// The form of the face ghosts of cell (p, i, j) as a sort key: 0 off every
// face; else 1 + (x-face form, east) * 16 + (y-face form, north), a form
// being the side's boundary type or its record kind
static unsigned char face_form(const HLevel& L, int p, int i, int j) {
  const PatchInfo& P = L.pi[p];
  const int mx = P.mx, my = P.my;
  auto form = [&](int side, int t) -> int {  // 1 DIR, 2 NEU, 3 COPY, 4 QIF, 5 QIF_END, 6 LIF(_CORNER)
    if (P.bc[side]) return P.bc[side] == 1 ? 1 : 2;
    const int off = side == 0 ? 0 : side == 1 ? my : side == 2 ? 2 * my : 2 * my + mx;
    const int kind = L.ghost[(size_t)P.gbase + off + t].kind;
    return kind == G_COPY ? 3 : kind == G_QIF ? 4 : kind == G_QIF_END ? 5 : 6;
  };
  int x = 0, y = 0;
  if (i == mx - 1) x = 2 * form(1, j) + 1;
  else if (i == 0) x = 2 * form(0, j);
  if (j == my - 1) y = 2 * form(3, i) + 1;
  else if (j == 0) y = 2 * form(2, i);
  if (!x && !y) return 0;
  return (unsigned char)(1 + x * 15 + y);
}

bool build_region_tiles(const HLevel& L, int rdepth, RegionBuild& out) {
  out = RegionBuild();
  if (L.np == 0 || rdepth < 2 || rdepth > kRegionMaxDepth || rdepth % 2) return false;
  for (const PatchInfo& P : L.pi)
    if (P.mx < 2 || P.my < 2 || P.mx > 4095 || P.my > 4095) return false;
  OwnerMap own;
  own.build(L);
  const int n0 = L.n[0];
  // cell id = global index (0-based) gi + gj * n0
  auto cell_of = [&](int p, int i, int j) -> long long {
    return (long long)(L.box[p].ilo - 1 + i) + (long long)(L.box[p].jlo - 1 + j) * n0;
  };
  // dependencies of cell (p, i, j); false when a source is not where the
  // fixed stencil puts it
  std::vector<long long> dep;  // the last deps() result, global ids
  auto deps = [&](int p, int i, int j) -> bool {
    dep.clear();
    const PatchInfo& P = L.pi[p];
    const Box& b = L.box[p];
    const int mx = P.mx, my = P.my;
    dep.push_back(cell_of(p, i, j));
    const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
    for (int d = 0; d < 4; ++d) {
      const int ni = i + di[d], nj = j + dj[d];
      if (ni >= 0 && ni < mx && nj >= 0 && nj < my) {
        dep.push_back(cell_of(p, ni, nj));
        continue;
      }
      const int ii = i - di[d], ij = j - dj[d];  // inward neighbour
      if (P.bc[d]) {
        if (P.bc[d] == 1) dep.push_back(cell_of(p, ii, ij));
        continue;
      }
      const int t = d < 2 ? j : i, len = d < 2 ? my : mx, u = t + 1;
      const int off = d == 0 ? 0 : d == 1 ? my : d == 2 ? 2 * my : 2 * my + mx;
      const GhostRec& r = L.ghost[P.gbase + off + t];
      const int ai = d < 2 ? 0 : 1, aj = d < 2 ? 1 : 0;  // along-face step
      auto inpatch = [&](int a, int c) { return a >= 0 && a < mx && c >= 0 && c < my; };
      switch (r.kind) {
        case G_COPY: {
          const int q = own.at(b.ilo + ni, b.jlo + nj) - 1;
          if (q < 0 || q == p) return false;
          const int si = b.ilo + ni - L.box[q].ilo, sj = b.jlo + nj - L.box[q].jlo;
          if (r.s[0] != qidx(L.pi[q], si + 1, sj + 1)) return false;
          dep.push_back(cell_of(q, si, sj));
          break;
        }
        case G_QIF: {
          if (u == 1 || u == len) return false;
          const int s = (u & 1) == 0 ? 1 : -1;
          if (!inpatch(i + s * ai, j + s * aj) || !inpatch(ii + 2 * s * ai, ij + 2 * s * aj)) return false;
          dep.push_back(cell_of(p, i + s * ai, j + s * aj));
          dep.push_back(cell_of(p, ii + 2 * s * ai, ij + 2 * s * aj));
          break;
        }
        case G_QIF_END:
          dep.push_back(cell_of(p, ii, ij));
          break;
        case G_LIF: {
          const int s = (u & 1) == 0 ? -1 : 1;
          if (!inpatch(i + s * ai, j + s * aj)) return false;
          dep.push_back(cell_of(p, i + s * ai, j + s * aj));
          break;
        }
        default:
          return false;
      }
    }
    return true;
  };
  // the dependencies of every cell, once (CSR over the level's cells in
  // patch order, entries as global ids)
  const long long nglob = (long long)n0 * L.n[1];
  std::vector<int> cell_ix((size_t)nglob, -1);  // global id -> level cell index (patch order)
  std::vector<long long> gid;                    // level cell index -> global id
  // face: 0 off every face, else the form of the cell's face ghosts (below),
  // so that a wave of face cells takes one branch of the update
  std::vector<unsigned char> colour, face;
  std::vector<int> cpatch, ci_, cj_;  // owner patch and local position of each cell
  gid.reserve((size_t)L.ncell);
  for (int p = 0; p < L.np; ++p)
    for (int j = 0; j < L.pi[p].my; ++j)
      for (int i = 0; i < L.pi[p].mx; ++i) {
        const long long g = cell_of(p, i, j);
        cell_ix[(size_t)g] = (int)gid.size();
        gid.push_back(g);
        colour.push_back((unsigned char)((i + j + L.pi[p].par0) & 1));
        face.push_back(face_form(L, p, i, j));
        cpatch.push_back(p);
        ci_.push_back(i);
        cj_.push_back(j);
      }
  const int ncell = (int)gid.size();
  std::vector<int> dptr(ncell + 1, 0), dlist;
  dlist.reserve((size_t)ncell * 6);
  {
    int x = 0;
    for (int p = 0; p < L.np; ++p)
      for (int j = 0; j < L.pi[p].my; ++j)
        for (int i = 0; i < L.pi[p].mx; ++i, ++x) {
          if (!deps(p, i, j)) return false;
          for (long long y : dep) dlist.push_back(cell_ix[(size_t)y]);
          dptr[x + 1] = (int)dlist.size();
        }
  }
  std::vector<int> depth_of(ncell, -1);  // reset after each draft through `order`
  std::vector<char> mark(ncell, 0);      // 1: own, 2: virtual own (reset after each draft)
  // level cell of an interior Q index (-1: a ghost slot)
  std::vector<int> qcell((size_t)std::max(1, L.nq), -1);
  for (int p = 0; p < L.np; ++p)
    for (int j = 0; j < L.pi[p].my; ++j)
      for (int i = 0; i < L.pi[p].mx; ++i) qcell[(size_t)qidx(L.pi[p], i + 1, j + 1)] = cell_ix[(size_t)cell_of(p, i, j)];
  // footprint item of each coarse cell, keyed by its first fine cell
  std::vector<int> foot_at(ncell, -1);
  for (size_t f = 0; f < L.foot.size(); ++f) {
    const FootRec& fr = L.foot[f];
    foot_at[(size_t)cell_ix[(size_t)cell_of(fr.p, 2 * fr.ci, 2 * fr.cj)]] = (int)f;
  }
  // the ghost records next to interior cell (p, i, j): its face positions and
  // corner (records per patch: W my, E my, S mx, N mx, then the 4 corners)
  auto cell_records = [&](int p, int i, int j, std::vector<int>& recs) {
    const PatchInfo& P = L.pi[p];
    const int b = P.gbase, mx = P.mx, my = P.my;
    if (i == 0) recs.push_back(b + j);
    if (i == mx - 1) recs.push_back(b + my + j);
    if (j == 0) recs.push_back(b + 2 * my + i);
    if (j == my - 1) recs.push_back(b + 2 * my + mx + i);
    const int c0 = b + 2 * my + 2 * mx;
    if (i == 0 && j == 0) recs.push_back(c0);
    if (i == mx - 1 && j == 0) recs.push_back(c0 + 1);
    if (i == 0 && j == my - 1) recs.push_back(c0 + 2);
    if (i == mx - 1 && j == my - 1) recs.push_back(c0 + 3);
  };
  // the fine Q indices a record's value reads (-1 entries unused); false for
  // a kind the tile does not evaluate
  auto rec_sources = [&](int rec, int q[2]) -> bool {
    const GhostRec& r = L.ghost[(size_t)rec];
    q[0] = q[1] = -1;
    switch (r.kind) {
      case G_COPY: case G_LIF_CORNER: case G_NEU: q[0] = r.s[0]; return true;
      case G_QIF: case G_QIF_END: case G_LIF: case G_DIR: q[0] = r.s[0]; q[1] = r.s[1]; return true;
      default: return false;
    }
  };
  // a tile drafted from a set of own cells (level cell indices)
  struct Draft {
    std::vector<int> own;
    int cx0 = 0, cy0 = 0, cw = 0, ch = 0;
    std::vector<int> pl;  // region patches
    std::vector<unsigned> load, red, black;
    std::vector<int> rcell, rdep;  // the region's cells and their depths
    std::vector<RegionGhost> gh;   // ghost records the tile writes
    std::vector<RegionRestrict> rs;// coarse cells it restricts
  };
  const int per_colour = kRegionThreads * kRegionK, per_load = kRegionThreads * kRegionKL;
  // false when the region does not fit a workgroup (or a ghost record is of
  // a kind the tile does not evaluate: `bad`)
  bool bad = false;
  auto draft = [&](Draft& D) -> bool {
    std::vector<int> order;  // level cell indices of the region, in discovery order
    for (int c : D.own) {
      depth_of[c] = rdepth;
      mark[c] = 1;
      order.push_back(c);
    }
    auto need_full = [&](int c) {  // a cell this tile's writes read at its final value
      if (c < 0 || mark[c]) return;
      mark[c] = 2;
      depth_of[c] = rdepth;
      order.push_back(c);
    };
    // ghost records next to own cells, and the coarse cells whose first fine
    // cell is own: their fine sources are needed at their final values
    std::vector<int> recs, blocks;
    for (int c : D.own) {
      cell_records(cpatch[c], ci_[c], cj_[c], recs);
      if (foot_at[c] >= 0) blocks.push_back(c);
    }
    for (int rec : recs) {
      const GhostRec& r = L.ghost[(size_t)rec];
      int q[2];
      if (r.kind == G_PCORNER) {
        for (int side = 0; side < 2; ++side) {
          if (!rec_sources(r.s[side], q)) { bad = true; return false; }
          for (int t = 0; t < 2; ++t)
            if (q[t] >= 0) need_full(qcell[(size_t)q[t]]);
        }
        need_full(qcell[(size_t)r.s[2]]);
      } else {
        if (!rec_sources(rec, q)) { bad = true; return false; }
        for (int t = 0; t < 2; ++t)
          if (q[t] >= 0) {
            if (qcell[(size_t)q[t]] < 0) { bad = true; return false; }
            need_full(qcell[(size_t)q[t]]);
          }
      }
    }
    for (int c : blocks) {
      const int p = cpatch[c], i = ci_[c], j = cj_[c];
      need_full(cell_ix[(size_t)cell_of(p, i + 1, j)]);
      need_full(cell_ix[(size_t)cell_of(p, i, j + 1)]);
      need_full(cell_ix[(size_t)cell_of(p, i + 1, j + 1)]);
    }
    // pass q updates colour (q - 1) & 1 (0: colour 1)
    for (int q = rdepth; q >= 1; --q) {
      const int col = (q - 1) & 1;
      const size_t n = order.size();
      for (size_t x = 0; x < n; ++x) {
        const int c = order[x];
        if (depth_of[c] < q || colour[c] != col) continue;
        for (int e = dptr[c]; e < dptr[c + 1]; ++e) {
          const int y = dlist[e];
          if (depth_of[y] < 0) {
            depth_of[y] = q - 1;
            order.push_back(y);
          } else if (depth_of[y] < q - 1) {
            depth_of[y] = q - 1;
          }
        }
      }
    }
    // canvas: bounding box of the region plus one slot of margin
    int gx0 = 1 << 30, gx1 = -1, gy0 = 1 << 30, gy1 = -1;
    for (int c : order) {
      const int gi = (int)(gid[c] % n0), gj = (int)(gid[c] / n0);
      gx0 = std::min(gx0, gi); gx1 = std::max(gx1, gi);
      gy0 = std::min(gy0, gj); gy1 = std::max(gy1, gj);
    }
    D.cx0 = gx0 - 1; D.cy0 = gy0 - 1; D.cw = gx1 - gx0 + 3; D.ch = gy1 - gy0 + 3;
    auto slot_of = [&](int c) -> unsigned short {
      const int gi = (int)(gid[c] % n0), gj = (int)(gid[c] / n0);
      return (unsigned short)((gi - D.cx0) + (gj - D.cy0) * D.cw);
    };
    // level cell indices run in patch order, row by row: each list is sorted
    // by the cells' face form, then in (patch, j, i) order -- the cells off
    // every face first (the kernel's short path), then the face cells grouped
    // by the form of their ghosts, so that the waves of a pass take one
    // branch each instead of the union of their lanes' branches
    std::vector<int> sorted(order);
    std::sort(sorted.begin(), sorted.end(), [&](int a, int b) { return face[a] != face[b] ? face[a] < face[b] : a < b; });
    D.pl.clear(); D.load.clear(); D.red.clear(); D.black.clear();
    D.rcell = order;
    D.rdep.resize(order.size());
    for (size_t x = 0; x < order.size(); ++x) D.rdep[x] = depth_of[order[x]];
    bool fit = (long long)D.cw * D.ch <= kRegionMaxSlots;
    for (int c : sorted) {
      const int cp = cpatch[c], ci = ci_[c], cj = cj_[c];
      int ps = -1;
      for (size_t k = 0; k < D.pl.size(); ++k)
        if (D.pl[k] == cp) ps = (int)k;
      if (ps < 0) {
        if ((int)D.pl.size() >= kRegionMaxPatches) { fit = false; break; }
        D.pl.push_back(cp);
        ps = (int)D.pl.size() - 1;
      }
      const int d = mark[c] == 2 ? kRegionVirtual : depth_of[c];
      const unsigned code = region_code(ps, ci, cj, d);
      if (d == 0) D.load.push_back(code);
      else if (colour[c] == 0) D.red.push_back(code);
      else D.black.push_back(code);
    }
    D.gh.clear();
    D.rs.clear();
    if (fit) {
      auto is_cf = [](int kind) { return kind == G_QIF || kind == G_QIF_END || kind == G_LIF || kind == G_LIF_CORNER; };
      for (int rec : recs) {
        const GhostRec& r = L.ghost[(size_t)rec];
        RegionGhost g;
        std::memset(&g, 0, sizeof g);
        g.dst = r.dst;
        g.rec = g.reca = g.recb = -1;
        unsigned short sl[6];
        for (int t = 0; t < 6; ++t) sl[t] = kRegionNoSlot;
        int q[2];
        if (r.kind == G_PCORNER) {
          const GhostRec& ra = L.ghost[(size_t)r.s[0]];
          const GhostRec& rb = L.ghost[(size_t)r.s[1]];
          g.kinds = G_PCORNER | (ra.kind << 4) | (rb.kind << 8);
          g.ca = ra.c;
          g.cb = rb.c;
          if (is_cf(ra.kind)) g.reca = r.s[0];
          if (is_cf(rb.kind)) g.recb = r.s[1];
          for (int side = 0; side < 2; ++side) {
            rec_sources(r.s[side], q);
            for (int t = 0; t < 2; ++t)
              if (q[t] >= 0) sl[2 * side + t] = slot_of(qcell[(size_t)q[t]]);
          }
          sl[4] = slot_of(qcell[(size_t)r.s[2]]);
        } else {
          g.kinds = r.kind;
          g.c = r.c;
          if (is_cf(r.kind)) g.rec = rec;
          rec_sources(rec, q);
          for (int t = 0; t < 2; ++t)
            if (q[t] >= 0) sl[t] = slot_of(qcell[(size_t)q[t]]);
        }
        g.s01 = (unsigned)sl[0] | ((unsigned)sl[1] << 16);
        g.s23 = (unsigned)sl[2] | ((unsigned)sl[3] << 16);
        g.s45 = (unsigned)sl[4] | ((unsigned)sl[5] << 16);
        D.gh.push_back(g);
      }
      for (int c : blocks) {
        const FootRec& fr = L.foot[(size_t)foot_at[c]];
        RegionRestrict e;
        e.qc = fr.qcoff + (fr.ci + 1) + (fr.cj + 1) * (fr.mx / 2 + 2);
        e.pq = fr.pq;
        e.slot = slot_of(c);
        e.pad = foot_at[c];
        D.rs.push_back(e);
      }
    }
    for (int c : order) {
      depth_of[c] = -1;
      mark[c] = 0;
    }
    return fit && (int)D.red.size() <= per_colour && (int)D.black.size() <= per_colour &&
           (int)D.load.size() <= per_load && (int)recs.size() <= kRegionThreads &&
           (int)blocks.size() <= kRegionThreads;
  };
  // 1. every patch, halved along its longer side until the region of every
  //    piece fits a workgroup
  std::vector<Draft> drafts;
  for (int p = 0; p < L.np; ++p) {
    std::vector<std::array<int, 5>> stack{{p, 0, L.pi[p].mx, 0, L.pi[p].my}};
    while (!stack.empty()) {
      const std::array<int, 5> r = stack.back();
      stack.pop_back();
      Draft D;
      for (int j = r[3]; j < r[4]; ++j)
        for (int i = r[1]; i < r[2]; ++i) D.own.push_back(cell_ix[(size_t)cell_of(r[0], i, j)]);
      if (draft(D)) {
        drafts.push_back(std::move(D));
        continue;
      }
      if (bad) return false;
      const int w0 = r[2] - r[1], h0 = r[4] - r[3];
      if (w0 <= 2 && h0 <= 2) return false;
      // the second half is pushed first so that pieces stay in spatial order
      if (w0 >= h0) {
        const int m = r[1] + w0 / 2;
        stack.push_back({r[0], m, r[2], r[3], r[4]});
        stack.push_back({r[0], r[1], m, r[3], r[4]});
      } else {
        const int m = r[3] + h0 / 2;
        stack.push_back({r[0], r[1], r[2], m, r[4]});
        stack.push_back({r[0], r[1], r[2], r[3], m});
      }
    }
  }
  // 2. adjacent pieces merged while the merged region still fits: fewer,
  //    fuller workgroups, and the cells between merged pieces are computed
  //    once instead of in both halos.  Smallest first, each with the smallest
  //    neighbour that fits; passes until nothing merges.
  std::vector<int> tile_of(ncell, -1);
  for (size_t t = 0; t < drafts.size(); ++t)
    for (int c : drafts[t].own) tile_of[c] = (int)t;
  std::vector<char> alive(drafts.size(), 1);
  auto neighbours = [&](int t) {
    std::vector<int> nb;
    for (int c : drafts[t].own) {
      const long long g = gid[c];
      const int gi = (int)(g % n0), gj = (int)(g / n0);
      const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
      for (int d = 0; d < 4; ++d) {
        const int ni = gi + di[d], nj = gj + dj[d];
        if (ni < 0 || nj < 0 || ni >= n0 || nj >= L.n[1]) continue;
        const int y = cell_ix[(size_t)ni + (size_t)nj * n0];
        if (y < 0 || tile_of[y] == t) continue;
        if (std::find(nb.begin(), nb.end(), tile_of[y]) == nb.end()) nb.push_back(tile_of[y]);
      }
    }
    std::sort(nb.begin(), nb.end(), [&](int a, int b) {
      return drafts[a].own.size() != drafts[b].own.size() ? drafts[a].own.size() < drafts[b].own.size() : a < b;
    });
    return nb;
  };
  // whether the region of A's and B's own cells fits: that region is the
  // union of theirs, each cell at the larger of its two depths (the walk is
  // a max over own cells), so the counts follow without a walk
  auto can_merge = [&](const Draft& A, const Draft& B) -> bool {
    int npl = (int)A.pl.size();
    for (int p : B.pl)
      if (std::find(A.pl.begin(), A.pl.end(), p) == A.pl.end() && ++npl > kRegionMaxPatches) return false;
    const long long x0 = std::min(A.cx0, B.cx0), x1 = std::max(A.cx0 + A.cw, B.cx0 + B.cw);
    const long long y0 = std::min(A.cy0, B.cy0), y1 = std::max(A.cy0 + A.ch, B.cy0 + B.ch);
    if ((x1 - x0) * (y1 - y0) > kRegionMaxSlots) return false;
    for (size_t x = 0; x < A.rcell.size(); ++x) depth_of[A.rcell[x]] = A.rdep[x];
    for (size_t x = 0; x < B.rcell.size(); ++x) depth_of[B.rcell[x]] = std::max(depth_of[B.rcell[x]], B.rdep[x]);
    int nred = 0, nblack = 0, nload = 0;
    auto count = [&](int c) {
      if (depth_of[c] < 0) return;
      if (depth_of[c] == 0) ++nload;
      else if (colour[c] == 0) ++nred;
      else ++nblack;
      depth_of[c] = -1;
    };
    for (int c : A.rcell) count(c);
    for (int c : B.rcell) count(c);
    return nred <= per_colour && nblack <= per_colour && nload <= per_load;
  };
  for (bool changed = true; changed;) {
    changed = false;
    std::vector<int> order;
    for (size_t t = 0; t < drafts.size(); ++t)
      if (alive[t]) order.push_back((int)t);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return drafts[a].own.size() < drafts[b].own.size(); });
    for (int t : order) {
      if (!alive[t]) continue;
      for (int u : neighbours(t)) {
        if ((long long)drafts[t].own.size() + (long long)drafts[u].own.size() > 2LL * per_colour) continue;
        if (!can_merge(drafts[t], drafts[u])) continue;
        Draft M;
        M.own = drafts[t].own;
        M.own.insert(M.own.end(), drafts[u].own.begin(), drafts[u].own.end());
        if (!draft(M)) return false;  // can_merge said it fits
        for (int c : drafts[t].own) tile_of[c] = u;
        drafts[u] = std::move(M);
        drafts[t] = Draft();
        alive[t] = 0;
        changed = true;
        break;
      }
    }
  }
  // 3. the tiles in spatial order (the launch gives each XCD a contiguous
  //    range): by the first own cell of each, in level cell order
  std::vector<int> keep;
  for (size_t t = 0; t < drafts.size(); ++t)
    if (alive[t]) {
      std::sort(drafts[t].own.begin(), drafts[t].own.end());
      keep.push_back((int)t);
    }
  std::sort(keep.begin(), keep.end(), [&](int a, int b) { return drafts[a].own[0] < drafts[b].own[0]; });
  for (int t : keep) {
    const Draft& D = drafts[t];
    RegionTile rt;
    std::memset(&rt, 0, sizeof rt);
    rt.cx0 = D.cx0; rt.cy0 = D.cy0; rt.cw = D.cw; rt.ch = D.ch;
    rt.np = (int)D.pl.size();
    rt.nload = (int)D.load.size(); rt.nred = (int)D.red.size(); rt.nblack = (int)D.black.size();
    rt.gofs = (int)out.ghosts.size(); rt.ngh = (int)D.gh.size();
    rt.rofs = (int)out.rs.size(); rt.nrs = (int)D.rs.size();
    out.ghosts.insert(out.ghosts.end(), D.gh.begin(), D.gh.end());
    out.rs.insert(out.rs.end(), D.rs.begin(), D.rs.end());
    out.max_gh = std::max(out.max_gh, rt.ngh);
    out.max_rs = std::max(out.max_rs, rt.nrs);
    for (int k = 0; k < kRegionMaxPatches; ++k) {
      RegionPatch rp;
      std::memset(&rp, 0, sizeof rp);
      if (k < (int)D.pl.size()) {
        rp.pi = L.pi[D.pl[k]];
        rp.ox = L.box[D.pl[k]].ilo - 1 - D.cx0;
        rp.oy = L.box[D.pl[k]].jlo - 1 - D.cy0;
      }
      out.patches.push_back(rp);
      out.pid.push_back(k < (int)D.pl.size() ? D.pl[k] : -1);
    }
    const size_t base = out.cells.size();
    out.cells.resize(base + kRegionStride, kRegionNone);
    std::copy(D.red.begin(), D.red.end(), out.cells.begin() + base);
    std::copy(D.black.begin(), D.black.end(), out.cells.begin() + base + per_colour);
    std::copy(D.load.begin(), D.load.end(), out.cells.begin() + base + 2 * per_colour);
    out.tiles.push_back(rt);
    out.max_slots = std::max(out.max_slots, D.cw * D.ch);
  }
  return !out.tiles.empty();
}

// Parent-window sources of the coarse-fine part of ghost record `rec` of
// level L (the stencils of cf_rec and of the coarse-fine corners): n window
// coordinates (wx, wy) of the record's patch, in the order the record's
// parent indices s[2..4] hold them; 0 for a record without a parent part.
// rp: each record's patch.
static int cf_window_sources(const HLevel& L, const std::vector<int>& rp, int rec, int wx[3], int wy[3]) {
  const GhostRec& r = L.ghost[(size_t)rec];
  if (r.kind != G_QIF && r.kind != G_QIF_END && r.kind != G_LIF && r.kind != G_LIF_CORNER) return 0;
  const PatchInfo& P = L.pi[(size_t)rp[(size_t)rec]];
  const int mx = P.mx, my = P.my, cmx = mx / 2, cmy = my / 2;
  int t = rec - P.gbase, side;
  if (t < my) side = 0;
  else if ((t -= my) < my) side = 1;
  else if ((t -= my) < mx) side = 2;
  else if ((t -= mx) < mx) side = 3;
  else {
    t -= mx;  // corner t: SW SE NW NE
    if (t < 0 || t > 3) return -1;
    wx[0] = (t & 1) ? cmx + 1 : 0;
    wy[0] = (t & 2) ? cmy + 1 : 0;
    return 1;
  }
  const int len = side < 2 ? my : mx, clen = side < 2 ? cmy : cmx, u = t + 1;
  const int fixed = side == 0 ? 0 : side == 1 ? cmx + 1 : side == 2 ? 0 : cmy + 1;
  int along[3], n = 1;
  if (r.kind == G_QIF_END) {
    n = 3;
    for (int k = 0; k < 3; ++k) along[k] = u == len ? clen + 1 - k : k;
  } else {
    along[0] = (u & 1) == 0 ? u / 2 : (u + 1) / 2;
  }
  for (int k = 0; k < n; ++k) {
    wx[k] = side < 2 ? fixed : along[k];
    wy[k] = side < 2 ? along[k] : fixed;
  }
  return n;
}

// The tables of the coarse-grid correction folded into a level's
// post-smoothing region launch (ebTransfer.m:57-71 before ebRelaxLevel.m:28-38):
// for every tile, per patch, the rectangle of the patch's QC window that its
// region cells' corrections (the 2 x 2 window cells of each cell) and the
// parent parts of its face cells' and ghost records' coarse-fine stencils
// read, staged as (QC index, parent index) cells; for every ghost record of
// the tile the window slots of its parent sources.  Every stencil is checked
// against the record's parent indices; false (the level keeps the separate
// correction launch) when one differs or a tile's window outgrows the
// workgroup.
bool build_region_windows(const HLevel& L, int parent_nq, RegionBuild& rb) {
  rb.win.clear();
  rb.gw.clear();
  rb.max_win = 0;
  if (L.wmap.empty() || rb.tiles.empty()) return false;
  std::vector<int> rp(L.ghost.size(), -1);
  for (int p = 0; p < L.np; ++p) {
    const PatchInfo& P = L.pi[(size_t)p];
    for (int t = 0; t < 2 * (P.mx + P.my) + 4; ++t) rp[(size_t)P.gbase + t] = p;
  }
  std::vector<int> rec_at((size_t)std::max(1, L.nq), -1);
  for (size_t r = 0; r < L.ghost.size(); ++r)
    if (L.ghost[r].dst >= 0 && L.ghost[r].dst < L.nq) rec_at[(size_t)L.ghost[r].dst] = (int)r;
  // every coarse-fine record's stencil against its parent indices
  for (size_t r = 0; r < L.ghost.size(); ++r) {
    int wx[3], wy[3];
    const int n = cf_window_sources(L, rp, (int)r, wx, wy);
    if (n < 0) return false;
    const PatchInfo& P = L.pi[(size_t)rp[r]];
    const int ldc = P.mx / 2 + 2;
    for (int k = 0; k < n; ++k) {
      if (wx[k] < 0 || wy[k] < 0 || wx[k] > P.mx / 2 + 1 || wy[k] > P.my / 2 + 1) return false;
      const int w = P.qcoff + wx[k] + wy[k] * ldc;
      if (w < 0 || w >= L.nqc || L.wmap[(size_t)w] != L.ghost[r].s[2 + k]) return false;
    }
  }
  const int per_colour = kRegionThreads * kRegionK;
  for (size_t x = 0; x < rb.tiles.size(); ++x) {
    RegionTile& T = rb.tiles[x];
    RegionPatch* sp = rb.patches.data() + x * kRegionMaxPatches;
    const int* pid = rb.pid.data() + x * kRegionMaxPatches;
    int lo[kRegionMaxPatches][2], hi[kRegionMaxPatches][2];
    for (int k = 0; k < kRegionMaxPatches; ++k) lo[k][0] = lo[k][1] = 1 << 30, hi[k][0] = hi[k][1] = -1;
    auto add = [&](int ps, int wx, int wy) {
      lo[ps][0] = std::min(lo[ps][0], wx); hi[ps][0] = std::max(hi[ps][0], wx);
      lo[ps][1] = std::min(lo[ps][1], wy); hi[ps][1] = std::max(hi[ps][1], wy);
    };
    auto slot_of_patch = [&](int p) {
      for (int k = 0; k < T.np; ++k)
        if (pid[k] == p) return k;
      return -1;
    };
    const unsigned* cells = rb.cells.data() + x * kRegionStride;
    auto each_cell = [&](auto&& fn) {
      for (int e = 0; e < kRegionStride; ++e) {
        if (cells[e] == kRegionNone) continue;
        const unsigned c = cells[e];
        fn((int)(c & 15), (int)((c >> 4) & 4095), (int)((c >> 16) & 4095), e < 2 * per_colour);
      }
    };
    // the cells' corrections and the updated face cells' parent parts
    bool ok = true;
    each_cell([&](int ps, int i, int j, bool updated) {
      const int wx = (i >> 1) + 1, wy = (j >> 1) + 1;
      add(ps, wx, wy);
      add(ps, wx + ((i & 1) ? 1 : -1), wy + ((j & 1) ? 1 : -1));
      if (!updated) return;
      const PatchInfo& P = sp[ps].pi;
      const int recs[4] = {i == 0 ? P.gbase + j : -1, i == P.mx - 1 ? P.gbase + P.my + j : -1,
                           j == 0 ? P.gbase + 2 * P.my + i : -1, j == P.my - 1 ? P.gbase + 2 * P.my + P.mx + i : -1};
      for (int side = 0; side < 4; ++side) {
        if (recs[side] < 0 || P.bc[side]) continue;
        int sx[3], sy[3];
        const int n = cf_window_sources(L, rp, recs[side], sx, sy);
        if (n < 0) ok = false;
        for (int k = 0; k < n; ++k) add(ps, sx[k], sy[k]);
      }
    });
    if (!ok) return false;
    // the ghost records' parent parts
    std::vector<RegionGhostW> gw((size_t)T.ngh);
    for (int g = 0; g < T.ngh; ++g) {
      RegionGhostW& o = gw[(size_t)g];
      for (int k = 0; k < 8; ++k) o.s[k] = kRegionNoSlot;
      const RegionGhost& G = rb.ghosts[(size_t)T.gofs + g];
      const int rec = rec_at[(size_t)G.dst];
      if (rec < 0) return false;
      const GhostRec& r = L.ghost[(size_t)rec];
      const int parts[2] = {r.kind == G_PCORNER ? r.s[0] : rec, r.kind == G_PCORNER ? r.s[1] : -1};
      for (int q = 0; q < 2; ++q) {
        if (parts[q] < 0) continue;
        int sx[3], sy[3];
        const int n = cf_window_sources(L, rp, parts[q], sx, sy);
        if (n < 0) return false;
        if (n == 0) continue;
        const int ps = slot_of_patch(rp[(size_t)parts[q]]);
        if (ps < 0) return false;
        for (int k = 0; k < n; ++k) add(ps, sx[k], sy[k]);
      }
    }
    // the rectangles, their window cells and the records' slots
    T.wofs = (int)rb.win.size();
    int nwin = 0;
    for (int ps = 0; ps < T.np; ++ps) {
      RegionPatch& R = sp[ps];
      if (hi[ps][0] < 0) { R.eoff = nwin; R.ex0 = R.ey0 = 0; R.ew = 1; continue; }
      const int ldc = R.pi.mx / 2 + 2;
      if (lo[ps][0] < 0 || lo[ps][1] < 0 || hi[ps][0] > R.pi.mx / 2 + 1 || hi[ps][1] > R.pi.my / 2 + 1) return false;
      R.eoff = nwin; R.ex0 = lo[ps][0]; R.ey0 = lo[ps][1]; R.ew = hi[ps][0] - lo[ps][0] + 1;
      for (int wy = lo[ps][1]; wy <= hi[ps][1]; ++wy)
        for (int wx = lo[ps][0]; wx <= hi[ps][0]; ++wx) {
          const int w = R.pi.qcoff + wx + wy * ldc;
          if (w < 0 || w >= L.nqc) return false;
          const int wm = L.wmap[(size_t)w];
          if (wm < 0 || wm >= parent_nq) return false;
          rb.win.push_back(RegionWin{w, wm});
          ++nwin;
        }
    }
    if (nwin > kRegionMaxWin) return false;
    T.nwin = nwin;
    rb.max_win = std::max(rb.max_win, nwin);
    for (int g = 0; g < T.ngh; ++g) {
      RegionGhostW& o = gw[(size_t)g];
      const RegionGhost& G = rb.ghosts[(size_t)T.gofs + g];
      const GhostRec& r = L.ghost[(size_t)rec_at[(size_t)G.dst]];
      const int parts[2] = {r.kind == G_PCORNER ? r.s[0] : rec_at[(size_t)G.dst], r.kind == G_PCORNER ? r.s[1] : -1};
      for (int q = 0; q < 2; ++q) {
        if (parts[q] < 0) continue;
        int sx[3], sy[3];
        const int n = cf_window_sources(L, rp, parts[q], sx, sy);
        if (n <= 0) continue;
        const RegionPatch& R = sp[slot_of_patch(rp[(size_t)parts[q]])];
        for (int k = 0; k < n; ++k)
          o.s[3 * q + k] = (unsigned short)(R.eoff + (sx[k] - R.ex0) + (sy[k] - R.ey0) * R.ew);
      }
    }
    rb.gw.insert(rb.gw.end(), gw.begin(), gw.end());
  }
  return true;
}

// The parent ghost records behind the FAS load's parent reads of level F
// (window cells and ring items whose source is a ghost slot of C), and
// whether all of them are physical: side records of a Dirichlet or Neumann
// face, or physical corners between two such sides.
void parent_ghost_sources(HLevel& F, const HLevel& C) {
  std::vector<int> rec_at((size_t)std::max(1, C.nq), -1);
  for (size_t r = 0; r < C.ghost.size(); ++r)
    if (C.ghost[r].dst >= 0 && C.ghost[r].dst < C.nq) rec_at[(size_t)C.ghost[r].dst] = (int)r;
  auto physical = [&](int r) {
    const int kind = C.ghost[(size_t)r].kind;
    return kind == G_DIR || kind == G_NEU;
  };
  bool ok = true;
  auto note = [&](int q) -> int {
    if (q < 0 || q >= C.nq) return -1;
    const int r = rec_at[(size_t)q];
    if (r < 0) return -1;
    const GhostRec& g = C.ghost[(size_t)r];
    if (g.kind == G_PCORNER) {
      if (!physical(g.s[0]) || !physical(g.s[1])) ok = false;
    } else if (!physical(r)) {
      ok = false;
    }
    return r;
  };
  F.wgr.assign(F.wmap.size(), -1);
  for (size_t w = 0; w < F.wmap.size(); ++w) F.wgr[w] = note(F.wmap[w]);
  F.ring_rec.assign(F.ring_all.size(), -1);
  for (size_t t = 0; t < F.ring_all.size(); ++t) F.ring_rec[t] = note(F.ring_all[t].src);
  F.pgh_ok = ok ? 1 : 0;
}

// Every index the region kernel dereferences, against its buffer: each cell
// entry inside its tile's patches and canvas, each ghost record's
// destination and slots, each footprint item and its block's slots; every
// ghost record and footprint item is written by exactly one tile.
int check_region_tables(const HLevel& L, const RegionBuild& rb) {
  auto bad = [](const char* what) {
    return fail(BSAM_E_ARG, "bsamgpu:internal: region table %s out of range", what);
  };
  std::vector<int> gseen(L.ghost.size(), 0), fseen(L.foot.size(), 0);
  std::vector<int> rec_at((size_t)std::max(1, L.nq), -1);  // ghost record of each Q destination
  for (size_t r = 0; r < L.ghost.size(); ++r)
    if (L.ghost[r].dst >= 0 && L.ghost[r].dst < L.nq) rec_at[(size_t)L.ghost[r].dst] = (int)r;
  for (size_t t = 0; t < rb.tiles.size(); ++t) {
    const RegionTile& T = rb.tiles[t];
    const long long slots = (long long)T.cw * T.ch;
    if (T.cw < 3 || T.ch < 3 || slots > kRegionMaxSlots || T.np < 1 || T.np > kRegionMaxPatches) return bad("tile");
    const unsigned* c = rb.cells.data() + t * kRegionStride;
    const int per_colour = kRegionThreads * kRegionK;
    for (int e = 0; e < kRegionStride; ++e) {
      if (c[e] == kRegionNone) continue;
      const int ps = c[e] & 15, i = (c[e] >> 4) & 4095, j = (c[e] >> 16) & 4095;
      if (ps >= T.np) return bad("patch slot");
      const RegionPatch& rp = rb.patches[t * kRegionMaxPatches + ps];
      const int x = rp.ox + i, y = rp.oy + j;
      if (i >= rp.pi.mx || j >= rp.pi.my || x < 1 || y < 1 || x > T.cw - 2 || y > T.ch - 2) return bad("cell");
      (void)per_colour;
    }
    if (T.gofs < 0 || T.ngh < 0 || (size_t)T.gofs + T.ngh > rb.ghosts.size()) return bad("ghost list");
    if (T.ngh > kRegionThreads || T.nrs > kRegionThreads) return bad("list length");
    for (int g = 0; g < T.ngh; ++g) {
      const RegionGhost& G = rb.ghosts[(size_t)T.gofs + g];
      if (G.dst < 0 || G.dst >= L.nq) return bad("ghost destination");
      for (int rec : {G.rec, G.reca, G.recb})
        if (rec < -1 || rec >= (int)L.ghost.size()) return bad("ghost record");
      const int kind = G.kinds & 15;
      if (kind > G_PCORNER || region_slot(G, 0) == kRegionNoSlot ||
          (kind == G_PCORNER && (region_slot(G, 2) == kRegionNoSlot || region_slot(G, 4) == kRegionNoSlot)))
        return bad("ghost form");
      for (int k = 0; k < 6; ++k)
        if (region_slot(G, k) != kRegionNoSlot && (long long)region_slot(G, k) >= slots) return bad("ghost slot");
      if (rec_at[(size_t)G.dst] < 0) return bad("ghost destination record");
      ++gseen[(size_t)rec_at[(size_t)G.dst]];
    }
    if (T.rofs < 0 || T.nrs < 0 || (size_t)T.rofs + T.nrs > rb.rs.size()) return bad("restriction list");
    for (int e = 0; e < T.nrs; ++e) {
      const RegionRestrict& r = rb.rs[(size_t)T.rofs + e];
      if (r.pad < 0 || (size_t)r.pad >= L.foot.size() || r.slot < 0 || (long long)r.slot + T.cw + 1 >= slots ||
          r.qc < 0 || r.qc >= L.nqc)
        return bad("restriction entry");
      ++fseen[(size_t)r.pad];
    }
  }
  for (int v : gseen)
    if (v != 1) return bad("ghost assignment");
  for (int v : fseen)
    if (v != 1) return bad("restriction assignment");
  return 0;
}

void build_phys_level(HLevel& L) {
  L.ghost.clear();
  L.klass.clear();
  L.klass_off.clear();
  for (int p = 0; p < L.np; ++p) {
    L.pi[p].gbase = (int)L.ghost.size();
    const PatchInfo& P = L.pi[p];
    L.klass_off.push_back((int)L.klass.size());
    int first[4];
    for (int side = 0; side < 4; ++side) {
      first[side] = (int)L.ghost.size();
      const int len = side < 2 ? P.my : P.mx;
      for (int t = 0; t < len; ++t) {
        L.ghost.push_back(phys_rec(L, P, side, t));
        L.klass.push_back(1);
      }
    }
    const int mx = P.mx, my = P.my;
    const int cdst[4] = {qidx(P, 0, 0), qidx(P, mx + 1, 0), qidx(P, 0, my + 1), qidx(P, mx + 1, my + 1)};
    const int cs0[4] = {first[2], first[2] + mx - 1, first[3], first[3] + mx - 1};
    const int cs1[4] = {first[0], first[1], first[0] + my - 1, first[1] + my - 1};
    const int cs2[4] = {qidx(P, 1, 1), qidx(P, mx, 1), qidx(P, 1, my), qidx(P, mx, my)};
    for (int c = 0; c < 4; ++c) {
      GhostRec r{};
      r.dst = cdst[c]; r.kind = G_PCORNER; r.s[0] = cs0[c]; r.s[1] = cs1[c]; r.s[2] = cs2[c];
      L.ghost.push_back(r);
      L.klass.push_back(1);
    }
  }
}

// Tables of level k (F) against its parent level k-1 (C): ebBuildOps.m:17-199
// and ebGatherParentCoeffs.m:12-63.
int build_level_tables(HLevel& F, HLevel& C, bool qif) {
  OwnerMap OF, OC;
  OF.build(F);
  OC.build(C);
  F.ghost.clear(); F.ring_all.clear(); F.ring_cf.clear(); F.foot.clear(); F.corr.clear();
  F.corr_target.clear(); F.klass.clear(); F.klass_off.clear();
  F.wmap.assign(F.nqc, -1);
  F.wmap_patch.assign(F.nqc, -1);
  std::vector<uint8_t> cov(C.ncell, 0);
  for (int p = 0; p < F.np; ++p) {
    const Box& b = F.box[p];
    const PatchInfo& P = F.pi[p];
    if (b.ilo % 2 != 1 || b.ihi % 2 != 0 || b.jlo % 2 != 1 || b.jhi % 2 != 0)
      return fail(BSAM_E_ALIGN, "eb_build_ops:align: patch not aligned to parent cells.");
    const int mx = P.mx, my = P.my;
    const int Il = (b.ilo + 1) / 2, Ih = b.ihi / 2, Jl = (b.jlo + 1) / 2, Jh = b.jhi / 2;
    const int cmx = Ih - Il + 1, cmy = Jh - Jl + 1;
    const int ldc = cmx + 2;
    // ---- parent window: ring (with overhang into parent ghosts) + interior
    auto set_w = [&](int wi, int wj, int q, int li, int lj) {
      const PatchInfo& Q = C.pi[q];
      F.wmap[P.qcoff + wi + wj * ldc] = qidx(Q, li, lj);
      F.wmap_patch[P.qcoff + wi + wj * ldc] = q;
    };
    for (int wj = 0; wj < cmy + 2; ++wj) {
      for (int wi = 0; wi < cmx + 2; ++wi) {
        const int gi = Il - 1 + wi, gj = Jl - 1 + wj;
        const bool ringcell = wi == 0 || wj == 0 || wi == cmx + 1 || wj == cmy + 1;
        if (ringcell) {
          const int gic = std::min(std::max(gi, 1), C.n[0]);
          const int gjc = std::min(std::max(gj, 1), C.n[1]);
          const int q = OC.at(gic, gjc);
          if (q == 0)
            return fail(BSAM_E_RING, "eb_build_ops:ring: parent cell (%d,%d) not covered - proper nesting violated.", gi, gj);
          const Box& qb = C.box[q - 1];
          const int li = gi - qb.ilo + 1, lj = gj - qb.jlo + 1;
          if (li < 0 || li > qb.ihi - qb.ilo + 2 || lj < 0 || lj > qb.jhi - qb.jlo + 2)
            return fail(BSAM_E_OVERHANG, "eb_build_ops:overhang: ring cell overhang exceeds ghost ring.");
          set_w(wi, wj, q - 1, li, lj);
        } else {
          const int q = OC.at(gi, gj);
          if (q == 0)
            return fail(BSAM_E_NESTING,
                        "eb_build_ops:nesting: patch %d at level: parent window not fully covered - proper nesting violated.", p + 1);
          const Box& qb = C.box[q - 1];
          set_w(wi, wj, q - 1, gi - qb.ilo + 1, gj - qb.jlo + 1);
          cov[cidx(C.pi[q - 1], gi - qb.ilo, gj - qb.jlo)] = 1;
        }
      }
    }
    auto win = [&](int wi, int wj) { return F.wmap[P.qcoff + wi + wj * ldc]; };
    // ---- ghost classes and records (land / cornerClass / same-level)
    F.pi[p].gbase = (int)F.ghost.size();
    F.klass_off.push_back((int)F.klass.size());
    int first[4];
    bool anycf = false;
    for (int side = 0; side < 4; ++side) {
      first[side] = (int)F.ghost.size();
      const int len = side < 2 ? my : mx;
      for (int t = 0; t < len; ++t) {
        if (P.bc[side] != 0) {
          F.ghost.push_back(phys_rec(F, P, side, t));
          F.klass.push_back(1);
          continue;
        }
        int gi, gj;
        switch (side) {
          case 0: gi = b.ilo - 1; gj = b.jlo + t; break;
          case 1: gi = b.ihi + 1; gj = b.jlo + t; break;
          case 2: gi = b.ilo + t; gj = b.jlo - 1; break;
          default: gi = b.ilo + t; gj = b.jhi + 1; break;
        }
        const int s = OF.at(gi, gj);
        if (s > 0) {
          const Box& sb = F.box[s - 1];
          GhostRec r{};
          r.kind = G_COPY;
          r.s[0] = qidx(F.pi[s - 1], gi - sb.ilo + 1, gj - sb.jlo + 1);
          // the sibling's ghost slot on the far side of its boundary cell
          // r.s[0] is where this patch's boundary cell is copied to
          const int lds = F.pi[s - 1].mx + 2;
          r.gi = side == 0 ? r.s[0] + 1 : side == 1 ? r.s[0] - 1 : side == 2 ? r.s[0] + lds : r.s[0] - lds;
          const GhostRec tmp = cf_rec(P, side, t, qif, win);  // for dst only
          r.dst = tmp.dst;
          F.ghost.push_back(r);
          F.klass.push_back(2);
        } else {
          F.ghost.push_back(cf_rec(P, side, t, qif, win));
          F.klass.push_back(3);
        }
      }
    }
    const int cgi[4] = {b.ilo - 1, b.ihi + 1, b.ilo - 1, b.ihi + 1};
    const int cgj[4] = {b.jlo - 1, b.jlo - 1, b.jhi + 1, b.jhi + 1};
    const int cdst[4] = {qidx(P, 0, 0), qidx(P, mx + 1, 0), qidx(P, 0, my + 1), qidx(P, mx + 1, my + 1)};
    const int cs0[4] = {first[2], first[2] + mx - 1, first[3], first[3] + mx - 1};
    const int cs1[4] = {first[0], first[1], first[0] + my - 1, first[1] + my - 1};
    const int cs2[4] = {qidx(P, 1, 1), qidx(P, mx, 1), qidx(P, 1, my), qidx(P, mx, my)};
    // diagonal stencil sources of a coarse-fine corner (ebInterpCf.m:123-153)
    const int cq2[4] = {qidx(P, 2, 2), qidx(P, mx - 1, 2), qidx(P, 2, my - 1), qidx(P, mx - 1, my - 1)};
    const int cwi[4] = {0, cmx + 1, 0, cmx + 1};
    const int cwj[4] = {0, 0, cmy + 1, cmy + 1};
    for (int c = 0; c < 4; ++c) {
      GhostRec r{};
      r.dst = cdst[c];
      r.gi = -1;
      const int gi = cgi[c], gj = cgj[c];
      int klass;
      if (gi < 1 || gi > F.n[0] || gj < 1 || gj > F.n[1]) {
        klass = 1;
        r.kind = G_PCORNER; r.s[0] = cs0[c]; r.s[1] = cs1[c]; r.s[2] = cs2[c];
      } else {
        const int s = OF.at(gi, gj);
        if (s > 0) {
          klass = 2;
          const Box& sb = F.box[s - 1];
          r.kind = G_COPY;
          r.s[0] = qidx(F.pi[s - 1], gi - sb.ilo + 1, gj - sb.jlo + 1);
        } else {
          klass = 3;
          anycf = true;
          if (qif) {
            r.kind = G_QIF; r.s[0] = cq2[c]; r.s[1] = cs2[c]; r.s[2] = win(cwi[c], cwj[c]);
          } else {
            r.kind = G_LIF_CORNER; r.s[0] = cs2[c]; r.s[2] = win(cwi[c], cwj[c]);
          }
        }
      }
      F.ghost.push_back(r);
      F.klass.push_back((int8_t)klass);
    }
    const bool fint = P.bc[0] == 0 || P.bc[1] == 0 || P.bc[2] == 0 || P.bc[3] == 0;
    F.pi[p].cf = (fint || anycf) ? 1 : 0;
    // ---- ring gathers (ebBuildOps.m:120-122)
    for (int wj = 0; wj < cmy + 2; ++wj)
      for (int wi = 0; wi < cmx + 2; ++wi) {
        if (!(wi == 0 || wj == 0 || wi == cmx + 1 || wj == cmy + 1)) continue;
        RingRec rr{P.qcoff + wi + wj * ldc, win(wi, wj)};
        F.ring_all.push_back(rr);
        if (F.pi[p].cf) F.ring_cf.push_back(rr);
      }
    // ---- footprint: parent cell, F index, parent face coefficients
    //      (ebBuildOps.m:144-156, ebGatherParentCoeffs.m:20-46; the last
    //      parent patch in order wins on shared faces, as in the reference)
    std::vector<int> pdew((size_t)(cmx + 1) * cmy, -1), pdns((size_t)cmx * (cmy + 1), -1);
    for (int q = 0; q < C.np; ++q) {
      const Box& qb = C.box[q];
      const PatchInfo& Q = C.pi[q];
      const int mqx = qb.ihi - qb.ilo + 1;
      for (int gj = std::max(Jl, qb.jlo); gj <= std::min(Jh, qb.jhi); ++gj)
        for (int f = std::max(Il, qb.ilo); f <= std::min(Ih + 1, qb.ihi + 1); ++f)
          pdew[(f - Il) + (size_t)(gj - Jl) * (cmx + 1)] = Q.xoff + (f - qb.ilo) + (gj - qb.jlo) * (mqx + 1);
      for (int g = std::max(Jl, qb.jlo); g <= std::min(Jh + 1, qb.jhi + 1); ++g)
        for (int gi = std::max(Il, qb.ilo); gi <= std::min(Ih, qb.ihi); ++gi)
          pdns[(gi - Il) + (size_t)(g - Jl) * cmx] = Q.yoff + (gi - qb.ilo) + (g - qb.jlo) * mqx;
    }
    for (int v : pdew)
      if (v < 0) return fail(BSAM_E_COVER, "eb_gather_parent_coeffs:cover: footprint of patch %d not covered by parent level.", p + 1);
    for (int v : pdns)
      if (v < 0) return fail(BSAM_E_COVER, "eb_gather_parent_coeffs:cover: footprint of patch %d not covered by parent level.", p + 1);
    for (int cj = 0; cj < cmy; ++cj)
      for (int ci = 0; ci < cmx; ++ci) {
        FootRec fr{};
        fr.p = p; fr.ci = ci; fr.cj = cj;
        fr.qoff = P.qoff; fr.coff = P.coff; fr.xoff = P.xoff; fr.yoff = P.yoff; fr.qcoff = P.qcoff;
        fr.mx = P.mx; fr.my = P.my;
        fr.pq = win(ci + 1, cj + 1);
        const int q = OC.at(Il + ci, Jl + cj) - 1;
        const Box& qb = C.box[q];
        fr.pc = cidx(C.pi[q], Il + ci - qb.ilo, Jl + cj - qb.jlo);
        fr.dew_w = pdew[ci + (size_t)cj * (cmx + 1)];
        fr.dew_e = pdew[ci + 1 + (size_t)cj * (cmx + 1)];
        fr.dns_s = pdns[ci + (size_t)cj * cmx];
        fr.dns_n = pdns[ci + (size_t)(cj + 1) * cmx];
        F.foot.push_back(fr);
      }
    // ---- reflux entries (ebBuildOps.m:158-166, ebFaceCorrections.m:29-55)
    const int ko = F.klass_off.back();
    for (int side = 0; side < 4; ++side) {
      const int clen = side < 2 ? cmy : cmx;
      const int koff = ko + (side == 0 ? 0 : side == 1 ? my : side == 2 ? 2 * my : 2 * my + mx);
      for (int j0 = 0; j0 < clen; ++j0) {
        if (F.klass[koff + 2 * j0 + 1] != 3) continue;
        int gi, gj;
        switch (side) {
          case 0: gi = Il - 1; gj = Jl + j0; break;
          case 1: gi = Ih + 1; gj = Jl + j0; break;
          case 2: gi = Il + j0; gj = Jl - 1; break;
          default: gi = Il + j0; gj = Jh + 1; break;
        }
        const int q = OC.at(gi, gj);
        if (q == 0) return fail(BSAM_E_CORR, "eb_build_ops:corr: mass-correction target cell uncovered.");
        const Box& qb = C.box[q - 1];
        CorrRec e{};
        e.face = side;
        const int a = 2 * j0, bb = 2 * j0 + 1;  // fine rows/cols of the two sub-faces (0-based)
        switch (side) {
          case 0:
            e.d1 = P.xoff + 0 + a * (mx + 1); e.d2 = P.xoff + 0 + bb * (mx + 1);
            e.in1 = qidx(P, 1, a + 1); e.g1 = qidx(P, 0, a + 1); e.in2 = qidx(P, 1, bb + 1); e.g2 = qidx(P, 0, bb + 1);
            e.ring = P.qcoff + 0 + (j0 + 1) * ldc; e.qcin = P.qcoff + 1 + (j0 + 1) * ldc;
            e.cfidx = pdew[0 + (size_t)j0 * (cmx + 1)];
            e.cf = C.DeW[e.cfidx];
            break;
          case 1:
            e.d1 = P.xoff + mx + a * (mx + 1); e.d2 = P.xoff + mx + bb * (mx + 1);
            e.in1 = qidx(P, mx, a + 1); e.g1 = qidx(P, mx + 1, a + 1); e.in2 = qidx(P, mx, bb + 1); e.g2 = qidx(P, mx + 1, bb + 1);
            e.ring = P.qcoff + (cmx + 1) + (j0 + 1) * ldc; e.qcin = P.qcoff + cmx + (j0 + 1) * ldc;
            e.cfidx = pdew[cmx + (size_t)j0 * (cmx + 1)];
            e.cf = C.DeW[e.cfidx];
            break;
          case 2:
            e.d1 = P.yoff + a + 0 * mx; e.d2 = P.yoff + bb + 0 * mx;
            e.in1 = qidx(P, a + 1, 1); e.g1 = qidx(P, a + 1, 0); e.in2 = qidx(P, bb + 1, 1); e.g2 = qidx(P, bb + 1, 0);
            e.ring = P.qcoff + (j0 + 1) + 0 * ldc; e.qcin = P.qcoff + (j0 + 1) + 1 * ldc;
            e.cfidx = pdns[j0 + (size_t)0 * cmx];
            e.cf = C.DnS[e.cfidx];
            break;
          default:
            e.d1 = P.yoff + a + my * mx; e.d2 = P.yoff + bb + my * mx;
            e.in1 = qidx(P, a + 1, my); e.g1 = qidx(P, a + 1, my + 1); e.in2 = qidx(P, bb + 1, my); e.g2 = qidx(P, bb + 1, my + 1);
            e.ring = P.qcoff + (j0 + 1) + (cmy + 1) * ldc; e.qcin = P.qcoff + (j0 + 1) + cmy * ldc;
            e.cfidx = pdns[j0 + (size_t)cmy * cmx];
            e.cf = C.DnS[e.cfidx];
            break;
        }
        F.corr.push_back(e);
        F.corr_target.push_back(cidx(C.pi[q - 1], gi - qb.ilo, gj - qb.jlo));
      }
    }
  }
  // ---- targets on the parent level (sequential order per target cell)
  C.tgt_cell.clear(); C.tgt_ptr.clear(); C.tgt_ent.clear();
  C.tgt_of_cell.assign(C.ncell, -1);
  std::vector<std::vector<int>> rows;
  for (int e = 0; e < (int)F.corr.size(); ++e) {
    const int cell = F.corr_target[e];
    int row = C.tgt_of_cell[cell];
    if (row < 0) {
      row = (int)rows.size();
      C.tgt_of_cell[cell] = row;
      C.tgt_cell.push_back(cell);
      rows.emplace_back();
    }
    rows[row].push_back(e);
  }
  C.tgt_ptr.push_back(0);
  for (auto& r : rows) {
    for (int e : r) C.tgt_ent.push_back(e);
    C.tgt_ptr.push_back((int)C.tgt_ent.size());
  }
  C.vis.assign(C.ncell, 1);
  for (int c = 0; c < C.ncell; ++c)
    if (cov[c]) C.vis[c] = 0;
  return 0;
}

// ------------------------------------------------------------------ coarsest
// Dense ghost-eliminated level-1 matrix (ebCoarsest.m:14-58, duplicates summed
// in triplet order) and its explicit inverse by LU with partial pivoting.
int coarsest_assemble(const HLevel& L, bool singular, std::vector<double>& Ainv,
                      std::vector<double>& bcRhs, int& M) {
  const int n1 = L.n[0], n2 = L.n[1], N = n1 * n2;
  const double hx = L.h[0], hy = L.h[1], hx2 = hx * hx, hy2 = hy * hy;
  const PatchInfo& P = L.pi[0];
  M = singular ? N + 1 : N;
  std::vector<double> A((size_t)M * M, 0.0);  // row-major during elimination
  auto at = [&](int r, int c) -> double& { return A[(size_t)r * M + c]; };
  bcRhs.assign(N, 0.0);
  auto idx = [&](int i, int j) { return i + j * n1; };
  const double* DeW = &L.DeW[P.xoff];
  const double* DnS = &L.DnS[P.yoff];
  const double* CC = &L.CC[P.coff];
  for (int c = 0; c < N; ++c) at(c, c) = at(c, c) + CC[c];
  // internal x-faces: (iL,iL,+c) all, (iR,iR,+c) all, (iL,iR,-c) all, (iR,iL,-c) all
  {
    std::vector<double> cv;
    std::vector<int> iL, iR;
    for (int j = 0; j < n2; ++j)
      for (int f = 1; f < n1; ++f) {
        cv.push_back(DeW[f + j * (n1 + 1)] / hx2);
        iL.push_back(idx(f - 1, j));
        iR.push_back(idx(f, j));
      }
    for (size_t t = 0; t < cv.size(); ++t) at(iL[t], iL[t]) = at(iL[t], iL[t]) + cv[t];
    for (size_t t = 0; t < cv.size(); ++t) at(iR[t], iR[t]) = at(iR[t], iR[t]) + cv[t];
    for (size_t t = 0; t < cv.size(); ++t) at(iL[t], iR[t]) = at(iL[t], iR[t]) + (-cv[t]);
    for (size_t t = 0; t < cv.size(); ++t) at(iR[t], iL[t]) = at(iR[t], iL[t]) + (-cv[t]);
  }
  {
    std::vector<double> cv;
    std::vector<int> iL, iR;
    for (int f = 1; f < n2; ++f)
      for (int i = 0; i < n1; ++i) {
        cv.push_back(DnS[i + f * n1] / hy2);
        iL.push_back(idx(i, f - 1));
        iR.push_back(idx(i, f));
      }
    for (size_t t = 0; t < cv.size(); ++t) at(iL[t], iL[t]) = at(iL[t], iL[t]) + cv[t];
    for (size_t t = 0; t < cv.size(); ++t) at(iR[t], iR[t]) = at(iR[t], iR[t]) + cv[t];
    for (size_t t = 0; t < cv.size(); ++t) at(iL[t], iR[t]) = at(iL[t], iR[t]) + (-cv[t]);
    for (size_t t = 0; t < cv.size(); ++t) at(iR[t], iL[t]) = at(iR[t], iL[t]) + (-cv[t]);
  }
  const double* G = &L.G[P.goff];
  for (int side = 0; side < 4; ++side) {
    const int code = P.bc[side];
    const int len = side < 2 ? n2 : n1;
    const double h = side < 2 ? hx : hy, h2 = side < 2 ? hx2 : hy2;
    const double* g = G + (side == 0 ? 0 : side == 1 ? n2 : side == 2 ? 2 * n2 : 2 * n2 + n1);
    std::vector<double> d(len);
    std::vector<int> i1(len), i2(len);
    for (int t = 0; t < len; ++t) {
      switch (side) {
        case 0: d[t] = DeW[0 + t * (n1 + 1)]; i1[t] = idx(0, t); i2[t] = idx(1, t); break;
        case 1: d[t] = DeW[n1 + t * (n1 + 1)]; i1[t] = idx(n1 - 1, t); i2[t] = idx(n1 - 2, t); break;
        case 2: d[t] = DnS[t + 0 * n1]; i1[t] = idx(t, 0); i2[t] = idx(t, 1); break;
        default: d[t] = DnS[t + n2 * n1]; i1[t] = idx(t, n2 - 1); i2[t] = idx(t, n2 - 2); break;
      }
    }
    if (code == 1) {
      for (int t = 0; t < len; ++t) at(i1[t], i1[t]) = at(i1[t], i1[t]) + 3 * d[t] / h2;
      for (int t = 0; t < len; ++t) at(i1[t], i2[t]) = at(i1[t], i2[t]) + (-d[t] / (3 * h2));
      for (int t = 0; t < len; ++t) bcRhs[i1[t]] = bcRhs[i1[t]] + (8.0 / 3.0) * d[t] * g[t] / h2;
    } else if (code == 2) {
      for (int t = 0; t < len; ++t) bcRhs[i1[t]] = bcRhs[i1[t]] + d[t] * g[t] / h;
    }
  }
  if (singular) {
    for (int c = 0; c < N; ++c) {
      at(c, N) = 1.0;
      at(N, c) = 1.0;
    }
  }
  Ainv.assign((size_t)M * M, 0.0);
  if (!bsam_lu_inverse(A.data(), M, Ainv.data()))
    return fail(BSAM_E_SINGULAR, "bsamgpu:coarsest: level-1 matrix is singular");
  return 0;
}

// ------------------------------------------------------------------ validation
// Every index a kernel dereferences is checked against its buffer before any
// upload, so a table bug is reported as an error instead of reaching the GPU.
int validate_level(const HLevel& L, const HLevel* C, const HLevel* Fn) {
  auto bad = [&](const char* what, long long v, long long n) {
    return fail(BSAM_E_ARG, "bsamgpu:internal: table %s index %lld outside [0,%lld)", what, v, n);
  };
  auto in = [](long long v, long long n) { return v >= 0 && v < n; };
  for (const TileRec& t : L.tiles) {
    if (!in(t.p, L.np)) return bad("tile patch", t.p, L.np);
    if (!in(t.start, (long long)L.pi[t.p].mx * L.pi[t.p].my))
      return bad("tile cell", t.start, (long long)L.pi[t.p].mx * L.pi[t.p].my);
    if (std::memcmp(&t.pi, &L.pi[t.p], sizeof(PatchInfo)) != 0)
      return fail(BSAM_E_ARG, "bsamgpu:internal: stale patch record in a tile");
  }
  for (const STile& t : L.stiles) {
    if (!in(t.p, L.np)) return bad("smoother tile patch", t.p, L.np);
    const PatchInfo& P = L.pi[t.p];
    if (t.i0 < 0 || t.i1 >= P.mx || t.j0 < 0 || t.j1 >= P.my || t.i0 > t.i1 || t.j0 > t.j1 ||
        (P.mx + 2) * (P.my + 2) > L.stile_lds)
      return fail(BSAM_E_ARG, "bsamgpu:internal: smoother tile outside its patch");
  }
  for (const PatchInfo& P : L.pi) {
    if (P.qoff + (P.mx + 2) * (P.my + 2) > L.nq || P.coff + P.mx * P.my > L.ncell ||
        P.xoff + (P.mx + 1) * P.my > L.nx || P.yoff + P.mx * (P.my + 1) > L.ny || P.goff + 2 * (P.mx + P.my) > L.ng)
      return fail(BSAM_E_ARG, "bsamgpu:internal: patch offsets exceed level buffers");
    if (C && P.qcoff + (P.mx / 2 + 2) * (P.my / 2 + 2) > L.nqc)
      return fail(BSAM_E_ARG, "bsamgpu:internal: QC offsets exceed level buffer");
  }
  const long long nrec = (long long)L.ghost.size();
  const long long pq = C ? C->nq : 0;
  for (const GhostRec& r : L.ghost) {
    if (!in(r.dst, L.nq)) return bad("ghost dst", r.dst, L.nq);
    switch (r.kind) {
      case G_COPY: case G_NEU: if (!in(r.s[0], L.nq)) return bad("ghost src", r.s[0], L.nq); break;
      case G_DIR: if (!in(r.s[0], L.nq) || !in(r.s[1], L.nq)) return bad("ghost src", r.s[0], L.nq); break;
      case G_QIF: case G_LIF:
        if (!in(r.s[0], L.nq) || !in(r.s[1], L.nq)) return bad("ghost src", r.s[1], L.nq);
        if (!in(r.s[2], pq)) return bad("ghost parent src", r.s[2], pq);
        break;
      case G_LIF_CORNER:
        if (!in(r.s[0], L.nq)) return bad("ghost src", r.s[0], L.nq);
        if (!in(r.s[2], pq)) return bad("ghost parent src", r.s[2], pq);
        break;
      case G_QIF_END:
        if (!in(r.s[0], L.nq) || !in(r.s[1], L.nq)) return bad("ghost src", r.s[0], L.nq);
        for (int i = 2; i < 5; ++i) if (!in(r.s[i], pq)) return bad("ghost parent src", r.s[i], pq);
        break;
      case G_PCORNER:
        if (!in(r.s[0], nrec) || !in(r.s[1], nrec)) return bad("corner record", r.s[0], nrec);
        if (L.ghost[r.s[0]].kind == G_PCORNER || L.ghost[r.s[1]].kind == G_PCORNER)
          return fail(BSAM_E_ARG, "bsamgpu:internal: corner refers to a corner");
        if (!in(r.s[2], L.nq)) return bad("corner src", r.s[2], L.nq);
        break;
      default: return fail(BSAM_E_ARG, "bsamgpu:internal: ghost kind %d", r.kind);
    }
  }
  for (const auto* rl : {&L.ring_all, &L.ring_cf})
    for (const RingRec& r : *rl) {
      if (!in(r.dst, L.nqc)) return bad("ring dst", r.dst, L.nqc);
      if (!in(r.src, pq)) return bad("ring src", r.src, pq);
    }
  for (int v : L.wmap) if (!in(v, pq)) return bad("window", v, pq);
  for (const FootRec& f : L.foot) {
    if (!in(f.p, L.np)) return bad("foot patch", f.p, L.np);
    if (!in(f.ci, L.pi[f.p].mx / 2) || !in(f.cj, L.pi[f.p].my / 2)) return bad("foot cell", f.ci, L.pi[f.p].mx / 2);
    if (!in(f.pq, pq)) return bad("foot parent Q", f.pq, pq);
    if (!in(f.pc, C->ncell)) return bad("foot parent cell", f.pc, C->ncell);
    if (!in(f.dew_w, C->nx) || !in(f.dew_e, C->nx)) return bad("foot DeW", f.dew_w, C->nx);
    if (!in(f.dns_s, C->ny) || !in(f.dns_n, C->ny)) return bad("foot DnS", f.dns_s, C->ny);
  }
  for (const CorrRec& e : L.corr) {
    const long long nd = e.face < 2 ? L.nx : L.ny;
    if (!in(e.d1, nd) || !in(e.d2, nd)) return bad("corr D", e.d1, nd);
    for (int v : {e.in1, e.g1, e.in2, e.g2}) if (!in(v, L.nq)) return bad("corr Q", v, L.nq);
    if (!in(e.ring, L.nqc) || !in(e.qcin, L.nqc)) return bad("corr QC", e.ring, L.nqc);
  }
  const long long nfc = Fn ? (long long)Fn->corr.size() : 0;
  if (!L.tgt_cell.empty() && (long long)L.tgt_ptr.size() != (long long)L.tgt_cell.size() + 1)
    return fail(BSAM_E_ARG, "bsamgpu:internal: target row pointer size");
  for (int c : L.tgt_cell) if (!in(c, L.ncell)) return bad("target cell", c, L.ncell);
  for (int e : L.tgt_ent) if (!in(e, nfc)) return bad("target entry", e, nfc);
  for (size_t i = 0; i + 1 < L.tgt_ptr.size(); ++i)
    if (L.tgt_ptr[i] > L.tgt_ptr[i + 1] || L.tgt_ptr[i + 1] > (int)L.tgt_ent.size())
      return fail(BSAM_E_ARG, "bsamgpu:internal: target rows not monotone");
  for (int v : L.tgt_of_cell) if (v >= (long long)L.tgt_cell.size()) return bad("target row", v, (long long)L.tgt_cell.size());
  if ((long long)L.vis.size() != L.ncell || (long long)L.tgt_of_cell.size() != L.ncell)
    return fail(BSAM_E_ARG, "bsamgpu:internal: per-cell map sizes");
  return 0;
}

// ------------------------------------------------------------------ upload
template <typename T, typename U>
int upload(bsam_hier* h, T** dst, const std::vector<U>& src) {
  int rc = h->mem.alloc(dst, src.size());
  if (rc) return rc;
  if (!src.empty()) HIPCHK(hipMemcpy(*dst, src.data(), src.size() * sizeof(U), hipMemcpyHostToDevice));
  return 0;
}

template <typename T>
int upload_cast(bsam_hier* h, T** dst, const std::vector<double>& src) {
  std::vector<T> tmp(src.begin(), src.end());
  return upload(h, dst, tmp);
}

template <typename T>
int make_dev_level(bsam_hier* h, HLevel& L, LevelDev<T>& D, bool tables_from64, const LevelDev<double>* d64) {
  D.n[0] = L.n[0]; D.n[1] = L.n[1]; D.h[0] = L.h[0]; D.h[1] = L.h[1];
  D.ih2[0] = 1.0 / (L.h[0] * L.h[0]);
  D.ih2[1] = 1.0 / (L.h[1] * L.h[1]);
  D.pow2 = is_pow2(L.h[0]) && is_pow2(L.h[1]) ? 1 : 0;
  D.even_mx = 1;
  for (const PatchInfo& P : L.pi)
    if (P.mx % 2) D.even_mx = 0;
  D.arith = tables_from64 ? d64->arith : (face_records_arith(L) ? 1 : 0);
  D.np = L.np; D.ncell = L.ncell; D.nq = L.nq; D.nx = L.nx; D.ny = L.ny; D.nqc = L.nqc; D.ng = L.ng;
  D.ntile = (int)L.tiles.size(); D.nghost = (int)L.ghost.size();
  D.nring_all = (int)L.ring_all.size(); D.nring_cf = (int)L.ring_cf.size();
  D.nfoot = (int)L.foot.size(); D.ncorr = (int)L.corr.size(); D.ntgt = (int)L.tgt_cell.size();
  D.nstile = (int)L.stiles.size(); D.stile_lds = L.stile_lds;
  int rc;
  if (tables_from64) {  // share the index tables of the FP64 level
    D.P = d64->P; D.tiles = d64->tiles; D.ghost = d64->ghost; D.ring_all = d64->ring_all;
    D.ring_cf = d64->ring_cf; D.foot = d64->foot; D.corr = d64->corr; D.tgt_cell = d64->tgt_cell;
    D.tgt_ptr = d64->tgt_ptr; D.tgt_ent = d64->tgt_ent; D.tgt_of_cell = d64->tgt_of_cell;
    D.wmap = d64->wmap; D.vis = d64->vis; D.stiles = d64->stiles;
    D.wgr = d64->wgr; D.ring_rec = d64->ring_rec; D.pcov = d64->pcov;
  } else {
    if ((rc = upload(h, &D.P, L.pi))) return rc;
    if ((rc = upload(h, &D.tiles, L.tiles))) return rc;
    if ((rc = upload(h, &D.stiles, L.stiles))) return rc;
    if ((rc = upload(h, &D.ghost, L.ghost))) return rc;
    if ((rc = upload(h, &D.ring_all, L.ring_all))) return rc;
    if ((rc = upload(h, &D.ring_cf, L.ring_cf))) return rc;
    if ((rc = upload(h, &D.foot, L.foot))) return rc;
    if ((rc = upload(h, &D.corr, L.corr))) return rc;
    if ((rc = upload(h, &D.tgt_cell, L.tgt_cell))) return rc;
    std::vector<int> ptr = L.tgt_ptr.empty() ? std::vector<int>{0} : L.tgt_ptr;
    if ((rc = upload(h, &D.tgt_ptr, ptr))) return rc;
    if ((rc = upload(h, &D.tgt_ent, L.tgt_ent))) return rc;
    if ((rc = upload(h, &D.tgt_of_cell, L.tgt_of_cell))) return rc;
    if ((rc = upload(h, &D.wmap, L.wmap))) return rc;
    if ((rc = upload(h, &D.vis, L.vis))) return rc;
    if ((rc = upload(h, &D.wgr, L.wgr))) return rc;
    if ((rc = upload(h, &D.ring_rec, L.ring_rec))) return rc;
    // the parent's covered cells -> this level's 2 x 2 blocks (fused filldown steps)
    D.pcov = nullptr;
    const long long kl = &L - h->H.data();
    if (kl > 0 && !L.foot.empty()) {
      const HLevel& C = h->H[(size_t)kl - 1];
      std::vector<int2> cov((size_t)C.nq, make_int2(-1, 0));
      for (const FootRec& fr : L.foot) {
        const int ldq = fr.mx + 2;
        if (fr.pq < 0 || fr.pq >= C.nq) return fail(BSAM_E_ARG, "bsamgpu:internal: footprint parent index out of range");
        cov[(size_t)fr.pq] = make_int2(fr.qoff + (2 * fr.ci + 1) + (2 * fr.cj + 1) * ldq, ldq);
      }
      if ((rc = upload(h, &D.pcov, cov))) return rc;
    }
  }
  D.pgh_ok = L.pgh_ok;
  if ((rc = upload_cast(h, &D.Q, L.Q))) return rc;
  if ((rc = upload_cast(h, &D.FT, L.FT))) return rc;
  if ((rc = upload_cast(h, &D.F, L.FT))) return rc;
  if ((rc = upload_cast(h, &D.CC, L.CC))) return rc;
  if ((rc = upload_cast(h, &D.ID, L.ID))) return rc;
  if ((rc = upload_cast(h, &D.DeW, L.DeW))) return rc;
  if ((rc = upload_cast(h, &D.DnS, L.DnS))) return rc;
  if ((rc = upload_cast(h, &D.G, L.G))) return rc;
  std::vector<double> zqc(L.nqc, 0.0);
  if ((rc = upload_cast(h, &D.QC, zqc))) return rc;
  // fused sweep tiles (levels with the standard face layout, even widths);
  // the FP32 copy shares them, each precision has its own second Q
  D.nsweep = 0; D.sweep = nullptr; D.Qs = nullptr;
  // streaming sweeps beat colour passes from about a million cells up
  D.stream = L.np == 1 && L.pi[0].mx % 2 == 0 && L.pi[0].bc[0] && L.pi[0].bc[1] && L.pi[0].bc[2] &&
             L.pi[0].bc[3] && (long long)L.pi[0].mx * L.pi[0].my >= kStreamMinCells;
  if (D.arith && D.even_mx && L.np > 0) {
    if (tables_from64) {
      D.nsweep = d64->nsweep; D.sweep = d64->sweep;
    } else {
      // a level takes fused sweeps when its tiles fit the device in one round
      // (the sweep kernel holds three 256-lane workgroups per CU); a larger
      // level runs colour passes, whose launches are all resident at once
      std::vector<SweepTile> st;
      const long long round = (long long)kSweepWGsPerCU * h->ncu;
      if (build_sweep_tiles(L, st) && ((long long)st.size() <= round || (long long)st.size() >= 16 * round)) {
        if ((rc = upload(h, &D.sweep, st))) return rc;
        D.nsweep = (int)st.size();
      }
    }
  }
  // whole relax calls per launch (region tiles) on levels with the standard
  // face layout that do not stream; the FP32 copy shares the tables
  D.nregion = 0; D.rdepth = 0; D.rslots = 0;
  D.region = nullptr; D.rpatch = nullptr; D.rcell = nullptr; D.rghost = nullptr; D.rrs = nullptr;
  D.rghmax = 0; D.rrsmax = 0;
  D.rgw = nullptr; D.rwin = nullptr; D.rcorr = 0; D.rwinmax = 0;
  if (D.arith && !D.stream && L.np > 0) {
    if (tables_from64) {
      D.nregion = d64->nregion; D.rdepth = d64->rdepth; D.rslots = d64->rslots;
      D.region = d64->region; D.rpatch = d64->rpatch; D.rcell = d64->rcell;
      D.rghost = d64->rghost; D.rrs = d64->rrs; D.rghmax = d64->rghmax; D.rrsmax = d64->rrsmax;
      D.rgw = d64->rgw; D.rwin = d64->rwin; D.rcorr = d64->rcorr; D.rwinmax = d64->rwinmax;
    } else {
      const int rdepth = 2 * std::max(3, std::max(h->opts.npre, h->opts.npost));
      RegionBuild rb;
      if (rdepth <= kRegionMaxDepth && build_region_tiles(L, rdepth, rb)) {
        if ((rc = check_region_tables(L, rb))) return rc;
        // the correction folded into the post-smoothing launch (levels with a parent)
        const long long kl = &L - h->H.data();
        // the post-smoothing launch's LDS with the window: canvas, ghost
        // records (at most 80 bytes each), ring staging, window values
        auto a16 = [](long long b) { return (b + 15) / 16 * 16; };
        const long long ntile = (long long)rb.tiles.size();
        const long long rring = std::min<long long>(kRegionThreads, ((long long)L.ring_cf.size() + ntile - 1) / ntile);
        auto lds_with = [&](int win) {
          return a16(8LL * rb.max_slots) + a16(80LL * rb.max_gh) + a16(4 * rring) + a16(8 * rring) + 2 * a16(8LL * win);
        };
        if (kl > 0 && build_region_windows(L, h->H[(size_t)kl - 1].nq, rb) &&
            lds_with(rb.max_win) + 8192 <= kStageLdsLimit) {
          if ((rc = upload(h, &D.rwin, rb.win))) return rc;
          if ((rc = upload(h, &D.rgw, rb.gw))) return rc;
          D.rcorr = 1;
          D.rwinmax = rb.max_win;
        }
        if ((rc = upload(h, &D.region, rb.tiles))) return rc;
        if ((rc = upload(h, &D.rpatch, rb.patches))) return rc;
        if ((rc = upload(h, &D.rcell, rb.cells))) return rc;
        if ((rc = upload(h, &D.rghost, rb.ghosts))) return rc;
        L.rghost = rb.ghosts;
        if ((rc = upload(h, &D.rrs, rb.rs))) return rc;
        D.nregion = (int)rb.tiles.size();
        D.rdepth = rdepth;
        D.rslots = rb.max_slots;
        D.rghmax = rb.max_gh;
        D.rrsmax = rb.max_rs;
      }
    }
  }
  if (D.nsweep || D.stream || D.nregion) {  // the sweeps' and region relax's second copy of Q
    std::vector<double> zq(L.nq, 0.0);
    if ((rc = upload_cast(h, &D.Qs, zq))) return rc;
  }
  std::vector<FaceAux<T>> aux(L.ghost.size());
  for (size_t t = 0; t < L.ghost.size(); ++t) {
    aux[t].cf = T(0);
    aux[t].kind = L.ghost[t].kind;
    aux[t].push = (L.ghost[t].kind == G_COPY && D.arith) ? L.ghost[t].gi : -1;
  }
  if ((rc = upload(h, &D.aux, aux))) return rc;
  return 0;
}

// ------------------------------------------------------------------ sequences
// Coarse-cycle cut: levels 1..kc (each at most `cells` cells) run their whole
// V-cycle inside one workgroup; kc < 2 disables it.
int coarse_cut(const bsam_hier* h, int cells) {
  if (cells < 0) return 1;
  if (cells == 0) cells = 1100;
  int kc = 1;
  while (kc < h->nlev && h->H[kc].ncell <= cells) ++kc;
  return kc;
}

// Filldown cut: levels 1..kb have at most kFilldownSmallCells cells each;
// their restrict + fill steps (k = kb..2) run in one workgroup.
constexpr int kFilldownSmallCells = 16384;
int filldown_cut(const bsam_hier* h) {
  int kb = 1;
  while (kb < h->nlev && h->H[kb].ncell <= kFilldownSmallCells) ++kb;
  return kb;
}

// The reference's call sequence as level-wide launches.  Three fusions keep
// the arithmetic and the final state identical to the reference order:
//  * inside one relax call, passes after the first evaluate internal-face
//    ghosts from the ghost records (what the fill after the previous pass
//    would have written) and the level is filled once after the last pass;
//  * ring gather + FAS load + reflux run as one launch (disjoint items; the
//    load reads ring values from the parent, which is what the gather wrote);
//  * after the correction the fill is folded into the first post-smoothing
//    pass (nothing reads the level's ghosts in between).
template <typename T>
struct Seq {
  bsam_hier* h;
  std::vector<LevelDev<T>>& L;
  const LevelDev<T>* LVd;  // device array of the same descriptors
  int homog;
  hipStream_t s;
  // the level descriptors as the buffers stand: in the fused cycle a level
  // that smooths with region launches swaps Q and Qs at every relax call
  // (pre-smoothing into Qs, post-smoothing back into Q)
  std::vector<LevelDev<T>> cur;
  // levels whose post-smoothing launch applied the correction and left out
  // the fill's QC ring refresh (other workgroups of that launch read the old
  // ring for their corrections); flush_rings refreshes them
  std::vector<char> ringp;
  // the top level's parent parts are current (the previous cycle's residual
  // refreshed them after the filldown, or the solve's first fill stored them)
  bool top_cf_current = false;
  // the FAS loads' reflux sums start from FTrue (a loop that resets F once,
  // before its first cycle: see reflux_item)
  int reflux_ft = 0;
  void flush_ring(int k) {
    if (k >= 2 && k <= (int)ringp.size() && ringp[k - 1]) {
      launch_ring_cf<T>(lev(k), lev(k - 1), s);
      ringp[k - 1] = 0;
    }
  }
  void flush_rings() {
    for (int k = 2; k <= (int)ringp.size(); ++k) flush_ring(k);
  }
  const LevelDev<T>& lev(int k) {
    if (cur.empty()) cur = L;
    return cur[k - 1];
  }
  const LevelDev<T>* par(int k) {
    if (cur.empty()) cur = L;
    return k > 1 ? &cur[k - 2] : nullptr;
  }
  void fill(int k) { launch_fill<T>(lev(k), par(k), homog, s); }
  // ebRelaxLevel.m:28-38; first_inline: 1 = the first pass evaluates ghosts
  // from the records (the preceding fill was skipped), 2 = the preceding
  // correction prepared the level for assembled ghosts (correct(k, true));
  // cf_current: the level was filled after its parent last changed, so the
  // parent parts that fill stored are those of the current parent
  // the whole relax call runs as one region launch (plus the fill that
  // follows every relax call)
  bool region(int k, int nsweeps) const {
    const LevelDev<T>& Lk = L[k - 1];
    return nsweeps > 0 && Lk.nregion > 0 && 2 * nsweeps <= Lk.rdepth && h->opts.smoother == 0;
  }
  // both relax calls of level k in the fused cycle take region launches that
  // write the level's whole new state into the other buffer
  bool swaps(int k) const {
    return h->opts.npre > 0 && h->opts.npost > 0 && region(k, h->opts.npre) && region(k, h->opts.npost);
  }
  // defer_fill: a region relax leaves its fill to the caller's fused launch.
  // swap: 1 = a region launch writes the own cells and every ghost into Qs
  // with the QC ring refresh (the fill), and the level's Q and Qs swap; 2 =
  // the same with the restriction into the parent (and rdup) in place of the
  // ring refresh, which the FAS load's gather redoes
  // corr: the launch applies the pending coarse-grid correction of level k
  // first (swap 1 only, levels with L.rcorr)
  // fload: k_relax_region's fused FAS load into the parent (a level of one
  // tile smoothing into its other buffer with the restriction)
  void relax(int k, int nsweeps, int first_inline = 0, bool cf_current = false, bool defer_fill = false,
             int swap = 0, T* rdup = nullptr, bool corr = false, int fload = 0) {
    const bool qs_current = qs_ready == k;
    qs_ready = 0;
    if (region(k, nsweeps)) {
      // pass 1 reads the frame when the reference's first pass does (the
      // preceding fill ran), else assembles its ghosts like every later pass.
      // The parent parts in aux are current after a fill that followed the
      // parent's last change or a preparing correction; else refresh them
      if (first_inline == 1 || (first_inline == 0 && !cf_current)) launch_refresh_cf<T>(lev(k), par(k), s);
      if (swap) {
        const LevelDev<T>& Lk = lev(k);
        T* rq = swap == 2 ? cur[k - 2].Q : nullptr;
        const bool cr = corr && swap == 1;
        launch_relax_region<T>(Lk, par(k), nsweeps, first_inline == 0 ? 1 : 0, (T)h->opts.omega, homog, s, Lk.Qs, 1,
                               rq, swap == 2 ? rdup : nullptr, swap == 1 && !cr ? 1 : 0, cr ? 1 : 0,
                               fload ? LVd + (k - 2) : nullptr, fload);
        if (cr) {
          if ((int)ringp.size() < h->nlev) ringp.assign((size_t)h->nlev, 0);
          ringp[(size_t)k - 1] = 1;
        }
        std::swap(cur[k - 1].Q, cur[k - 1].Qs);
        return;
      }
      launch_relax_region<T>(lev(k), par(k), nsweeps, first_inline == 0 ? 1 : 0, (T)h->opts.omega, homog, s);
      if (!defer_fill) launch_fill_copy<T>(lev(k), par(k), homog, lev(k).Qs, s);
      return;
    }
    if (nsweeps > 0 && h->opts.smoother == 1 && tiled_fits<T>(lev(k))) {
      // tiled smoother: pass 1 reads the materialised ghost frame
      if (first_inline) fill(k);
      launch_relax_tiled<T>(lev(k), par(k), nsweeps, (T)h->opts.omega, homog, s);
      fill(k);
      return;
    }
    // Pass 1 reads the materialised frame or, after a correction, the
    // records.  On levels with the standard face layout every later pass
    // assembles coarse-fine ghosts from neighbours and the parent parts of
    // the current parent, and reads sibling ghosts that the previous pass
    // pushed; other levels read the records.  The parent parts come from the
    // preceding fill (cf_current), from the records pass, or from the
    // other colour's face cells of passes 1 and 2, which store them for the
    // pass that follows.
    const bool ar = lev(k).arith && lev(k).aux;
    if (first_inline == 2 && !ar) first_inline = 1;
    int it0 = 0;
    const bool stream = lev(k).stream && h->opts.smoother != 3;
    if (stream && nsweeps > 0) {
      // streaming sweeps, Q and Qs in turn, ending in Q: an odd count starts
      // from a copy of Q in Qs (a copy costs a third of a sweep, two colour
      // passes cost nearly two).  Physical boundaries only: no ghost source,
      // parent parts or pushes
      const LevelDev<T>& Lk = lev(k);
      const int odd = nsweeps & 1;
      // (a failure is reported by the sequence's hipGetLastError check)
      if (odd && !qs_current)
        (void)hipMemcpyAsync(Lk.Qs, Lk.Q, (size_t)Lk.nq * sizeof(T), hipMemcpyDeviceToDevice, s);
      T* buf[2] = {odd ? Lk.Qs : Lk.Q, odd ? Lk.Q : Lk.Qs};
      for (int it = 0; it < nsweeps; ++it)
        launch_sweep_stream<T>(Lk, h->H[k - 1].pi[0], buf[it & 1], buf[(it + 1) & 1], (T)h->opts.omega, homog,
                               h->ncu, h->opts.stream_rows, s);
      fill(k);
      return;
    }
    if (nsweeps > 1 && h->opts.smoother == 2 && ar && lev(k).nsweep) {
      // one launch per red+black sweep for an even number of sweeps, Q and
      // Qs in turn, so that the last of them writes Q; an odd count does its
      // last sweep as two colour passes on Q.  The parent parts are current
      // after the preceding fill or the preparing correction, else the first
      // sweep computes and stores them (and, in record mode, its black half
      // needs them from the records)
      const LevelDev<T>& Lk = lev(k);
      T* buf[2] = {Lk.Q, Lk.Qs};
      it0 = nsweeps & ~1;
      for (int it = 0; it < it0; ++it) {
        const int mode = it ? kInlArith : (first_inline == 1 ? kInlRecords : first_inline == 2 ? kInlArith : 0);
        const int keep = it == 0 && (first_inline == 1 || (first_inline == 0 && !cf_current));
        launch_relax_sweep<T>(Lk, par(k), buf[it & 1], buf[(it + 1) & 1], mode, keep, (T)h->opts.omega, homog, s);
      }
    }
    for (int it = it0; it < nsweeps; ++it)
      for (int colour = 1; colour <= 2; ++colour) {
        const bool first = it == 0 && colour == 1, last = it == nsweeps - 1 && colour == 2;
        int mode = first ? (first_inline ? kInlRecords : 0) : kInlRecords;
        if (ar) {
          if (!first || first_inline == 2) mode = kInlArith;
          if (first_inline == 1 ? first : (first_inline == 0 && !cf_current && it == 0)) mode |= kKeepCf;
          if (!last) mode |= kPush;
        }
        launch_relax<T>(lev(k), par(k), colour, (T)h->opts.omega, homog, mode, s);
      }
    if (nsweeps > 0) fill(k);
  }
  void restrict_(int k) { launch_restrict<T>(lev(k), lev(k - 1), 0, s); }
  // A level whose odd sweep count starts from a copy of Q in Qs gets the copy
  // from the launch that last wrote its interior instead: the restriction
  // into a level the child covers entirely (below the root), or the
  // correction (a one-patch level, every cell).  qs_ready names that level.
  int qs_ready = 0;
  int unfilled = 0;  // the level whose fill after the restriction was left out
  bool odd_stream(int k, int nsweeps) const {  // relax(k, nsweeps) takes the streaming path and copies
    return (nsweeps & 1) && L[k - 1].stream && h->opts.smoother != 3 &&
           !(h->opts.smoother == 1 && tiled_fits<T>(L[k - 1]));
  }
  void ring(int k) { launch_ring<T>(lev(k), lev(k - 1), s); }
  void coarse_load(int k) { launch_coarse_load<T>(lev(k), lev(k - 1), h->opts.mass_correction, s); }
  void correct(int k, bool prep = false) { launch_correct<T>(lev(k), lev(k - 1), 0, s, prep ? 1 : 0); }
  void coarsest() {
    launch_coarsest<T>(h->cst, lev(1), homog, s);
    if (!h->cst.gtab) fill(1);  // else the solve's blocks wrote the ghosts
  }
  void vcycle_ref(int k) {  // ebVcycle.m:15-35, one launch per reference step
    if (k == 1) {
      coarsest();
      return;
    }
    relax(k, h->opts.npre, 0, k < h->nlev);
    restrict_(k);
    fill(k - 1);
    ring(k);
    coarse_load(k);
    vcycle_ref(k - 1);
    correct(k);
    fill(k);
    relax(k, h->opts.npost, 0, true);
  }
  void vcycle(int k, int kc) {  // ebVcycle.m:15-35, fused
    if (k <= kc && kc >= 2) {
      launch_coarse_cycle<T>(LVd, L.data(), k, h->cst, h->opts.npre, h->opts.npost, (T)h->opts.omega, homog,
                             h->opts.mass_correction, s, reflux_ft);
      return;
    }
    if (k == 1) {
      coarsest();
      return;
    }
    // below the top level the fill in the previous step is current.  A level
    // whose relax calls both run as region launches smooths into its other
    // buffer with its ghosts and the restriction in the same launch;
    // otherwise after a region relax its fill and the restriction run as one
    // launch (the fill's QC ring refresh is redone by the load's ring gather)
    const bool sw = swaps(k);
    const bool fuse = !sw && region(k, h->opts.npre);
    T* dup = nullptr;
    if (k <= h->root && k - 1 > kc && odd_stream(k - 1, h->opts.npre)) dup = lev(k - 1).Qs;
    // a level whose fill after the restriction was left out starts its
    // pre-smoothing from assembled ghosts (what that fill would have written)
    const bool first_assembled = unfilled == k;
    unfilled = 0;
    // the parent's fill after the restriction is left out when none of its
    // readers needs it: the FAS load evaluates the parent's physical ghosts
    // from the interior (pgh_ok: every parent ghost it reads is physical), and
    // the parent's pre-smoothing is a buffer-swapping region launch that
    // assembles its first pass's ghosts, or the parent is the coarsest level,
    // whose solve reads no ghost (it is filled after the solve)
    const bool pskip = lev(k).pgh_ok && (k - 1 == 1 ? kc <= 1 : (k - 1 > kc && swaps(k - 1)));
    // a one-tile level above the coarsest does the FAS load in its
    // pre-smoothing launch (the coarsest level's descriptor never swaps)
    const bool fl = sw && k == 2 && pskip && !dup && lev(k).nregion == 1 && region(k, h->opts.npre);
    const int fload = fl ? 1 | (h->opts.mass_correction ? 2 : 0) | 4 | (reflux_ft ? 8 : 0) : 0;
    relax(k, h->opts.npre, first_assembled ? 2 : 0, k < h->nlev || top_cf_current, fuse, sw ? 2 : 0, dup, false,
          fload);
    if (dup) qs_ready = k - 1;
    if (sw) {
    } else if (fuse) {
      launch_fill_copy_restrict<T>(lev(k), lev(k - 1), homog, lev(k).Qs, dup, s);
    } else if (dup) {
      launch_restrict<T>(lev(k), lev(k - 1), 0, s, dup);
    } else {
      restrict_(k);
    }
    if (!pskip) fill(k - 1);
    if (!fl)
      launch_load<T>(lev(k), lev(k - 1), h->opts.mass_correction, s, pskip ? 1 : 0, homog,
                     pskip && k - 1 > 1 ? (const T*)lev(k - 2).Q : nullptr, reflux_ft);
    unfilled = pskip ? k - 1 : 0;
    vcycle(k - 1, kc);
    const bool prep = h->opts.npost > 0 && lev(k).arith && lev(k).aux && h->opts.smoother != 1;
    if (sw && lev(k).rcorr) {
      // the post-smoothing launch applies the correction itself, with the
      // parent parts of its coarse-fine ghosts from the corrected parent
      relax(k, h->opts.npost, 2, false, false, 1, nullptr, true);
      return;
    }
    if (odd_stream(k, h->opts.npost)) {
      launch_correct<T>(lev(k), lev(k - 1), 0, s, prep ? 1 : 0, lev(k).Qs);
      qs_ready = k;
    } else {
      correct(k, prep);
    }
    if (h->opts.npost > 0) {
      relax(k, h->opts.npost, prep ? 2 : 1, false, false, sw ? 1 : 0);
    } else {
      fill(k);
    }
  }
  // ebTransfer.m:73-79 for the levels kk = nlev..klo (restrict kk into kk-1,
  // fill kk-1); the steps of the levels with at most kFilldownSmallCells cells
  // run as one single-workgroup launch
  //
  // qc: each step also does the composite residual's QC refresh
  // (ebCompResidual.m:23-39) of the level above the one it restricts, whose
  // parent is final after the step before: the rings of levels >= max(2,
  // root+1) and the top level's coarse-fine parent parts.  qc_done tells
  // whether every one of them found a step (the residual then skips its own
  // refresh launch).
  bool qc_done = false;
  void filldown(int klo = 2, bool qc = false) {
    const int kb = std::max(filldown_cut(h), klo - 1);
    const int kqc = std::max(2, h->root + 1), top = h->nlev;
    // level kqc's ring is gathered by step kqc-1, which must exist
    qc_done = qc && kqc - 1 >= std::max(klo, 2);
    for (int kk = h->nlev; kk > kb; --kk) {
      const int kc = kk + 1;
      const bool ring = qc_done && kc >= kqc && kc <= top;
      if (lev(kk).pcov) {  // the step as one launch
        launch_restrict_fill<T>(lev(kk), lev(kk - 1), par(kk - 1), homog, s, ring ? &lev(kc) : nullptr,
                                kc == top ? 1 : 0);
        continue;
      }
      restrict_(kk);
      fill(kk - 1);
      if (ring) qc_done = false;
    }
    launch_filldown_small<T>(LVd, kb, klo, homog, s, qc_done ? kqc : 1 << 30, top, qc_done ? top : 0);
  }
};

// F = FTrue on every level in one launch (ebSolveCycles.m:30-34)
template <typename T>
void enqueue_reset_force(bsam_hier* h, const LevelDev<T>* LVd, hipStream_t s) {
  int mc = 0;
  for (const HLevel& L : h->H) mc = std::max(mc, L.ncell);
  launch_reset_force<T>(LVd, h->nlev, mc, s);
}

// composite residual (ebCompResidual.m:23-84) on the FP64 hierarchy: one
// launch refreshes every QC (ring + interior), one evaluates every level's
// residual with its reflux chunks, one reduces the parts in a fixed order into
// dres and the pinned hres; rout (device array of per-level FP32 pointers)
// receives r when given.
// qc_fresh: the filldown did the QC refresh (Seq::filldown's qc_done)
void enqueue_residual(bsam_hier* h, float* const* rout, hipStream_t s, bool qc_fresh = false) {
  if (h->opts.mass_correction && !qc_fresh) launch_qc_refresh_all<double>(h->lv64, h->qc_items, h->nqc_items, s);
  launch_residual_all<double, float>(h->lv64, h->nlev, h->res_tiles, h->nres_tiles, h->opts.mass_correction,
                                     h->parts, rout, h->root, s);
  launch_reduce(h->parts, std::min(h->nres_tiles, kResidualParts), h->dres, s, h->hres, h->dseq);
}

}  // namespace

// the region tables of a level as bsam_export(k, 5) returns them:
// [ntile, rdepth, then per tile cx0 cy0 cw ch np nload nred nblack,
//  (pid ox oy) x np, the nload + nred + nblack cell entries]; [0, 0] when the
// level takes other smoother forms
static void region_export(const bsam_hier* h, const HLevel& L, std::vector<long long>& v) {
  v.assign(2, 0);
  const int rdepth = 2 * std::max(3, std::max(h->opts.npre, h->opts.npost));
  const bool stream = L.np == 1 && L.pi[0].mx % 2 == 0 && L.pi[0].bc[0] && L.pi[0].bc[1] && L.pi[0].bc[2] &&
                      L.pi[0].bc[3] && (long long)L.pi[0].mx * L.pi[0].my >= kStreamMinCells;
  RegionBuild rb;
  if (stream || !face_records_arith(L) || rdepth > kRegionMaxDepth || !build_region_tiles(L, rdepth, rb)) return;
  v[0] = (long long)rb.tiles.size();
  v[1] = rdepth;
  const int per_colour = kRegionThreads * kRegionK;
  for (size_t x = 0; x < rb.tiles.size(); ++x) {
    const RegionTile& t = rb.tiles[x];
    const long long hd[8] = {t.cx0, t.cy0, t.cw, t.ch, t.np, t.nload, t.nred, t.nblack};
    v.insert(v.end(), hd, hd + 8);
    for (int q = 0; q < t.np; ++q) {
      const RegionPatch& rp = rb.patches[x * kRegionMaxPatches + q];
      v.push_back(rb.pid[x * kRegionMaxPatches + q]); v.push_back(rp.ox); v.push_back(rp.oy);
    }
    const unsigned* c = rb.cells.data() + x * kRegionStride;
    for (int e = 0; e < t.nload; ++e) v.push_back(c[2 * per_colour + e]);
    for (int e = 0; e < t.nred; ++e) v.push_back(c[e]);
    for (int e = 0; e < t.nblack; ++e) v.push_back(c[per_colour + e]);
    v.push_back(t.ngh);
    for (int g = 0; g < t.ngh; ++g) {
      const RegionGhost& G = rb.ghosts[(size_t)t.gofs + g];
      v.push_back(G.dst);
      for (int k = 0; k < 6; ++k) v.push_back(region_slot(G, k) == kRegionNoSlot ? -1 : (long long)region_slot(G, k));
    }
    v.push_back(t.nrs);
    for (int e = 0; e < t.nrs; ++e) {
      v.push_back(rb.rs[(size_t)t.rofs + e].pad);
      v.push_back(rb.rs[(size_t)t.rofs + e].slot);
    }
  }
}

// the thread's last error message, for the host-only sources of the library
int bsam_set_error(int code, const char* msg) { return fail(code, "%s", msg); }

// ============================================================== C ABI
extern "C" {

void bsam_default_opts(bsam_solve_opts* o) {
  if (!o) return;
  o->npre = 3; o->npost = 3; o->omega = 1.0; o->rel_tol = 1e-9; o->max_vcycles = 60;
  o->ghost_interp = 0; o->mass_correction = 1; o->precision = BSAM_FP64; o->inner_cycles = 1;
  o->use_graph = 1; o->coarse_cells = 0; o->smoother = 0; o->stream_rows = 0;
}

const char* bsam_last_error(void) { return g_err.c_str(); }
const char* bsam_version(void) { return "bsamgpu 0.1 (gfx950)"; }

int bsam_ctx_create(int hip_device, bsam_ctx** out) {
  if (!out) return fail(BSAM_E_ARG, "bsamgpu:arg: null output");
  if (hip_device == -1) {  // host-only context: tables and exports, no device
    auto* c = new bsam_ctx;
    c->device = -1;
    *out = c;
    return 0;
  }
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (hip_device < 0 || hip_device >= n) return fail(BSAM_E_ARG, "bsamgpu:arg: device %d of %d", hip_device, n);
  HIPCHK(hipSetDevice(hip_device));
  auto* c = new bsam_ctx;
  c->device = hip_device;
  *out = c;
  return 0;
}

void bsam_ctx_destroy(bsam_ctx* ctx) { delete ctx; }

int build_coarsest_ghosts(bsam_hier* h);

static int hier_create_impl(bsam_ctx* ctx, const bsam_hier_desc* d, const bsam_solve_opts* o, bsam_hier* h) {
  if (!d || d->nlev < 1 || d->root_idx < 1 || d->root_idx > d->nlev || !d->n || !d->np || !d->boxes ||
      !d->bc_code || !d->Q || !d->FTrue || !d->DeW || !d->DnS || !d->CC || !d->InvDiagS || !d->g)
    return fail(BSAM_E_ARG, "bsamgpu:arg: incomplete hierarchy descriptor");
  h->ctx = ctx;
  h->nlev = d->nlev;
  h->root = d->root_idx;
  for (int i = 0; i < 4; ++i) h->dom[i] = d->domain[i];
  if (o) h->opts = *o; else bsam_default_opts(&h->opts);
  if (ctx->device >= 0) {
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
  }
  h->H.resize(h->nlev);
  size_t pb = 0, oq = 0, oc = 0, ox = 0, oy = 0, og = 0;
  for (int k = 0; k < h->nlev; ++k) {
    HLevel& L = h->H[k];
    L.n[0] = d->n[2 * k]; L.n[1] = d->n[2 * k + 1];
    if (L.n[0] < 2 || L.n[1] < 2) return fail(BSAM_E_ARG, "bsamgpu:arg: level %d grid too small", k + 1);
    L.h[0] = (d->domain[1] - d->domain[0]) / L.n[0];
    L.h[1] = (d->domain[3] - d->domain[2]) / L.n[1];
    L.np = d->np[k];
    if (L.np < 1) return fail(BSAM_E_ARG, "bsamgpu:arg: level %d has no patches", k + 1);
    for (int p = 0; p < L.np; ++p, ++pb) {
      Box b{d->boxes[4 * pb], d->boxes[4 * pb + 1], d->boxes[4 * pb + 2], d->boxes[4 * pb + 3]};
      if (b.ilo < 1 || b.ihi > L.n[0] || b.jlo < 1 || b.jhi > L.n[1] || b.ilo > b.ihi || b.jlo > b.jhi)
        return fail(BSAM_E_BOX, "eb_make_level:box: invalid patch box.");
      PatchInfo P{};
      P.mx = b.ihi - b.ilo + 1; P.my = b.jhi - b.jlo + 1;
      P.magic = row_magic(P.mx, P.my);
      if (k > 0 && (P.mx % 2 || P.my % 2))
        return fail(BSAM_E_ALIGN, "eb_build_ops:align: patch not aligned to parent cells.");
      if (P.mx < 2 || P.my < 2) return fail(BSAM_E_BOX, "eb_make_level:box: patch narrower than 2 cells.");
      P.par0 = (b.ilo + b.jlo) % 2;
      for (int s = 0; s < 4; ++s) {
        P.bc[s] = d->bc_code[4 * pb + s];
        if (P.bc[s] < 0 || P.bc[s] > 2) return fail(BSAM_E_ARG, "bsamgpu:arg: bad bc code");
      }
      const bool edge[4] = {b.ilo == 1, b.ihi == L.n[0], b.jlo == 1, b.jhi == L.n[1]};
      for (int s = 0; s < 4; ++s)
        if ((P.bc[s] != 0) != edge[s])
          return fail(BSAM_E_ARG, "bsamgpu:arg: bc code of side %d inconsistent with the box", s);
      P.qoff = L.nq; P.coff = L.ncell; P.xoff = L.nx; P.yoff = L.ny; P.goff = L.ng;
      P.qcoff = L.nqc;
      L.nq += (P.mx + 2) * (P.my + 2);
      L.ncell += P.mx * P.my;
      L.nx += (P.mx + 1) * P.my;
      L.ny += P.mx * (P.my + 1);
      L.ng += 2 * (P.mx + P.my);
      if (k > 0) L.nqc += (P.mx / 2 + 2) * (P.my / 2 + 2);
      L.box.push_back(b);
      L.pi.push_back(P);
    }
    L.Q.assign(d->Q + oq, d->Q + oq + L.nq);
    L.FT.assign(d->FTrue + oc, d->FTrue + oc + L.ncell);
    L.CC.assign(d->CC + oc, d->CC + oc + L.ncell);
    L.ID.assign(d->InvDiagS + oc, d->InvDiagS + oc + L.ncell);
    L.DeW.assign(d->DeW + ox, d->DeW + ox + L.nx);
    L.DnS.assign(d->DnS + oy, d->DnS + oy + L.ny);
    L.G.assign(d->g + og, d->g + og + L.ng);
    oq += L.nq; oc += L.ncell; ox += L.nx; oy += L.ny; og += L.ng;
    for (double v : L.DeW) {
      if (!std::isfinite(v)) return fail(BSAM_E_NONFINITE, "eb_make_level:eval: coefficient function returned non-finite values.");
      if (v <= 0) return fail(BSAM_E_D, "eb_make_level:D: D(x,y) must be strictly positive at faces.");
    }
    for (double v : L.DnS) {
      if (!std::isfinite(v)) return fail(BSAM_E_NONFINITE, "eb_make_level:eval: coefficient function returned non-finite values.");
      if (v <= 0) return fail(BSAM_E_D, "eb_make_level:D: D(x,y) must be strictly positive at faces.");
    }
    for (const auto* a : {&L.CC, &L.FT, &L.ID, &L.Q, &L.G})
      for (double v : *a)
        if (!std::isfinite(v)) return fail(BSAM_E_NONFINITE, "eb_make_level:eval: coefficient function returned non-finite values.");
    L.vis.assign(L.ncell, 1);
    L.tgt_of_cell.assign(L.ncell, -1);
    make_tiles(L);
  }
  if (h->H[0].np != 1) return fail(BSAM_E_ARG, "bsamgpu:arg: level 1 must be one full-domain patch");
  const bool qif = h->opts.ghost_interp == 0;
  build_phys_level(h->H[0]);
  for (int k = 1; k < h->nlev; ++k) {
    int rc = build_level_tables(h->H[k], h->H[k - 1], qif);
    if (rc) return rc;
  }
  for (int k = 1; k < h->nlev; ++k) parent_ghost_sources(h->H[k], h->H[k - 1]);
  for (auto& L : h->H)  // patch records are final now: copy them into the tiles
    for (TileRec& t : L.tiles) t.pi = L.pi[t.p];
  // singular iff all-Neumann level 1 and C == 0 everywhere (ebCoarsest.m:58-66)
  bool allneu = true;
  for (int s = 0; s < 4; ++s) allneu = allneu && h->H[0].pi[0].bc[s] != 1;
  double cmax = 0;
  for (auto& L : h->H)
    for (double v : L.CC) cmax = std::max(cmax, std::fabs(v));
  h->singular = allneu && cmax == 0;
  std::vector<double> Ainv, bcRhs;
  int M = 0;
  int rc = coarsest_assemble(h->H[0], h->singular, Ainv, bcRhs, M);
  if (rc) return rc;
  for (int k = 0; k < h->nlev; ++k)
    if ((rc = validate_level(h->H[k], k > 0 ? &h->H[k - 1] : nullptr, k + 1 < h->nlev ? &h->H[k + 1] : nullptr)))
      return rc;
  h->cst.N = h->H[0].n[0] * h->H[0].n[1];
  h->cst.M = M;
  if (ctx->device < 0) {  // host-only: no device state; the host inverse is kept for export
    h->ainv_host = std::move(Ainv);
    return 0;
  }
  if ((rc = upload(h, &h->cst.Ainv, Ainv))) return rc;
  if ((rc = upload(h, &h->cst.bcRhs, bcRhs))) return rc;
  if ((rc = build_coarsest_ghosts(h))) return rc;
  // device levels
  h->d64.resize(h->nlev);
  for (int k = 0; k < h->nlev; ++k)
    if ((rc = make_dev_level<double>(h, h->H[k], h->d64[k], false, nullptr))) return rc;
  h->Q0.resize(h->nlev);
  for (int k = 0; k < h->nlev; ++k)
    if ((rc = upload(h, &h->Q0[k], h->H[k].Q))) return rc;
  if ((rc = upload(h, &h->q0dev, h->Q0))) return rc;
  h->part_off.resize(h->nlev + 1);
  h->nparts = 0;
  std::vector<int2> res_tiles, qc_items;
  for (int k = 0; k < h->nlev; ++k) {
    h->part_off[k] = h->nparts;
    h->nparts += (int)h->H[k].tiles.size();
    for (int t = 0; t < (int)h->H[k].tiles.size(); ++t) res_tiles.push_back(make_int2(k, t));
    if (k + 1 >= std::max(2, h->root + 1)) {
      const int n = (int)h->H[k].ring_all.size() + (int)h->H[k].foot.size();
      for (int t = 0; t < n; ++t) qc_items.push_back(make_int2(k, t));
      // the top level's parent parts, from the parent as the filldown left it:
      // the next cycle's first relax call of the top level reads them
      if (k + 1 == h->nlev)
        for (int g = 0; g < (int)h->H[k].ghost.size(); ++g) {
          const int kind = h->H[k].ghost[(size_t)g].kind;
          if (kind == G_QIF || kind == G_QIF_END || kind == G_LIF || kind == G_LIF_CORNER)
            qc_items.push_back(make_int2(k, n + g));
        }
    }
  }
  h->part_off[h->nlev] = h->nparts;
  h->nres_tiles = (int)res_tiles.size();
  h->nqc_items = (int)qc_items.size();
  if ((rc = upload(h, &h->res_tiles, res_tiles))) return rc;
  if ((rc = upload(h, &h->qc_items, qc_items))) return rc;
  if ((rc = upload(h, &h->lv64, h->d64))) return rc;
  if ((rc = h->mem.alloc(&h->parts, std::max(1, h->nparts)))) return rc;
  if ((rc = h->mem.alloc(&h->dres, 1))) return rc;
  if ((rc = h->mem.alloc(&h->dseq, 1))) return rc;
  HIPCHK(hipMemset(h->dseq, 0, sizeof(unsigned long long)));
  HIPCHK(hipHostMalloc((void**)&h->hres, sizeof(ResidualOut), hipHostMallocDefault));
  // a recycled pinned block can hold an earlier hierarchy's flag word, which
  // the first wait_residual would take for this one's first result
  std::memset(h->hres, 0, sizeof(ResidualOut));
  HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  h->cell_base.resize(h->nlev + 1);
  h->ncell_all = 0;
  for (int k = 0; k < h->nlev; ++k) {
    h->cell_base[k] = h->ncell_all;
    h->ncell_all += h->H[k].ncell;
  }
  h->cell_base[h->nlev] = h->ncell_all;
  return 0;
}

static int need_dev(const bsam_hier* h) {
  if (!h) return fail(BSAM_E_ARG, "bsamgpu:arg: null hierarchy");
  if (h->ctx->device < 0) return fail(BSAM_E_ARG, "bsamgpu:arg: host-only hierarchy has no device state");
  return 0;
}

int bsam_hier_create(bsam_ctx* ctx, const bsam_hier_desc* d, const bsam_solve_opts* o, bsam_hier** out) {
  if (!ctx || !out) return fail(BSAM_E_ARG, "bsamgpu:arg: null context or output");
  auto h = std::make_unique<bsam_hier>();
  int rc = hier_create_impl(ctx, d, o, h.get());
  if (rc) return rc;
  *out = h.release();
  return 0;
}

void bsam_hier_destroy(bsam_hier* h) {
  if (!h) return;
  if (h->ctx->device >= 0) (void)hipSetDevice(h->ctx->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  delete h;
}

// FP32 copy of the hierarchy for the inner V-cycles (tables shared with the
// FP64 levels).  Q and QC of all levels live in one buffer each so that the
// per-step zeroing is a single launch.
static int ensure32(bsam_hier* h) {
  if (h->have32) return 0;
  h->d32.resize(h->nlev);
  long long nq = 0, nqc = 0;
  for (int k = 0; k < h->nlev; ++k) {
    int rc = make_dev_level<float>(h, h->H[k], h->d32[k], true, &h->d64[k]);
    if (rc) return rc;
    launch_fill_value<float>(h->d32[k].FT, h->H[k].ncell, 0.0f, h->stream);
    nq += h->H[k].nq;
    nqc += h->H[k].nqc;
  }
  int rc;
  if ((rc = h->mem.alloc(&h->q32all, std::max(1LL, nq)))) return rc;
  if ((rc = h->mem.alloc(&h->qc32all, std::max(1LL, nqc)))) return rc;
  h->nq32all = nq;
  h->nqc32all = nqc;
  long long oq = 0, oqc = 0;
  std::vector<float*> rout(h->nlev);
  for (int k = 0; k < h->nlev; ++k) {
    h->d32[k].Q = h->q32all + oq;
    h->d32[k].QC = h->qc32all + oqc;
    oq += h->H[k].nq;
    oqc += h->H[k].nqc;
    rout[k] = h->d32[k].FT;
  }
  launch_fill_value<float>(h->q32all, nq, 0.0f, h->stream);
  launch_fill_value<float>(h->qc32all, nqc, 0.0f, h->stream);
  if ((rc = upload(h, &h->lv32, h->d32))) return rc;
  if ((rc = upload(h, &h->rout32, rout))) return rc;
  for (int k = 0; k < h->nlev; ++k) rout[k] = h->d32[k].F;
  if ((rc = upload(h, &h->rout32f, rout))) return rc;
  h->have32 = true;
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

static int check_level(const bsam_hier* h, int k, int lo) {
  if (int rc_ = need_dev(h)) return rc_;
  if (h->ctx->device < 0) return fail(BSAM_E_ARG, "bsamgpu:arg: host-only hierarchy has no device state");
  if (k < lo || k > h->nlev) return fail(BSAM_E_ARG, "bsamgpu:arg: level %d out of range", k);
  return 0;
}

static int sync(bsam_hier* h) {
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

static Seq<double> seq64(bsam_hier* h, hipStream_t s) { return Seq<double>{h, h->d64, h->lv64, 0, s}; }
static Seq<float> seq32(bsam_hier* h, hipStream_t s) { return Seq<float>{h, h->d32, h->lv32, 1, s}; }

int bsam_fill_ghosts(bsam_hier* h, int k) {
  if (int rc = check_level(h, k, 1)) return rc;
  seq64(h, h->stream).fill(k);
  return sync(h);
}

int bsam_relax_level(bsam_hier* h, int k, int nsweeps) {
  if (int rc = check_level(h, k, 1)) return rc;
  seq64(h, h->stream).relax(k, nsweeps);
  return sync(h);
}

int bsam_transfer(bsam_hier* h, int k, int what) {
  if (int rc_ = need_dev(h)) return rc_;
  Seq<double> s = seq64(h, h->stream);
  if (what == BSAM_FILLDOWN) {
    s.filldown();
    return sync(h);
  }
  if (int rc = check_level(h, k, 2)) return rc;
  switch (what) {
    case BSAM_RESTRICT: s.restrict_(k); break;
    case BSAM_RING: s.ring(k); break;
    case BSAM_CORRECT: s.correct(k); break;
    default: return fail(BSAM_E_ARG, "eb_transfer:what: unknown transfer %d", what);
  }
  return sync(h);
}

int bsam_coarse_load(bsam_hier* h, int k) {
  if (int rc = check_level(h, k, 2)) return rc;
  seq64(h, h->stream).coarse_load(k);
  return sync(h);
}

int bsam_coarsest_solve(bsam_hier* h) {
  if (int rc_ = need_dev(h)) return rc_;
  seq64(h, h->stream).coarsest();
  return sync(h);
}

int bsam_vcycle(bsam_hier* h) {
  if (int rc_ = need_dev(h)) return rc_;
  Seq<double> q = seq64(h, h->stream);
  q.vcycle(h->nlev, coarse_cut(h, h->opts.coarse_cells));
  q.flush_rings();
  return sync(h);
}

int bsam_reset_force(bsam_hier* h) {
  if (int rc_ = need_dev(h)) return rc_;
  enqueue_reset_force<double>(h, h->lv64, h->stream);
  return sync(h);
}

int bsam_init_level_prolong(bsam_hier* h, int k) {
  if (int rc = check_level(h, k, 2)) return rc;
  Seq<double> s = seq64(h, h->stream);
  s.fill(k - 1);
  launch_correct<double>(h->d64[k - 1], h->d64[k - 2], 1, h->stream);
  s.fill(k);
  return sync(h);
}

// |r|_inf / max(|FTrue|_inf, realmin) (ebCompResidual.m:80-84); NaN when the
// residual holds non-finite values (the max-reductions skip NaN, the sum does not)
static double residual_rel(const ResidualOut& r) {
  if (!std::isfinite(r.l2sq)) return NAN;
  return r.linf / std::max(r.fnorm, 2.2250738585072014e-308);
}

// the scalars of the reduction the host enqueued last (one more than the
// count so far): polls the flag the reduction writes into the pinned hres
// after them, which returns as soon as the result lands rather than when the
// runtime reports the stream idle; after kWaitSpinNs without it, waits for
// the stream (which also reports a failed launch)
constexpr long long kWaitSpinNs = 20000000;
static int wait_residual(bsam_hier* h) {
  const long long want = ++h->seq_expect;
  volatile long long* flag = &h->hres->pad2[0];
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned n = 1; *flag != want; ++n) {
    if ((n & 1023u) == 0 &&
        std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count() >
            kWaitSpinNs) {
      HIPCHK(hipStreamSynchronize(h->stream));
      break;
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return 0;
}

// after enqueue_residual, whose reduction wrote the scalars to hres
static int read_residual(bsam_hier* h, bsam_residual* out) {
  if (int rc = wait_residual(h)) return rc;
  if (out) {
    out->linf = h->hres->linf;
    out->l2 = std::sqrt(h->hres->l2sq);
    out->fnorm = std::max(h->hres->fnorm, 2.2250738585072014e-308);
    out->rel = residual_rel(*h->hres);
    out->ncell = h->hres->ncell;
  }
  return 0;
}

int bsam_comp_residual(bsam_hier* h, bsam_residual* out) {
  if (int rc_ = need_dev(h)) return rc_;
  enqueue_residual(h, nullptr, h->stream);
  HIPCHK(hipGetLastError());
  return read_residual(h, out);
}

// one cycle of the plain FP64 loop: V-cycle, filldown, residual, residual
// scalars to pinned host memory.  The loop's F = FTrue before every cycle
// (ebSolveCycles.m:30-34) runs once, before the first (bsam_solve_cycles):
// between cycles F differs from FTrue only where the FAS loads write, and
// they overwrite covered cells and sum the reflux targets from FTrue.
static int enqueue_cycle64(bsam_hier* h, hipStream_t s) {
  Seq<double> q = seq64(h, s);
  q.reflux_ft = 1;
  q.top_cf_current = h->opts.mass_correction && h->nlev >= std::max(2, h->root + 1);
  q.vcycle(h->nlev, coarse_cut(h, h->opts.coarse_cells));
  // rings the correcting post-smoothing left out: the filldown's fills
  // refresh those of levels below the top, the residual's QC refresh the top
  // level's when it is an AMR level with flux corrections; else refresh it
  if (h->nlev <= h->root || !h->opts.mass_correction) q.flush_ring(h->nlev);
  q.filldown(2, h->opts.mass_correction != 0);
  enqueue_residual(h, nullptr, s, q.qc_done);
  return 0;
}

// where the FP64 residual writes the FP32 right side of the correction
// problem: straight into F for one inner cycle per step (levels >= root are
// rewritten every step; the FAS load sets the sub-root levels' F before it is
// read), into FTrue when F must be reset from it before each inner cycle
static float* const* mixed_rout(const bsam_hier* h) {
  return h->opts.inner_cycles <= 1 ? h->rout32f : h->rout32;
}

// one step of the mixed loop: FP32 V-cycles on the correction problem
// (homogeneous boundary data, zero initial guess, the current FP64 composite
// residual as right side), FP64 update, filldown, ghost refresh and composite
// residual (which also writes the next right side).  With a zero initial
// guess and homogeneous data every ghost is zero, so no initial fill is needed;
// covered cells of the correction are recomputed by the FP64 filldown, so the
// FP32 filldown only runs between inner cycles.  Levels below the root carry
// no visible cells and the next cycle's FP32 correction starts from zero, so
// the FP64 update, filldown and fill stop at the root; the sub-root levels
// are brought up to date once, after the last cycle (sub_root_filldown).
static int enqueue_cycle_mixed(bsam_hier* h, hipStream_t s) {
  Seq<float> q = seq32(h, s);
  const int kc = coarse_cut(h, h->opts.coarse_cells);
  launch_fill_value<float>(h->q32all, h->nq32all, 0.0f, s);
  launch_fill_value<float>(h->qc32all, h->nqc32all, 0.0f, s);
  const int inner = std::max(1, h->opts.inner_cycles);
  for (int it = 0; it < inner; ++it) {
    if (inner > 1) enqueue_reset_force<float>(h, h->lv32, s);  // else the residual wrote F itself
    q.vcycle(h->nlev, kc);
    if (it + 1 < inner) q.filldown();
  }
  for (int k = h->root; k <= h->nlev; ++k) launch_add_correction(h->d64[k - 1], h->d32[k - 1], s);
  Seq<double> d = seq64(h, s);
  d.filldown(h->root + 1);
  for (int k = h->root; k <= h->nlev; ++k) d.fill(k);
  enqueue_residual(h, mixed_rout(h), s);
  return 0;
}

// the full FP64 filldown after the last mixed cycle: levels >= root are
// rewritten with the values they hold (restriction of an unchanged child), the
// sub-root levels receive the restriction of the root, as after FP64 cycles
static void sub_root_filldown(bsam_hier* h) {
  if (h->root > 1) seq64(h, h->stream).filldown();
}

static int check_stream_rows(int r) {
  if (r != 0 && r != 8 && r != 16 && r != 32 && r != 64)
    return fail(BSAM_E_ARG, "bsamgpu:arg: stream_rows must be 0, 8, 16, 32 or 64");
  return 0;
}

static int run_cycle(bsam_hier* h, bool mixed) {
  hipStream_t s = h->stream;
  if (!h->opts.use_graph) {
    int rc = mixed ? enqueue_cycle_mixed(h, s) : enqueue_cycle64(h, s);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return 0;
  }
  // every option the enqueued sequence depends on (omega is baked into the
  // kernel arguments)
  const std::vector<double> key = {mixed ? 1.0 : 0.0,
                                   (double)h->opts.inner_cycles,
                                   (double)h->opts.npre,
                                   (double)h->opts.npost,
                                   (double)h->opts.mass_correction,
                                   h->opts.omega,
                                   (double)coarse_cut(h, h->opts.coarse_cells),
                                   (double)h->opts.smoother,
                                   (double)h->opts.stream_rows};
  if (!h->gexec || h->graph_key != key) {
    if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
    if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = mixed ? enqueue_cycle_mixed(h, s) : enqueue_cycle64(h, s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : fail(BSAM_E_HIP, "bsamgpu:hip: graph capture failed: %s", hipGetErrorString(e));
    }
    h->graph = g;
    HIPCHK(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
    h->graph_key = key;
  }
  HIPCHK(hipGraphLaunch(h->gexec, s));
  return 0;
}

int bsam_solve_cycles(bsam_hier* h, const bsam_solve_opts* o, bsam_stage_stats* out) {
  if (int rc_ = need_dev(h)) return rc_;
  if (o) {
    if (o->ghost_interp != h->opts.ghost_interp)
      return fail(BSAM_E_ARG, "bsamgpu:arg: GhostInterp is fixed at hierarchy creation");
    h->opts = *o;
  }
  if (h->opts.npre < 0 || h->opts.npost < 0 || h->opts.max_vcycles < 0)
    return fail(BSAM_E_ARG, "bsamgpu:arg: negative sweep or cycle count");
  if (h->opts.smoother < 0 || h->opts.smoother > 4)
    return fail(BSAM_E_ARG, "bsamgpu:arg: smoother must be 0, 1, 2, 3 or 4");
  if (int rc = check_stream_rows(h->opts.stream_rows)) return rc;
  const bool mixed = h->opts.precision == BSAM_MIXED;
  if (mixed)
    if (int rc = ensure32(h)) return rc;
  HIPCHK(hipSetDevice(h->ctx->device));
  EventPair ev;
  HIPCHK(ev.create());
  hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
  HIPCHK(hipEventRecord(e0, h->stream));
  // ebSolveCycles.m:11-19: fill every level, baseline residual
  Seq<double> d = seq64(h, h->stream);
  for (int k = 1; k <= h->nlev; ++k) d.fill(k);
  enqueue_residual(h, mixed ? mixed_rout(h) : nullptr, h->stream);
  HIPCHK(hipGetLastError());
  bsam_residual res{};
  if (int rc = read_residual(h, &res)) return rc;
  std::vector<double> hist{res.linf};
  double rel = res.rel;
  long long ncell = res.ncell;
  int it = 0;
  while (it < h->opts.max_vcycles && rel > h->opts.rel_tol) {
    ++it;
    if (it == 1 && !mixed) enqueue_reset_force<double>(h, h->lv64, h->stream);  // see enqueue_cycle64
    if (int rc = run_cycle(h, mixed)) return rc;
    if (int rc = wait_residual(h)) return rc;
    const ResidualOut r = *h->hres;
    hist.push_back(r.linf);
    rel = residual_rel(r);
    ncell = r.ncell;
    if (!std::isfinite(rel)) break;
  }
  if (mixed && it > 0) sub_root_filldown(h);
  HIPCHK(hipEventRecord(e1, h->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  if (out) {
    out->amr_levels = h->nlev - h->root + 1;
    out->nlev = h->nlev;
    out->vcycles = it;
    out->final_rel = rel;
    out->visible_cells = ncell;
    long long tot = 0;
    for (int k = h->root; k <= h->nlev; ++k) tot += h->H[k - 1].ncell;
    out->total_cells = tot;
    out->avg_factor = hist.size() >= 2
                          ? std::pow(hist.back() / std::max(hist.front(), 2.2250738585072014e-308), 1.0 / (hist.size() - 1))
                          : NAN;
    out->res_hist_len = 0;
    if (out->res_hist)
      for (size_t i = 0; i < hist.size() && (int)i < out->res_hist_cap; ++i) {
        out->res_hist[i] = hist[i];
        out->res_hist_len = (int)i + 1;
      }
    out->gpu_ms = ms;
  }
  return 0;
}

int bsam_composite_mean(bsam_hier* h, double* mean, double* area) {
  if (int rc_ = need_dev(h)) return rc_;
  for (int k = h->root; k <= h->nlev; ++k) launch_mass<double>(h->d64[k - 1], h->parts + h->part_off[k - 1], h->stream);
  launch_reduce(h->parts + h->part_off[h->root - 1], h->nparts - h->part_off[h->root - 1], h->dres, h->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(h->hres, h->dres, sizeof(ResidualOut), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  double ar = 0;
  for (int k = h->root; k <= h->nlev; ++k) {
    long long nv = 0;
    for (uint8_t v : h->H[k - 1].vis) nv += v;
    ar += nv * (h->H[k - 1].h[0] * h->H[k - 1].h[1]);
  }
  if (area) *area = ar;
  if (mean) *mean = h->hres->l2sq / std::max(ar, 2.2250738585072014e-308);
  return 0;
}

int bsam_tag_level(bsam_hier* h, int k, int indicator, double threshold, int buffer, unsigned char* tags,
                   unsigned char* raw, long long* ntag) {
  if (int rc = check_level(h, k, 1)) return rc;
  if (indicator != BSAM_TAG_ULAP && indicator != BSAM_TAG_RTE)
    return fail(BSAM_E_ARG, "bsamgpu:arg: indicator must be BSAM_TAG_ULAP or BSAM_TAG_RTE");
  const HLevel& L = h->H[k - 1];
  if (indicator == BSAM_TAG_RTE)
    for (const PatchInfo& P : L.pi)
      if ((P.mx | P.my) & 1) return fail(BSAM_E_ARG, "bsamgpu:arg: the rte indicator needs even patch widths");
  const long long n = (long long)L.n[0] * L.n[1];
  std::vector<int2> lo(L.np);
  for (int p = 0; p < L.np; ++p) lo[p] = make_int2(L.box[p].ilo - 1, L.box[p].jlo - 1);
  HIPCHK(hipSetDevice(h->ctx->device));
  DevBufs<double> mem;
  uint8_t* d = nullptr;
  int2* dlo = nullptr;
  if (int rc = mem.alloc(&d, (size_t)(3 * n))) return rc;
  if (int rc = mem.alloc(&dlo, lo.size())) return rc;
  HIPCHK(hipMemcpyAsync(dlo, lo.data(), lo.size() * sizeof(int2), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemsetAsync(d, 0, (size_t)(2 * n), h->stream));
  launch_tag_cells(h->d64[k - 1], dlo, indicator, threshold, d, d + n, h->stream);
  launch_tag_finish(d, d + n, L.n[0], L.n[1], buffer, d + 2 * n, h->stream);
  HIPCHK(hipGetLastError());
  std::vector<uint8_t> out;
  uint8_t* dst = tags;
  if (!dst && ntag) {
    out.resize(n);
    dst = out.data();
  }
  if (dst) HIPCHK(hipMemcpyAsync(dst, d + 2 * n, (size_t)n, hipMemcpyDeviceToHost, h->stream));
  if (raw) HIPCHK(hipMemcpyAsync(raw, d, (size_t)n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (ntag) {
    long long c = 0;
    for (long long i = 0; i < n; ++i) c += dst[i];
    *ntag = c;
  }
  return 0;
}

// ---- tagging without allocation
// The tag maps and the patch corner table of bsam_tag_level_static live in
// one file-scope device array.  In host code `scratch` names a host-side
// stand-in, not device memory: every use goes through the device address
// that hipGetSymbolAddress returns, looked up once per device.
constexpr long long kTagStaticMaxCells = 1LL << 22;    // n1*n2 of the largest level map (2048 x 2048)
constexpr int kTagStaticMaxPatches = 16384;
constexpr size_t kTagStaticLoBytes = (size_t)kTagStaticMaxPatches * sizeof(int2);
constexpr size_t kTagStaticBytes = kTagStaticLoBytes + 3 * (size_t)kTagStaticMaxCells;
static __device__ __align__(256) unsigned char scratch[kTagStaticBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// the calling thread's current device must be `dev`
int tag_scratch(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(BSAM_E_ARG, "bsamgpu:arg: device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(BSAM_E_HIP, "bsamgpu:hip: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

int bsam_tag_level_static(bsam_hier* h, int k, int indicator, double threshold, int buffer, unsigned char* tags,
                          unsigned char* raw, long long* ntag) {
  if (int rc = check_level(h, k, 1)) return rc;
  if (indicator != BSAM_TAG_ULAP && indicator != BSAM_TAG_RTE)
    return fail(BSAM_E_ARG, "bsamgpu:arg: indicator must be BSAM_TAG_ULAP or BSAM_TAG_RTE");
  const HLevel& L = h->H[k - 1];
  if (indicator == BSAM_TAG_RTE)
    for (const PatchInfo& P : L.pi)
      if ((P.mx | P.my) & 1) return fail(BSAM_E_ARG, "bsamgpu:arg: the rte indicator needs even patch widths");
  const long long n = (long long)L.n[0] * L.n[1];
  if (n > kTagStaticMaxCells)
    return fail(BSAM_E_ARG, "bsamgpu:arg: level %d map has %lld cells, above the %lld of bsam_tag_level_static", k, n,
                kTagStaticMaxCells);
  if (L.np > kTagStaticMaxPatches)
    return fail(BSAM_E_ARG, "bsamgpu:arg: level %d has %d patches, above the %d of bsam_tag_level_static", k, L.np,
                kTagStaticMaxPatches);
  HIPCHK(hipSetDevice(h->ctx->device));
  unsigned char* base = nullptr;
  if (int rc = tag_scratch(h->ctx->device, &base)) return rc;
  int2* dlo = (int2*)base;
  uint8_t* d = base + kTagStaticLoBytes;
  // patch corners through a fixed host buffer (synchronous copies)
  int2 chunk[256];
  for (int p0 = 0; p0 < L.np; p0 += 256) {
    const int m = std::min(256, L.np - p0);
    for (int q = 0; q < m; ++q) chunk[q] = make_int2(L.box[p0 + q].ilo - 1, L.box[p0 + q].jlo - 1);
    HIPCHK(hipMemcpyAsync(dlo + p0, chunk, (size_t)m * sizeof(int2), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  // raw and coverage maps start at zero: the array keeps the last call's bytes
  launch_fill_value<uint8_t>(d, 2 * n, 0, h->stream);
  launch_tag_cells(h->d64[k - 1], dlo, indicator, threshold, d, d + n, h->stream);
  launch_tag_finish(d, d + n, L.n[0], L.n[1], buffer, d + 2 * n, h->stream);
  HIPCHK(hipGetLastError());
  if (tags) HIPCHK(hipMemcpyAsync(tags, d + 2 * n, (size_t)n, hipMemcpyDeviceToHost, h->stream));
  if (raw) HIPCHK(hipMemcpyAsync(raw, d, (size_t)n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (ntag) {
    long long c = 0;
    if (tags) {
      for (long long i = 0; i < n; ++i) c += tags[i];
    } else {  // count through a fixed host buffer
      unsigned char buf[65536];
      for (long long i0 = 0; i0 < n; i0 += (long long)sizeof buf) {
        const size_t m = (size_t)std::min<long long>((long long)sizeof buf, n - i0);
        HIPCHK(hipMemcpy(buf, d + 2 * n + i0, m, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < m; ++i) c += buf[i];
      }
    }
    *ntag = c;
  }
  return 0;
}

int bsam_polygon_fields(int device, const double* vx, const double* vy, int nv, const double* x, const double* y,
                        long long npts, double eps, double* d, double* phi, double* delta) {
  if (nv < 4 || !vx || !vy || npts < 0 || (npts > 0 && (!x || !y || !d)))
    return fail(BSAM_E_ARG, "bsamgpu:arg: bad polygon or point buffers");
  if ((phi || delta) && !(eps > 0)) return fail(BSAM_E_ARG, "dd_transmission:epsilon: epsilon must be positive");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(BSAM_E_ARG, "bsamgpu:arg: no HIP device %d", device);
  if (npts == 0) return 0;
  const int ns = nv - 1;
  std::vector<double> seg(5 * (size_t)ns);
  for (int j = 0; j < ns; ++j) {
    const double ex = vx[j + 1] - vx[j], ey = vy[j + 1] - vy[j];
    seg[j] = vx[j];
    seg[ns + j] = vy[j];
    seg[2 * ns + j] = ex;
    seg[3 * ns + j] = ey;
    seg[4 * ns + j] = ex * ex + ey * ey;
  }
  HIPCHK(hipSetDevice(device));
  DevBufs<double> mem;
  double *dseg, *dx, *dy, *dd, *dphi = nullptr, *ddel = nullptr;
  if (int rc = mem.alloc(&dseg, seg.size())) return rc;
  if (int rc = mem.alloc(&dx, (size_t)npts)) return rc;
  if (int rc = mem.alloc(&dy, (size_t)npts)) return rc;
  if (int rc = mem.alloc(&dd, (size_t)npts)) return rc;
  if (phi)
    if (int rc = mem.alloc(&dphi, (size_t)npts)) return rc;
  if (delta)
    if (int rc = mem.alloc(&ddel, (size_t)npts)) return rc;
  const size_t nb = (size_t)npts * sizeof(double);
  HIPCHK(hipMemcpy(dseg, seg.data(), seg.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dx, x, nb, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dy, y, nb, hipMemcpyHostToDevice));
  launch_polygon_fields(dseg, ns, dx, dy, npts, eps, dd, dphi, ddel, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(d, dd, nb, hipMemcpyDeviceToHost));
  if (phi) HIPCHK(hipMemcpy(phi, dphi, nb, hipMemcpyDeviceToHost));
  if (delta) HIPCHK(hipMemcpy(delta, ddel, nb, hipMemcpyDeviceToHost));
  return 0;
}

int bsam_cluster_tags(const unsigned char* tags, const unsigned char* allowed, int n1, int n2, int min_width,
                      double fill_ratio, int* boxes, int cap, int* nboxes, int* ndropped) {
  if (!tags || !allowed || n1 < 1 || n2 < 1 || cap < 0 || (cap > 0 && !boxes) || !nboxes)
    return fail(BSAM_E_ARG, "bsamgpu:arg: bad tag map or box buffer");
  if (min_width < 1) return fail(BSAM_E_ARG, "eb_options:minbox: MinBoxWidth must be a positive integer");
  const int n = bsam_cluster_impl(tags, allowed, n1, n2, min_width, fill_ratio, boxes, cap, ndropped);
  *nboxes = n;
  if (n > cap) return fail(BSAM_E_ARG, "bsamgpu:arg: %d boxes do not fit the buffer of %d", n, cap);
  return 0;
}

int bsam_cluster_tags_per_parent(const unsigned char* tags, const unsigned char* allowed, int n1, int n2,
                                 const int* parents, int np, int min_width, double fill_ratio, int* boxes, int cap,
                                 int* nboxes, int* ndropped) {
  if (!tags || !allowed || n1 < 1 || n2 < 1 || np < 0 || (np > 0 && !parents) || cap < 0 ||
      (cap > 0 && !boxes) || !nboxes)
    return fail(BSAM_E_ARG, "bsamgpu:arg: bad tag map, parent list or box buffer");
  if (min_width < 1) return fail(BSAM_E_ARG, "eb_options:minbox: MinBoxWidth must be a positive integer");
  int total = 0, dropped = 0;
  std::vector<unsigned char> tq, aq;
  std::vector<int> bq;
  for (int q = 0; q < np; ++q) {
    const int ilo = parents[4 * q], ihi = parents[4 * q + 1], jlo = parents[4 * q + 2], jhi = parents[4 * q + 3];
    if (ilo < 1 || ihi > n1 || jlo < 1 || jhi > n2 || ilo > ihi || jlo > jhi)
      return fail(BSAM_E_ARG, "bsamgpu:arg: parent box %d lies outside the %d x %d level", q + 1, n1, n2);
    const int m1 = ihi - ilo + 1, m2 = jhi - jlo + 1;
    // this parent's own tags and allowed region, in its local index space
    tq.assign((size_t)m1 * m2, 0);
    aq.assign((size_t)m1 * m2, 0);
    bool any = false;
    for (int j = 0; j < m2; ++j)
      for (int i = 0; i < m1; ++i) {
        const size_t g = (size_t)(ilo - 1 + i) + (size_t)(jlo - 1 + j) * n1, l = (size_t)i + (size_t)j * m1;
        tq[l] = tags[g] != 0;
        aq[l] = allowed[g] != 0;
        any = any || tq[l];
      }
    if (!any) continue;
    bq.assign(4 * tq.size(), 0);  // every box holds a tag: at most m1*m2 boxes
    int dq = 0;
    const int nq = bsam_cluster_impl(tq.data(), aq.data(), m1, m2, min_width, fill_ratio, bq.data(),
                                     (int)tq.size(), &dq);
    dropped += dq;
    for (int b = 0; b < nq; ++b, ++total)
      if (total < cap) {
        boxes[5 * total] = q + 1;
        boxes[5 * total + 1] = bq[4 * b] + ilo - 1;
        boxes[5 * total + 2] = bq[4 * b + 1] + ilo - 1;
        boxes[5 * total + 3] = bq[4 * b + 2] + jlo - 1;
        boxes[5 * total + 4] = bq[4 * b + 3] + jlo - 1;
      }
  }
  *nboxes = total;
  if (ndropped) *ndropped = dropped;
  if (total > cap) return fail(BSAM_E_ARG, "bsamgpu:arg: %d boxes do not fit the buffer of %d", total, cap);
  return 0;
}

long long bsam_level_field_size(const bsam_hier* h, int k, int field) {
  if (!h || k < 1 || k > h->nlev) return -1;
  const HLevel& L = h->H[k - 1];
  switch (field) {
    case BSAM_Q: return L.nq;
    case BSAM_F: case BSAM_FTRUE: return L.ncell;
    case BSAM_QC: return L.nqc;
    case BSAM_TAGMAP: return (long long)L.n[0] * L.n[1];
    default: return -1;
  }
}

static double* field_ptr(bsam_hier* h, int k, int field) {
  LevelDev<double>& D = h->d64[k - 1];
  switch (field) {
    case BSAM_Q: return D.Q;
    case BSAM_F: return D.F;
    case BSAM_FTRUE: return D.FT;
    case BSAM_QC: return D.QC;
    default: return nullptr;
  }
}

int bsam_get_level(bsam_hier* h, int k, int field, double* out) {
  if (int rc = check_level(h, k, 1)) return rc;
  const long long n = bsam_level_field_size(h, k, field);
  if (n < 0 || !out || !field_ptr(h, k, field)) return fail(BSAM_E_ARG, "bsamgpu:arg: bad field");
  if (n == 0) return 0;
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(out, field_ptr(h, k, field), n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int bsam_set_level(bsam_hier* h, int k, int field, const double* in) {
  if (int rc = check_level(h, k, 1)) return rc;
  const long long n = bsam_level_field_size(h, k, field);
  if (n < 0 || !in || !field_ptr(h, k, field)) return fail(BSAM_E_ARG, "bsamgpu:arg: bad field");
  if (n == 0) return 0;
  for (long long i = 0; i < n; ++i)
    if (!std::isfinite(in[i])) return fail(BSAM_E_NONFINITE, "eb_make_level:eval: non-finite input values.");
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(field_ptr(h, k, field), in, n * sizeof(double), hipMemcpyHostToDevice));
  if (field == BSAM_FTRUE && h->have32)
    launch_convert<float, double>(h->d32[k - 1].FT, h->d64[k - 1].FT, n, h->stream);
  return sync(h);
}

long long bsam_export_size(const bsam_hier* h, int k, int what) {
  if (!h || k < 1 || k > h->nlev) return -1;
  const HLevel& L = h->H[k - 1];
  switch (what) {
    case 0: return (long long)L.klass.size();
    case 1: return k >= 2 ? (long long)L.nqc : 0;
    case 2: return 2 * (long long)L.corr.size();
    case 3: return 1;
    case 4: return 1;
    case 5: {
      std::vector<long long> v;
      region_export(h, L, v);
      return (long long)v.size();
    }
    case 6: return k == 1 ? (long long)h->cst.M * h->cst.M : -1;
    default: return -1;
  }
}

int bsam_export(const bsam_hier* h, int k, int what, long long* out) {
  const long long n = bsam_export_size(h, k, what);
  if (n < 0 || !out) return fail(BSAM_E_ARG, "bsamgpu:arg: bad export request");
  const HLevel& L = h->H[k - 1];
  switch (what) {
    case 0:
      for (long long i = 0; i < n; ++i) out[i] = L.klass[i];
      break;
    case 1:
      for (long long i = 0; i < n; ++i) {
        const int q = L.wmap_patch[i];
        out[i] = q < 0 ? -1 : (long long)(L.wmap[i] - h->H[k - 2].pi[q].qoff) + 1000000LL * q;
      }
      break;
    case 2:
      for (size_t e = 0; e < L.corr.size(); ++e) {
        out[2 * e] = L.corr_target[e];
        out[2 * e + 1] = L.corr[e].face;
      }
      break;
    case 3:
      out[0] = face_records_arith(L) ? 1 : 0;
      break;
    case 4: {
      std::vector<SweepTile> st;
      bool even = true;
      for (const PatchInfo& P : L.pi) even = even && P.mx % 2 == 0;
      out[0] = (face_records_arith(L) && even && L.np > 0 && build_sweep_tiles(L, st)) ? (long long)st.size() : 0;
      break;
    }
    case 5: {
      std::vector<long long> v;
      region_export(h, L, v);
      for (long long i = 0; i < n; ++i) out[i] = v[(size_t)i];
      break;
    }
    case 6: {  // the coarsest inverse's bit patterns
      std::vector<double> a((size_t)n);
      if (h->ctx->device >= 0) {
        HIPCHK(hipSetDevice(h->ctx->device));
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemcpy(a.data(), h->cst.Ainv, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
      } else {
        if (h->ainv_host.size() != (size_t)n) return fail(BSAM_E_ARG, "bsamgpu:arg: no coarsest inverse");
        a = h->ainv_host;
      }
      std::memcpy(out, a.data(), (size_t)n * sizeof(double));
      break;
    }
  }
  return 0;
}

// Level 1's ghost records for k_coarsest's blocks (CoarsestGhost): every
// record must be physical (Dirichlet, Neumann, or a corner of two), with all
// its source cells among the rows of one 64-row block of the solve; else no
// table is built and the fill after the solve stays a launch of its own.
int build_coarsest_ghosts(bsam_hier* h) {
  h->cst.gtab = nullptr;
  h->cst.gofs = nullptr;
  const HLevel& L = h->H[0];
  if (L.np != 1 || L.ghost.empty()) return 0;
  const int n1 = L.n[0], n2 = L.n[1], ldq = n1 + 2, N = h->cst.N, B = 64;
  const int nb = (N + B - 1) / B;
  auto row = [&](int q) {
    const int ci = q % ldq - 1, cj = q / ldq - 1;
    return (q >= 0 && ci >= 0 && ci < n1 && cj >= 0 && cj < n2) ? ci + cj * n1 : -1;
  };
  auto side = [&](const GhostRec& s, int& a, int& b) {
    if (s.kind == G_DIR) { a = row(s.s[0]); b = row(s.s[1]); return a >= 0 && b >= 0; }
    if (s.kind == G_NEU) { a = row(s.s[0]); b = a; return a >= 0; }
    return false;
  };
  std::vector<std::vector<CoarsestGhost>> per((size_t)nb);
  const int ng = (int)L.ghost.size();
  for (int g = 0; g < ng; ++g) {
    const GhostRec& r = L.ghost[(size_t)g];
    CoarsestGhost e{};
    e.rec = g;
    int* w = e.rows;
    if (r.kind == G_PCORNER) {
      if (r.s[0] < 0 || r.s[0] >= ng || r.s[1] < 0 || r.s[1] >= ng) return 0;
      if (!side(L.ghost[(size_t)r.s[0]], w[0], w[1]) || !side(L.ghost[(size_t)r.s[1]], w[2], w[3])) return 0;
      w[4] = row(r.s[2]);
      if (w[4] < 0) return 0;
    } else {
      if (!side(r, w[0], w[1])) return 0;
      w[2] = w[3] = w[4] = w[0];
    }
    const int b = w[0] / B;
    for (int k = 0; k < 5; ++k)
      if (w[k] / B != b) return 0;
    per[(size_t)b].push_back(e);
  }
  std::vector<CoarsestGhost> tab;
  std::vector<int> ofs{0};
  for (const auto& v : per) {
    tab.insert(tab.end(), v.begin(), v.end());
    ofs.push_back((int)tab.size());
  }
  int rc;
  if ((rc = upload(h, &h->cst.gtab, tab))) return rc;
  return upload(h, &h->cst.gofs, ofs);
}

int bsam_hier_reset(bsam_hier* h) {
  if (int rc_ = need_dev(h)) return rc_;
  int mx = 0;
  for (const LevelDev<double>& D : h->d64) mx = std::max(mx, std::max(D.nq, std::max(D.ncell, D.nqc)));
  launch_reset_levels<double>(h->lv64, (const double* const*)h->q0dev, h->nlev, mx, h->stream);
  return sync(h);
}

// level k runs its relax calls as region launches under the current options
static bool region_level(const bsam_hier* h, int k) {
  if ((int)h->d64.size() < k || k < 2) return false;
  const LevelDev<double>& D = h->d64[k - 1];
  return h->opts.smoother == 0 && D.nregion > 0 && 2 * std::max(h->opts.npre, h->opts.npost) <= D.rdepth;
}

double bsam_smoother_bytes(const bsam_hier* h, int k, int elem_bytes) {
  if (!h || k < 1 || k > h->nlev) return -1;
  double b = 0;
  for (const PatchInfo& P : h->H[k - 1].pi) {
    const double mx = P.mx, my = P.my;
    b += (mx + 2) * (my + 2) + 3 * mx * my + (mx + 1) * my + mx * (my + 1);
  }
  // a fused sweep launch does the whole red+black sweep, a colour pass half;
  // a region launch (a whole relax call) reads and writes the level once
  const bool sweep = (int)h->d64.size() >= k && ((h->d64[k - 1].stream && h->opts.smoother != 3) ||
                                                  (h->opts.smoother == 2 && h->d64[k - 1].nsweep > 0) ||
                                                  region_level(h, k));
  return sweep ? b * elem_bytes : b * elem_bytes / 2.0;
}

int bsam_level_streams(const bsam_hier* h, int k) {
  if (!h || k < 1 || k > h->nlev || (int)h->d64.size() < k) return -1;
  return h->d64[k - 1].stream ? 1 : 0;
}

int bsam_time_smoother(bsam_hier* h, int k, int precision, int nlaunch, double* ms_per_launch) {
  if (int rc = check_level(h, k, 1)) return rc;
  if (int rc = check_stream_rows(h->opts.stream_rows)) return rc;
  if (nlaunch < 1 || !ms_per_launch) return fail(BSAM_E_ARG, "bsamgpu:arg: bad smoother timing request");
  if (precision == BSAM_MIXED)
    if (int rc = ensure32(h)) return rc;
  EventPair ev;
  HIPCHK(ev.create());
  hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
  HIPCHK(hipEventRecord(e0, h->stream));
  const bool stream = h->d64[k - 1].stream && h->opts.smoother != 3;
  const bool sweep = stream || (h->opts.smoother == 2 && h->d64[k - 1].nsweep > 0);
  const bool region = region_level(h, k);
  for (int i = 0; i < nlaunch; ++i) {
    if (region) {
      // one relax call of npre sweeps per launch, Q -> Qs, in the solve's
      // form: with the level's ghosts and the QC ring refresh when the fused
      // cycle swaps the level's buffers (pre- and post-smoothing both run)
      const int gh = h->opts.npre > 0 && h->opts.npost > 0 ? 1 : 0;
      if (precision == BSAM_MIXED)
        launch_relax_region<float>(h->d32[k - 1], k > 1 ? &h->d32[k - 2] : nullptr, h->opts.npre, 0,
                                   (float)h->opts.omega, 1, h->stream, nullptr, gh, nullptr, nullptr, gh);
      else
        launch_relax_region<double>(h->d64[k - 1], k > 1 ? &h->d64[k - 2] : nullptr, h->opts.npre, 0,
                                    h->opts.omega, 0, h->stream, nullptr, gh, nullptr, nullptr, gh);
      continue;
    }
    if (stream) {  // streaming sweeps, Q and Qs in turn
      const PatchInfo& P = h->H[k - 1].pi[0];
      if (precision == BSAM_MIXED) {
        LevelDev<float>& D = h->d32[k - 1];
        launch_sweep_stream<float>(D, P, (i & 1) ? D.Qs : D.Q, (i & 1) ? D.Q : D.Qs, (float)h->opts.omega, 1,
                                   h->ncu, h->opts.stream_rows, h->stream);
      } else {
        LevelDev<double>& D = h->d64[k - 1];
        launch_sweep_stream<double>(D, P, (i & 1) ? D.Qs : D.Q, (i & 1) ? D.Q : D.Qs, h->opts.omega, 0, h->ncu,
                                    h->opts.stream_rows, h->stream);
      }
      continue;
    }
    if (sweep) {
      // the form of every sweep after the first: a fused red+black sweep,
      // Q and Qs in turn
      if (precision == BSAM_MIXED) {
        LevelDev<float>& D = h->d32[k - 1];
        launch_relax_sweep<float>(D, k > 1 ? &h->d32[k - 2] : nullptr, (i & 1) ? D.Qs : D.Q, (i & 1) ? D.Q : D.Qs,
                                  kInlArith, 0, (float)h->opts.omega, 1, h->stream);
      } else {
        LevelDev<double>& D = h->d64[k - 1];
        launch_relax_sweep<double>(D, k > 1 ? &h->d64[k - 2] : nullptr, (i & 1) ? D.Qs : D.Q,
                                   (i & 1) ? D.Q : D.Qs, kInlArith, 0, h->opts.omega, 0, h->stream);
      }
      continue;
    }
    const int colour = 1 + (i & 1);
    // the form of the passes between the first and the last of a relax call
    const int mode = h->d64[k - 1].arith ? (kInlArith | kPush) : kInlRecords;
    if (precision == BSAM_MIXED)
      launch_relax<float>(h->d32[k - 1], k > 1 ? &h->d32[k - 2] : nullptr, colour, (float)h->opts.omega, 1, mode,
                          h->stream);
    else
      launch_relax<double>(h->d64[k - 1], k > 1 ? &h->d64[k - 2] : nullptr, colour, h->opts.omega, 0, mode,
                           h->stream);
  }
  HIPCHK(hipEventRecord(e1, h->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_launch = ms / nlaunch;
  return sync(h);
}

// ------------------------------------------------------------------ coefficients
// ebpSetTime.m:19-51: the caller re-evaluated CC/DeW/DnS/g at the new time
// (with the implicit shift) and ran the coefficient fill-down; everything the
// library derived from those values is refreshed here: the physical ghost
// data of the ghost records, the interface diffusivities of the reflux
// records (ebGatherParentCoeffs.m:47-61) and the coarsest factorisation.
// The boundary values a region tile's ghost records carry (the record's own,
// or a physical corner's two side records') from the level's current ghost
// records, found through the Q index each one writes.
static void refresh_region_ghosts(HLevel& L) {
  if (L.rghost.empty()) return;
  std::vector<int> rec_at((size_t)std::max(1, L.nq), -1);
  for (size_t r = 0; r < L.ghost.size(); ++r)
    if (L.ghost[r].dst >= 0 && L.ghost[r].dst < L.nq) rec_at[(size_t)L.ghost[r].dst] = (int)r;
  for (RegionGhost& g : L.rghost) {
    const GhostRec& r = L.ghost[(size_t)rec_at[(size_t)g.dst]];
    if ((g.kinds & 15) == G_PCORNER) {
      g.ca = L.ghost[(size_t)r.s[0]].c;
      g.cb = L.ghost[(size_t)r.s[1]].c;
    } else {
      g.c = r.c;
    }
  }
}

int bsam_update_coeffs(bsam_hier* h, const bsam_hier_desc* d) {
  if (!h) return fail(BSAM_E_ARG, "bsamgpu:arg: null hierarchy");
  if (!d || d->nlev != h->nlev || d->root_idx != h->root || !d->n || !d->np || !d->boxes || !d->bc_code ||
      !d->DeW || !d->DnS || !d->CC || !d->InvDiagS || !d->g)
    return fail(BSAM_E_ARG, "bsamgpu:arg: coefficient update needs the hierarchy's own descriptor");
  size_t pb = 0, oc = 0, ox = 0, oy = 0, og = 0;
  for (int k = 0; k < h->nlev; ++k) {
    const HLevel& L = h->H[k];
    if (d->n[2 * k] != L.n[0] || d->n[2 * k + 1] != L.n[1] || d->np[k] != L.np)
      return fail(BSAM_E_ARG, "bsamgpu:arg: coefficient update changes level %d's grid", k + 1);
    for (int p = 0; p < L.np; ++p, ++pb) {
      const Box& b = L.box[p];
      if (d->boxes[4 * pb] != b.ilo || d->boxes[4 * pb + 1] != b.ihi || d->boxes[4 * pb + 2] != b.jlo ||
          d->boxes[4 * pb + 3] != b.jhi)
        return fail(BSAM_E_ARG, "bsamgpu:arg: coefficient update changes a patch box of level %d", k + 1);
      for (int sd = 0; sd < 4; ++sd)
        if (d->bc_code[4 * pb + sd] != L.pi[p].bc[sd])
          return fail(BSAM_E_ARG, "bsamgpu:arg: coefficient update changes a boundary type of level %d", k + 1);
    }
    for (size_t i = 0; i < (size_t)L.nx; ++i) {
      const double v = d->DeW[ox + i];
      if (!std::isfinite(v)) return fail(BSAM_E_NONFINITE, "eb_make_level:eval: coefficient function returned non-finite values.");
      if (v <= 0) return fail(BSAM_E_D, "eb_make_level:D: D(x,y) must be strictly positive at faces.");
    }
    for (size_t i = 0; i < (size_t)L.ny; ++i) {
      const double v = d->DnS[oy + i];
      if (!std::isfinite(v)) return fail(BSAM_E_NONFINITE, "eb_make_level:eval: coefficient function returned non-finite values.");
      if (v <= 0) return fail(BSAM_E_D, "eb_make_level:D: D(x,y) must be strictly positive at faces.");
    }
    for (size_t i = 0; i < (size_t)L.ncell; ++i)
      if (!std::isfinite(d->CC[oc + i]) || !std::isfinite(d->InvDiagS[oc + i]))
        return fail(BSAM_E_NONFINITE, "eb_make_level:eval: coefficient function returned non-finite values.");
    for (size_t i = 0; i < (size_t)L.ng; ++i)
      if (!std::isfinite(d->g[og + i]))
        return fail(BSAM_E_NONFINITE, "eb_make_level:eval: coefficient function returned non-finite values.");
    oc += L.ncell; ox += L.nx; oy += L.ny; og += L.ng;
  }
  // accepted: replace the host copies and re-derive the dependent records
  oc = ox = oy = og = 0;
  for (int k = 0; k < h->nlev; ++k) {
    HLevel& L = h->H[k];
    L.CC.assign(d->CC + oc, d->CC + oc + L.ncell);
    L.ID.assign(d->InvDiagS + oc, d->InvDiagS + oc + L.ncell);
    L.DeW.assign(d->DeW + ox, d->DeW + ox + L.nx);
    L.DnS.assign(d->DnS + oy, d->DnS + oy + L.ny);
    L.G.assign(d->g + og, d->g + og + L.ng);
    oc += L.ncell; ox += L.nx; oy += L.ny; og += L.ng;
    for (const PatchInfo& P : L.pi) {  // physical sides: W(my) E(my) S(mx) N(mx) from gbase
      int first = P.gbase;
      for (int side = 0; side < 4; ++side) {
        const int len = side < 2 ? P.my : P.mx;
        if (P.bc[side] != 0)
          for (int t = 0; t < len; ++t) L.ghost[first + t] = phys_rec(L, P, side, t);
        first += len;
      }
    }
  }
  for (int k = 1; k < h->nlev; ++k) {
    const HLevel& C = h->H[k - 1];
    for (CorrRec& e : h->H[k].corr) e.cf = e.face < 2 ? C.DeW[e.cfidx] : C.DnS[e.cfidx];
  }
  for (HLevel& L : h->H) refresh_region_ghosts(L);
  bool allneu = true;
  for (int sd = 0; sd < 4; ++sd) allneu = allneu && h->H[0].pi[0].bc[sd] != 1;
  double cmax = 0;
  for (auto& L : h->H)
    for (double v : L.CC) cmax = std::max(cmax, std::fabs(v));
  const bool singular = allneu && cmax == 0;
  std::vector<double> Ainv, bcRhs;
  int M = 0;
  if (int rc = coarsest_assemble(h->H[0], singular, Ainv, bcRhs, M)) return rc;
  h->singular = singular;
  if (h->ctx->device < 0) {
    h->cst.M = M;
    h->ainv_host = std::move(Ainv);
    return 0;
  }
  HIPCHK(hipSetDevice(h->ctx->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  // captured cycles hold the coarsest order and buffers by value
  if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
  if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
  h->graph_key.clear();
  if (M != h->cst.M) {
    if (int rc = upload(h, &h->cst.Ainv, Ainv)) return rc;
    h->cst.M = M;
  } else {
    HIPCHK(hipMemcpy(h->cst.Ainv, Ainv.data(), Ainv.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(h->cst.bcRhs, bcRhs.data(), bcRhs.size() * sizeof(double), hipMemcpyHostToDevice));
  for (int k = 0; k < h->nlev; ++k) {
    const HLevel& L = h->H[k];
    LevelDev<double>& D = h->d64[k];
    auto put = [&](double* dst, const std::vector<double>& src) -> int {
      if (!src.empty()) HIPCHK(hipMemcpy(dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice));
      return 0;
    };
    int rc;
    if ((rc = put(D.CC, L.CC)) || (rc = put(D.ID, L.ID)) || (rc = put(D.DeW, L.DeW)) || (rc = put(D.DnS, L.DnS)) ||
        (rc = put(D.G, L.G)))
      return rc;
    if (!L.ghost.empty())
      HIPCHK(hipMemcpy(D.ghost, L.ghost.data(), L.ghost.size() * sizeof(GhostRec), hipMemcpyHostToDevice));
    if (!L.corr.empty())
      HIPCHK(hipMemcpy(D.corr, L.corr.data(), L.corr.size() * sizeof(CorrRec), hipMemcpyHostToDevice));
    if (!L.rghost.empty() && D.rghost)  // shared with the FP32 level
      HIPCHK(hipMemcpy(D.rghost, L.rghost.data(), L.rghost.size() * sizeof(RegionGhost), hipMemcpyHostToDevice));
    if (h->have32) {  // the FP32 levels share the tables; their value arrays are cast copies
      LevelDev<float>& E = h->d32[k];
      auto put32 = [&](float* dst, const std::vector<double>& src) -> int {
        if (src.empty()) return 0;
        std::vector<float> tmp(src.begin(), src.end());
        HIPCHK(hipMemcpy(dst, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice));
        return 0;
      };
      if ((rc = put32(E.CC, L.CC)) || (rc = put32(E.ID, L.ID)) || (rc = put32(E.DeW, L.DeW)) ||
          (rc = put32(E.DnS, L.DnS)) || (rc = put32(E.G, L.G)))
        return rc;
    }
  }
  return 0;
}

// ------------------------------------------------------------------ parabolic
// Time-stepping state of ebpSolve.m (H.par.Uo, H.par.Uoo) and the source
// f(t1) live on the device, packed like FTrue over all levels.
static int ensure_history(bsam_hier* h) {
  if (h->Uo) return 0;
  int rc;
  if ((rc = h->mem.alloc(&h->Uo, std::max(1LL, h->ncell_all)))) return rc;
  if ((rc = h->mem.alloc(&h->Uoo, std::max(1LL, h->ncell_all)))) return rc;
  return 0;
}

int bsam_parabolic_init(bsam_hier* h) {
  if (int rc_ = need_dev(h)) return rc_;
  if (int rc = ensure_history(h)) return rc;
  for (int k = 0; k < h->nlev; ++k)
    launch_history(h->d64[k], h->Uo + h->cell_base[k], h->Uoo + h->cell_base[k], 0, h->stream);
  h->have_uoo = false;
  return sync(h);
}

// ------------------------------------------------------------------ regrid
static bool same_level(const HLevel& a, const HLevel& b) {
  if (a.n[0] != b.n[0] || a.n[1] != b.n[1] || a.np != b.np) return false;
  for (int p = 0; p < a.np; ++p)
    if (a.box[p].ilo != b.box[p].ilo || a.box[p].ihi != b.box[p].ihi || a.box[p].jlo != b.box[p].jlo ||
        a.box[p].jhi != b.box[p].jhi)
      return false;
  return true;
}

int bsam_hier_inherit(bsam_hier* h, const bsam_hier* src) {
  if (int rc_ = need_dev(h)) return rc_;
  if (!src || src->ctx->device != h->ctx->device)
    return fail(BSAM_E_ARG, "bsamgpu:arg: source hierarchy missing or on another device");
  int same = 0;
  while (same < std::min(h->nlev, src->nlev) && same_level(h->H[same], src->H[same])) ++same;
  HIPCHK(hipSetDevice(h->ctx->device));
  HIPCHK(hipStreamSynchronize(src->stream));
  for (int k = 0; k < same; ++k)
    HIPCHK(hipMemcpyAsync(h->d64[k].Q, src->d64[k].Q, (size_t)h->H[k].nq * sizeof(double), hipMemcpyDeviceToDevice,
                          h->stream));
  if (src->Uo && same > 0) {
    if (int rc = ensure_history(h)) return rc;
    const size_t nb = (size_t)h->cell_base[same] * sizeof(double);
    HIPCHK(hipMemcpyAsync(h->Uo, src->Uo, nb, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->Uoo, src->Uoo, nb, hipMemcpyDeviceToDevice, h->stream));
    h->have_uoo = src->have_uoo;
  }
  return sync(h);
}

int bsam_regrid_level(bsam_hier* h, int k, const bsam_hier* old) {
  if (int rc = check_level(h, k, h->root)) return rc;
  if (!old || old->ctx->device != h->ctx->device)
    return fail(BSAM_E_ARG, "bsamgpu:arg: old hierarchy missing or on another device");
  const HLevel& L = h->H[k - 1];
  const bool conslin = k > h->root;
  const bool hist = old->Uo && old->have_uoo;
  if (conslin && hist && !h->Uo)
    return fail(BSAM_E_ARG, "bsamgpu:arg: the parent level's history has not been transferred");
  if (hist)
    if (int rc = ensure_history(h)) return rc;
  // cells an old patch of the same level covers (ebpRegrid.m:72-88)
  std::vector<int> oq(std::max(1, L.ncell), -1), oc(std::max(1, L.ncell), -1);
  if (k <= old->nlev) {
    const HLevel& O = old->H[k - 1];
    if (O.n[0] != L.n[0] || O.n[1] != L.n[1]) return fail(BSAM_E_ARG, "bsamgpu:arg: old level %d differs in size", k);
    for (int p = 0; p < L.np; ++p) {
      const Box& b = L.box[p];
      const PatchInfo& P = L.pi[p];
      for (int q = 0; q < O.np; ++q) {
        const Box& qb = O.box[q];
        const PatchInfo& Qp = O.pi[q];
        const int i0 = std::max(b.ilo, qb.ilo), i1 = std::min(b.ihi, qb.ihi);
        const int j0 = std::max(b.jlo, qb.jlo), j1 = std::min(b.jhi, qb.jhi);
        for (int j = j0; j <= j1; ++j)
          for (int i = i0; i <= i1; ++i) {
            const int c = P.coff + (i - b.ilo) + (j - b.jlo) * P.mx;
            oq[c] = Qp.qoff + (i - qb.ilo + 1) + (j - qb.jlo + 1) * (Qp.mx + 2);
            oc[c] = (int)old->cell_base[k - 1] + Qp.coff + (i - qb.ilo) + (j - qb.jlo) * Qp.mx;
          }
      }
    }
  }
  if (!conslin)
    for (int c = 0; c < L.ncell; ++c)
      if (oq[c] < 0) return fail(BSAM_E_ARG, "bsamgpu:arg: the old root does not cover the new root");
  // parent history index of every footprint cell (parentFootprint, ebpRegrid.m:140-151)
  std::vector<int> fc(1, 0);
  if (conslin && hist) {
    const HLevel& Pl = h->H[k - 2];
    fc.assign(std::max(1, L.ncell / 4), -1);
    for (int p = 0; p < L.np; ++p) {
      const PatchInfo& P = L.pi[p];
      const int cmx = P.mx / 2, cmy = P.my / 2, ldc = cmx + 2;
      for (int cj = 0; cj < cmy; ++cj)
        for (int ci = 0; ci < cmx; ++ci) {
          const int w = P.qcoff + (ci + 1) + (cj + 1) * ldc;
          const int q = L.wmap_patch[w];
          if (q < 0) return fail(BSAM_E_ARG, "bsamgpu:internal: footprint cell without a parent");
          const PatchInfo& Q = Pl.pi[q];
          const int loc = L.wmap[w] - Q.qoff, li = loc % (Q.mx + 2), lj = loc / (Q.mx + 2);
          if (li < 1 || li > Q.mx || lj < 1 || lj > Q.my)
            return fail(BSAM_E_ARG, "bsamgpu:internal: footprint cell in a parent ghost");
          fc[P.coff / 4 + ci + cj * cmx] = (int)h->cell_base[k - 2] + Q.coff + (li - 1) + (lj - 1) * Q.mx;
        }
    }
  }
  HIPCHK(hipSetDevice(h->ctx->device));
  HIPCHK(hipStreamSynchronize(old->stream));
  DevBufs<double> mem;
  int *doq, *doc, *dfc;
  if (int rc = mem.alloc(&doq, oq.size())) return rc;
  if (int rc = mem.alloc(&doc, oc.size())) return rc;
  if (int rc = mem.alloc(&dfc, fc.size())) return rc;
  HIPCHK(hipMemcpyAsync(doq, oq.data(), oq.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(doc, oc.data(), oc.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(dfc, fc.data(), fc.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
  launch_regrid_level(h->d64[k - 1], conslin ? h->d64[k - 2].Q : nullptr, conslin ? 1 : 0, hist ? 1 : 0,
                      hist ? h->Uoo + h->cell_base[k - 1] : nullptr, h->Uoo, dfc, doq, doc,
                      k <= old->nlev ? old->d64[k - 1].Q : nullptr, old->Uoo, h->stream);
  if (hist) h->have_uoo = true;
  return sync(h);
}

int bsam_parabolic_uo_from_q(bsam_hier* h) {
  if (int rc_ = need_dev(h)) return rc_;
  if (int rc = ensure_history(h)) return rc;
  for (int k = 0; k < h->nlev; ++k) {
    launch_history(h->d64[k], h->Uo + h->cell_base[k], h->Uoo + h->cell_base[k], 2, h->stream);
    if (k + 1 < h->root)
      launch_fill_value<double>(h->Uoo + h->cell_base[k], h->cell_base[k + 1] - h->cell_base[k], 0.0, h->stream);
  }
  return sync(h);
}

int bsam_parabolic_set_history(bsam_hier* h, const double* Uo, const double* Uoo) {
  if (int rc_ = need_dev(h)) return rc_;
  if (int rc = ensure_history(h)) return rc;
  const size_t nb = (size_t)h->ncell_all * sizeof(double);
  if (Uo) HIPCHK(hipMemcpyAsync(h->Uo, Uo, nb, hipMemcpyHostToDevice, h->stream));
  if (Uoo) {
    HIPCHK(hipMemcpyAsync(h->Uoo, Uoo, nb, hipMemcpyHostToDevice, h->stream));
    h->have_uoo = true;
  }
  return sync(h);
}

int bsam_parabolic_get_history(bsam_hier* h, int which, double* out) {
  if (int rc_ = need_dev(h)) return rc_;
  if (!h->Uo) return fail(BSAM_E_ARG, "bsamgpu:arg: history not initialised");
  if ((which != 0 && which != 1) || !out) return fail(BSAM_E_ARG, "bsamgpu:arg: bad history request");
  HIPCHK(hipMemcpyAsync(out, which ? h->Uoo : h->Uo, (size_t)h->ncell_all * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  return sync(h);
}

int bsam_parabolic_set_source(bsam_hier* h, const double* f) {
  if (int rc_ = need_dev(h)) return rc_;
  if (!f) {
    h->have_src = false;
    return 0;
  }
  for (long long i = 0; i < h->ncell_all; ++i)
    if (!std::isfinite(f[i])) return fail(BSAM_E_NONFINITE, "eb_make_level:eval: coefficient function returned non-finite values.");
  if (!h->src)
    if (int rc = h->mem.alloc(&h->src, std::max(1LL, h->ncell_all))) return rc;
  HIPCHK(hipMemcpyAsync(h->src, f, (size_t)h->ncell_all * sizeof(double), hipMemcpyHostToDevice, h->stream));
  h->have_src = true;
  return sync(h);
}

int bsam_parabolic_set_shift(bsam_hier* h, double aDt) {
  if (int rc_ = need_dev(h)) return rc_;
  if (!std::isfinite(aDt)) return fail(BSAM_E_ARG, "bsamgpu:arg: the shift must be finite");
  h->par_shift = aDt;
  return 0;
}

// cn (ebpRhs.m:17-22, 41-43, 52-66): fill-down and ghost fill of u^n, the
// pointwise part 2 f + (2/dt + aDt) Uo into FTrue, then the composite
// residual machinery turns it into FTrue - L(u^n) + reflux chunks in place:
// the QC refresh (ring, and the interior R(u) the fill-down has just written)
// and the per-cell expression and chunk order are those of ebCompResidual.m,
// which accumulates the chunks in the same (level, patch, face, entry) order
// as the cn branch.
static int enqueue_cn_rhs(bsam_hier* h, double dt) {
  if (!std::isfinite(h->par_shift))
    return fail(BSAM_E_ARG, "ebp_rhs:shift: cn needs the operator's shift (bsam_parabolic_set_shift)");
  if (!h->rout64) {
    std::vector<double*> ft(h->nlev);
    for (int k = 0; k < h->nlev; ++k) ft[k] = h->d64[k].FT;
    if (int rc = upload(h, &h->rout64, ft)) return rc;
  }
  Seq<double> q = seq64(h, h->stream);
  q.filldown();
  for (int k = 1; k <= h->nlev; ++k) q.fill(k);
  const double c = 2.0 / dt + h->par_shift;
  for (int k = h->root; k <= h->nlev; ++k) {
    const long long b = h->cell_base[k - 1];
    launch_cn_rhs(h->d64[k - 1].FT, h->have_src ? h->src + b : nullptr, h->Uo + b, c, h->H[k - 1].ncell, h->stream);
  }
  if (h->opts.mass_correction) launch_qc_refresh_all<double>(h->lv64, h->qc_items, h->nqc_items, h->stream);
  launch_residual_all<double, double>(h->lv64, h->nlev, h->res_tiles, h->nres_tiles, h->opts.mass_correction,
                                      h->parts, h->rout64, h->root, h->stream);
  HIPCHK(hipGetLastError());
  return 0;
}

static int enqueue_rhs(bsam_hier* h, double dt, int scheme) {
  if (!h->Uo) return fail(BSAM_E_ARG, "bsamgpu:arg: history not initialised");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(BSAM_E_ARG, "ebp_options:Dt: Dt must be finite and positive");
  if (scheme != BSAM_BE && scheme != BSAM_BDF2 && scheme != BSAM_CN)
    return fail(BSAM_E_ARG, "ebp_rhs:scheme: unknown scheme %d", scheme);
  if (scheme == BSAM_BDF2 && !h->have_uoo)
    return fail(BSAM_E_ARG, "ebp_rhs:history: bdf2 needs u^{n-1} (take a be step first)");
  if (scheme == BSAM_CN) return enqueue_cn_rhs(h, dt);
  // sub-root levels keep their FTrue: the FAS load overwrites them (ebpRhs.m:25-27)
  for (int k = h->root; k <= h->nlev; ++k) {
    const long long b = h->cell_base[k - 1];
    launch_bdf2_rhs(h->d64[k - 1].FT, h->have_src ? h->src + b : nullptr, h->Uo + b, h->Uoo + b, dt,
                    scheme == BSAM_BE, h->H[k - 1].ncell, h->stream);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int bsam_parabolic_rhs(bsam_hier* h, double dt, int scheme) {
  if (int rc_ = need_dev(h)) return rc_;
  if (int rc = enqueue_rhs(h, dt, scheme)) return rc;
  return sync(h);
}

int bsam_parabolic_advance(bsam_hier* h) {
  if (int rc_ = need_dev(h)) return rc_;
  if (!h->Uo) return fail(BSAM_E_ARG, "bsamgpu:arg: history not initialised");
  for (int k = 0; k < h->nlev; ++k)
    launch_history(h->d64[k], h->Uo + h->cell_base[k], h->Uoo + h->cell_base[k], 1, h->stream);
  h->have_uoo = true;
  return sync(h);
}

int bsam_parabolic_step(bsam_hier* h, double dt, int scheme, const bsam_solve_opts* o, bsam_stage_stats* out,
                        double* mass) {
  if (int rc_ = need_dev(h)) return rc_;
  if (int rc = enqueue_rhs(h, dt, scheme)) return rc;
  if (int rc = bsam_solve_cycles(h, o, out)) return rc;
  if (int rc = bsam_parabolic_advance(h)) return rc;
  if (mass) {
    double m = 0, a = 0;
    if (int rc = bsam_composite_mean(h, &m, &a)) return rc;
    *mass = m * a;  // compMass: mbar * vol (ebpSolve.m:103-106)
  }
  return 0;
}

// ------------------------------------------------------------------ batches
// Many independent hierarchies advanced by one launch: one workgroup per
// hierarchy runs its stage preamble or one whole cycle (same arithmetic as
// the hierarchy's own solve), so requests of any size and level count share
// the device without changing each other's results.
// Batch items run as groups of workgroups: an item gets one workgroup per
// kBatchCellsPerWg cells of its largest level, at most kBatchMaxWg and at
// most what the device holds resident at once.  The count depends on the
// hierarchy and the device alone, so an item's arithmetic (the order of its
// residual partial sums included) is the same in any batch.
constexpr long long kBatchCellsPerWg = 8192;
constexpr int kBatchMaxWg = 16;

struct bsam_batch {
  bsam_ctx* ctx = nullptr;
  int cap = 0;
  int ncu = 0;
  hipStream_t stream = nullptr;
  DevBufs<char> mem;
  BatchItem* d_items = nullptr;
  ResidualOut* d_out = nullptr;
  ResidualOut* d_parts = nullptr;  // kBatchMaxWg per item
  int* d_grp = nullptr;            // workgroup -> item, kBatchMaxWg per item
  unsigned* d_ctr = nullptr;       // one barrier counter per item
  int* d_err = nullptr;
  BatchItem* h_items = nullptr;  // pinned
  ResidualOut* h_out = nullptr;  // pinned
  int* h_grp = nullptr;          // pinned
  int* h_err = nullptr;          // pinned
  int bpc_m = -1, bpc = 0;       // resident workgroups per CU for coarsest order bpc_m
  ~bsam_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (h_items) (void)hipHostFree(h_items);
    if (h_out) (void)hipHostFree(h_out);
    if (h_grp) (void)hipHostFree(h_grp);
    if (h_err) (void)hipHostFree(h_err);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

int bsam_batch_create(bsam_ctx* ctx, int capacity, bsam_batch** out) {
  if (!ctx || !out || capacity < 1) return fail(BSAM_E_ARG, "bsamgpu:arg: bad batch request");
  if (ctx->device < 0) return fail(BSAM_E_ARG, "bsamgpu:arg: host-only context has no device state");
  HIPCHK(hipSetDevice(ctx->device));
  auto b = std::make_unique<bsam_batch>();
  b->ctx = ctx;
  b->cap = capacity;
  HIPCHK(hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
  int rc;
  const size_t nw = (size_t)capacity * kBatchMaxWg;
  if ((rc = b->mem.alloc(&b->d_items, capacity))) return rc;
  if ((rc = b->mem.alloc(&b->d_out, capacity))) return rc;
  if ((rc = b->mem.alloc(&b->d_parts, nw))) return rc;
  if ((rc = b->mem.alloc(&b->d_grp, nw))) return rc;
  if ((rc = b->mem.alloc(&b->d_ctr, capacity))) return rc;
  if ((rc = b->mem.alloc(&b->d_err, 1))) return rc;
  HIPCHK(hipHostMalloc((void**)&b->h_items, capacity * sizeof(BatchItem), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&b->h_out, capacity * sizeof(ResidualOut), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&b->h_grp, nw * sizeof(int), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&b->h_err, sizeof(int), hipHostMallocDefault));
  HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  *out = b.release();
  return 0;
}

void bsam_batch_destroy(bsam_batch* b) { delete b; }

int bsam_batch_step(bsam_batch* b, bsam_hier* const* hs, const int* phase, int n, const bsam_solve_opts* o,
                    bsam_residual* out) {
  if (!b || !hs || !phase || !o || !out || n < 0) return fail(BSAM_E_ARG, "bsamgpu:arg: bad batch step");
  if (n > b->cap) return fail(BSAM_E_ARG, "bsamgpu:arg: %d items exceed the batch capacity %d", n, b->cap);
  if (o->npre < 0 || o->npost < 0) return fail(BSAM_E_ARG, "bsamgpu:arg: negative sweep count");
  if (n == 0) return 0;
  HIPCHK(hipSetDevice(b->ctx->device));
  const bool mixed = o->precision == BSAM_MIXED;
  int max_m = 1;
  for (int i = 0; i < n; ++i) {
    bsam_hier* h = hs[i];
    if (!h || h->ctx->device != b->ctx->device)
      return fail(BSAM_E_ARG, "bsamgpu:arg: batch item %d is not a hierarchy on the batch's device", i);
    if (phase[i] != 0 && phase[i] != 1) return fail(BSAM_E_ARG, "bsamgpu:arg: batch phase must be 0 or 1");
    for (int j = 0; j < i; ++j)
      if (hs[j] == h) return fail(BSAM_E_ARG, "bsamgpu:arg: hierarchy listed twice in one batch step");
    if (mixed)
      if (int rc = ensure32(h)) return rc;
    max_m = std::max(max_m, h->cst.M);
  }
  // resident workgroups of the launch (every group's workgroups must run at once)
  if (b->bpc_m != max_m) {
    b->bpc = batch_cycle_blocks_per_cu(max_m);
    b->bpc_m = max_m;
  }
  const int cap_wg = b->bpc * b->ncu;
  if (cap_wg < 1) return fail(BSAM_E_HIP, "bsamgpu:hip: the batch kernel cannot be resident on this device");
  // groups, then launches of whole groups that fit the resident capacity
  std::vector<int> first_item{0}, first_wg{0};
  int wg = 0, wg_chunk = 0;
  for (int i = 0; i < n; ++i) {
    bsam_hier* h = hs[i];
    long long cells = 1;
    for (int k = 0; k < h->nlev; ++k) cells = std::max(cells, (long long)h->H[k].ncell);
    const int g = (int)std::min<long long>({(cells + kBatchCellsPerWg - 1) / kBatchCellsPerWg,
                                            (long long)kBatchMaxWg, (long long)cap_wg});
    if (wg_chunk + g > cap_wg) {
      first_item.push_back(i);
      first_wg.push_back(wg);
      wg_chunk = 0;
    }
    BatchItem& it = b->h_items[i];
    it.LV64 = h->lv64;
    it.LV32 = mixed ? h->lv32 : nullptr;
    it.rout32 = mixed ? h->rout32 : nullptr;
    it.out = b->d_out + i;
    it.parts = b->d_parts + (size_t)i * kBatchMaxWg;
    it.cst = h->cst;
    it.nlev = h->nlev;
    it.root = h->root;
    it.phase = phase[i];
    it.wg0 = wg_chunk;
    it.nwg = g;
    it.pad = 0;
    for (int w = 0; w < g; ++w) b->h_grp[wg + w] = i;
    wg += g;
    wg_chunk += g;
  }
  first_item.push_back(n);
  first_wg.push_back(wg);
  hipStream_t s = b->stream;
  HIPCHK(hipMemcpyAsync(b->d_items, b->h_items, n * sizeof(BatchItem), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(b->d_grp, b->h_grp, (size_t)wg * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(b->d_ctr, 0, n * sizeof(unsigned), s));
  HIPCHK(hipMemsetAsync(b->d_err, 0, sizeof(int), s));
  for (size_t c = 0; c + 1 < first_item.size(); ++c) {
    const int e = launch_batch_cycle(b->d_items, b->d_grp + first_wg[c], first_wg[c + 1] - first_wg[c], b->d_ctr,
                                     b->d_err, max_m, o->npre, o->npost, o->omega, o->mass_correction,
                                     mixed ? 1 : 0, std::max(1, o->inner_cycles), s);
    if (e != (int)hipSuccess)
      return fail(BSAM_E_HIP, "bsamgpu:hip: batch launch of %d workgroups failed: %s",
                  first_wg[c + 1] - first_wg[c], hipGetErrorString((hipError_t)e));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->h_out, b->d_out, n * sizeof(ResidualOut), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(b->h_err, b->d_err, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (*b->h_err)
    return fail(BSAM_E_HIP, "bsamgpu:hip: a batch group's barrier gave up waiting (workgroups not resident)");
  for (int i = 0; i < n; ++i) {
    const ResidualOut& r = b->h_out[i];
    out[i].linf = r.linf;
    out[i].l2 = std::sqrt(r.l2sq);
    out[i].fnorm = std::max(r.fnorm, 2.2250738585072014e-308);
    out[i].rel = residual_rel(r);
    out[i].ncell = r.ncell;
  }
  return 0;
}

}  // extern "C"
