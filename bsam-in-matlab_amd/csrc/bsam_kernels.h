// Launch wrappers of the V-cycle kernels (bsam_kernels.hip).
#pragma once
#include "bsam_internal.h"

namespace bsam {

// one colour pass over a level.  Mode bits:
//   ghost source (mode & 3): 0 the materialised ghost frame; 1 the ghost
//     records; 2 sibling ghosts from the frame and coarse-fine ghosts from the
//     cell's neighbours and the stored parent part (levels with L.arith);
//   kKeepCf: the other colour's face cells store the parent parts of their
//     coarse-fine records (and, in record mode, so do this colour's);
//   kPush: after the update, write the new value into every sibling ghost
//     slot that copies this cell.
constexpr int kInlRecords = 1, kInlArith = 2, kKeepCf = 4, kPush = 8;

template <typename T>
void launch_relax(const LevelDev<T>& L, const LevelDev<T>* parent, int colour, T omega, int homog, int mode,
                  hipStream_t s);
// one fused red+black sweep of every sweep tile of a level, Qin -> Qout
// (mode: the red half's ghost source, 0 / kInlRecords / kInlArith; keep: the
// parent parts are computed from the parent and stored)
template <typename T>
void launch_relax_sweep(const LevelDev<T>& L, const LevelDev<T>* parent, const T* Qin, T* Qout, int mode, int keep,
                        T omega, int homog, hipStream_t s);
// one red+black sweep of a one-patch level with physical boundaries on every
// side, Qin -> Qout: 64-lane waves on strips of kStreamCols written columns
// (two per lane, one halo lane at each end) and blocks of StreamRows<T> rows
// (the block height: rows_req when positive, else the largest power of two
// from kStreamRowsMin up to StreamRows<T> that keeps about 32 waves per CU in
// FP64, 8 in FP32)
constexpr int kStreamCols = 124, kStreamRowsMin = 8;
constexpr int kXcds = 8;  // MI355X: workgroups are dealt to the 8 XCDs in turn
template <typename T> struct StreamRows { static constexpr int value = sizeof(T) == 8 ? 64 : 32; };
template <typename T>
void launch_sweep_stream(const LevelDev<T>& L, const PatchInfo& P, const T* Qin, T* Qout, T omega, int homog,
                         int ncu, int rows_req, hipStream_t s);
// A whole relax call of nsweeps red+black sweeps on the level's region tiles
// (levels with L.nregion, 2 nsweeps <= L.rdepth), own cells into Qout (null:
// L.Qs); frame: pass 1 reads internal-face ghosts from the ghost frame; the
// parent parts in L.aux must be those of the current parent.
// ghosts = 0: launch_fill_copy (Qsrc = Qout) must follow, which copies the
// interior into Q and fills the ghosts from it.  ghosts = 1: the launch also
// writes every ghost of the level into Qout (the fill without its QC ring
// refresh), so Qout holds the level's complete new state; ring adds the QC
// ring refresh (from the parent); Rq (with Rdup) adds the restriction into
// the parent and QC's interior.
template <typename T>
void launch_relax_region(const LevelDev<T>& L, const LevelDev<T>* parent, int nsweeps, int frame, T omega, int homog,
                         hipStream_t s, T* Qout = nullptr, int ghosts = 0, T* Rq = nullptr, T* Rdup = nullptr,
                         int ring = 0, int corr = 0, const LevelDev<T>* parent_dev = nullptr, int fload = 0);
template <typename T>
void launch_fill_copy(const LevelDev<T>& L, const LevelDev<T>* parent, int homog, const T* Qsrc, hipStream_t s);
// the fill after a region relax (parent parts from aux, no QC ring refresh)
// and the restriction into the parent (also into dup when given), reading the
// new interior from Qsrc, in one launch
template <typename T>
void launch_fill_copy_restrict(const LevelDev<T>& L, const LevelDev<T>& parent, int homog, const T* Qsrc, T* dup,
                               hipStream_t s);
// aux parent parts of every coarse-fine record of the level from the current parent
template <typename T> void launch_refresh_cf(const LevelDev<T>& L, const LevelDev<T>* parent, hipStream_t s);
// all passes of one relax call, one patch per workgroup staged in LDS
constexpr long long kTiledLdsBytes = 64 * 1024;
template <typename T> bool tiled_fits(const LevelDev<T>& L);
template <typename T>
void launch_relax_tiled(const LevelDev<T>& L, const LevelDev<T>* parent, int nsweeps, T omega, int homog,
                        hipStream_t s);
template <typename T> void launch_fill(const LevelDev<T>& L, const LevelDev<T>* parent, int homog, hipStream_t s);
// tag indicator of every cell of a level (0: undivided Laplacian, 1: relative
// truncation error) against thr, stored with the level's coverage into
// level-index bitmaps; lo: per patch, the box's lower corner (0-based)
void launch_tag_cells(const LevelDev<double>& L, const int2* lo, int indicator, double thr, uint8_t* raw,
                      uint8_t* cover, hipStream_t s);
// signed distance to a closed polygon (seg5: ax, ay, ex, ey, len2 of its ns
// segments, back to back) and, where phi / delta are not null, the diffuse-
// domain fields (1 + tanh(d/eps))/2 and sech(d/eps)^2/(2 eps)
void launch_polygon_fields(const double* seg5, int ns, const double* x, const double* y, long long npts, double eps,
                           double* d, double* phi, double* delta, hipStream_t s);
// buffer (dilation by a (2b+1)^2 square) and proper-nesting clip of the raw tags
void launch_tag_finish(const uint8_t* raw, const uint8_t* cover, int n1, int n2, int buffer, uint8_t* out,
                       hipStream_t s);
template <typename T> void launch_ring(const LevelDev<T>& L, const LevelDev<T>& parent, hipStream_t s);
template <typename T>
void launch_ring_cf(const LevelDev<T>& L, const LevelDev<T>& parent, hipStream_t s);
// C (may be null): F's child, whose QC ring (and, with cf, coarse-fine parent
// parts) the launch also gathers from F, as the composite residual's refresh
template <typename T>
void launch_restrict_fill(const LevelDev<T>& F, const LevelDev<T>& P, const LevelDev<T>* PP, int homog,
                          hipStream_t s, const LevelDev<T>* C = nullptr, int cf = 0);
// restrict k -> k-1 + fill k-1 for k = kfrom..2 in one workgroup; step k
// also gathers level k+1's QC ring from level k when kqc <= k+1 <= ktop, and
// its parent parts when k+1 == cftop
template <typename T>
void launch_filldown_small(const LevelDev<T>* LV_dev, int kfrom, int kto, int homog, hipStream_t s, int kqc = 1 << 30,
                           int ktop = 0, int cftop = 0);
// dup (may be null): a second parent array that receives the restricted values too
template <typename T>
void launch_restrict(const LevelDev<T>& L, const LevelDev<T>& parent, int qc_only, hipStream_t s, T* dup = nullptr);
// ebCoarseLoad alone (QC ring as stored) and the fused ring + load + reflux
template <typename T> void launch_coarse_load(const LevelDev<T>& L, const LevelDev<T>& parent, int mass_corr, hipStream_t s);
// pghost: the parent's frame was not filled after the restriction; the load
// evaluates the parent's physical ghosts it reads (requires L.pgh_ok).  PPq
// (the parent's parent Q): the same launch also stores the parent's parent
// parts in its aux, as that fill would have
// from_ft: the reflux sums start from FTrue instead of F (loops that reset F
// once before their first cycle; see reflux_item)
template <typename T> void launch_load(const LevelDev<T>& L, const LevelDev<T>& parent, int mass_corr, hipStream_t s,
                                       int pghost = 0, int homog = 0, const T* PPq = nullptr, int from_ft = 0);
// prep != 0 (levels with L.arith): also store the parent parts of the face
// records and push corrected boundary values into sibling ghosts, for a
// following smoother pass that assembles its ghosts (kInlArith); dup (may be
// null): a second array of the level that receives the corrected values too
template <typename T>
void launch_correct(const LevelDev<T>& L, const LevelDev<T>& parent, int init, hipStream_t s, int prep = 0,
                    T* dup = nullptr);
template <typename T> void launch_coarsest(const CoarsestDev& C, const LevelDev<T>& L1, int homog, hipStream_t s);
// V-cycle of levels 1..kc in one workgroup (LV_dev: device array of levels,
// LV_host: the same descriptors on the host).  When the levels' fields and
// tables fit the workgroup's LDS they are staged there for the whole cycle.
template <typename T>
void launch_coarse_cycle(const LevelDev<T>* LV_dev, const LevelDev<T>* LV_host, int kc, const CoarsestDev& C,
                         int npre, int npost, T omega, int homog, int mass_corr, hipStream_t s, int from_ft = 0);

// LDS staging of the coarse cycle: the arrays a V-cycle of levels 1..kc reads
// or writes, in a fixed order shared by the host (sizing) and the kernel.
constexpr int kStageMaxLevels = 8;
constexpr long long kStageLdsLimit = 160LL * 1024 - 8 * 1024;
__host__ __device__ inline long long stage_align(long long b) { return (b + 15) & ~15LL; }
#define BSAM_STAGE_FIELDS(X, L, T)                                                                    \
  X(Q, T, (L).nq) X(F, T, (L).ncell) X(QC, T, (L).nqc) X(ID, T, (L).ncell) X(CC, T, (L).ncell)        \
  X(DeW, T, (L).nx) X(DnS, T, (L).ny) X(G, T, (L).ng) X(tiles, TileRec, (L).ntile)                   \
  X(ghost, GhostRec, (L).nghost) X(foot, FootRec, (L).nfoot) X(ring_all, RingRec, (L).nring_all)     \
  X(ring_cf, RingRec, (L).nring_cf) X(corr, CorrRec, (L).ncorr) X(wmap, int, (L).nqc)
// fine_ncorr: reflux entries of level k+1 (the size of level k's tgt_ent);
// negative when level k's targets are not used inside the cycle (top level)
template <typename T>
__host__ __device__ inline long long stage_level_bytes(const LevelDev<T>& L, int fine_ncorr) {
  long long b = 0;
#define BSAM_STAGE_SIZE(field, type, count) b += stage_align((long long)(count) * (long long)sizeof(type));
  BSAM_STAGE_FIELDS(BSAM_STAGE_SIZE, L, T)
#undef BSAM_STAGE_SIZE
  if (fine_ncorr >= 0)
    b += stage_align(4LL * L.ntgt) + stage_align(4LL * (L.ntgt + 1)) + stage_align(4LL * fine_ncorr);
  return b;
}
template <typename T> void launch_qc_refresh_all(const LevelDev<T>* LV_dev, const int2* items, int nitems, hipStream_t s);
// the composite residual writes min(ntl, kResidualParts) parts
constexpr int kResidualParts = 2048;
template <typename T, typename TO>
void launch_residual_all(const LevelDev<T>* LV_dev, int nlev, const int2* tl, int ntl, int mass_corr,
                         ResidualOut* parts, TO* const* rout, int k0, hipStream_t s);
// out_host (pinned host memory), when given, receives the result as well;
// seq: a device counter the launch increments and writes into
// out_host->pad2[0] after the result (the host polls it)
void launch_reduce(const ResidualOut* parts, int nparts, ResidualOut* out, hipStream_t s,
                   ResidualOut* out_host = nullptr, unsigned long long* seq = nullptr);
template <typename T> void launch_mass(const LevelDev<T>& L, ResidualOut* parts, hipStream_t s);
void launch_add_correction(const LevelDev<double>& L, const LevelDev<float>& E, hipStream_t s);
template <typename T> void launch_fill_value(T* a, long long n, T v, hipStream_t s);
// F = FTrue on every level; max_cells: the largest level's cell count
template <typename T> void launch_reset_force(const LevelDev<T>* LV_dev, int nlev, int max_cells, hipStream_t s);
// Q = Q0[k], F = FTrue, QC = 0 on every level; max_elems: the largest of any
// level's nq, ncell and nqc
template <typename T>
void launch_reset_levels(const LevelDev<T>* LV_dev, const T* const* Q0_dev, int nlev, int max_elems, hipStream_t s);
template <typename TD, typename TS> void launch_convert(TD* d, const TS* src, long long n, hipStream_t s);
void launch_bdf2_rhs(double* FT, const double* f, const double* Uo, const double* Uoo, double dt, int first,
                     long long n, hipStream_t s);
// FT = 2 f + c Uo (f may be null: 0), the pointwise part of the CN right side
void launch_cn_rhs(double* FT, const double* f, const double* Uo, double c, long long n, hipStream_t s);
// rotate 0: Uoo = Uo = interior(Q); 1: Uoo = Uo, Uo = interior(Q); 2: Uo = interior(Q)
void launch_history(const LevelDev<double>& L, double* Uo, double* Uoo, int rotate, hipStream_t s);
// one level's regrid transfer (Q interior and, with hist, its history)
void launch_regrid_level(const LevelDev<double>& L, const double* Pq, int conslin, int hist, double* Uoo_k,
                         const double* Uoo_all, const int* fc, const int* oq, const int* oc, const double* Qold,
                         const double* Uoo_old, hipStream_t s);
// a group of nwg_i workgroups (1024 threads each) per item: the item's
// preamble or whole cycle; grp_item maps each of the launch's nwg workgroups
// to its item, ctr holds one zeroed barrier counter per item and *err (zero)
// is set when a group barrier gave up waiting; max_m: largest coarsest order
// among the items (LDS right side).  A cooperative launch: returns its
// hipError_t (every workgroup must be resident at once)
int launch_batch_cycle(const BatchItem* items, const int* grp_item, int nwg, unsigned* ctr, int* err, int max_m,
                       int npre, int npost, double omega, int mass_corr, int mixed, int inner, hipStream_t s);
// resident k_batch_cycle workgroups per CU for a coarsest order max_m (0: unknown)
int batch_cycle_blocks_per_cu(int max_m);

}  // namespace bsam
