/* bsamgpu -- MI355X-native adaptive FAS multigrid V-cycle (C ABI).
 *
 * The library owns a device-resident copy of a block-structured patch
 * hierarchy and runs the solve path of the reference MATLAB code on it:
 * red-black Gauss-Seidel smoothing, QIF/LIF coarse-fine ghost fill,
 * restriction, ring gather, FAS coarse loading with flux-balance (reflux)
 * corrections, bilinear correction, the coarsest direct solve and the
 * composite residual.  Tagging, clustering, regridding and coefficient
 * evaluation stay with the caller (MATLAB through the MEX gateway, the plain
 * C driver, or the Python host layer).
 *
 * Conventions
 *  - Levels are numbered 1..nlev as in the reference (level 1 is the coarsest
 *    16x16-type grid, level root_idx the root, levels above it adaptive).
 *  - Patch boxes are 1-based inclusive [ilo ihi jlo jhi] in the level's
 *    global cell indices (EllipticBSAM/src/ebMakeLevel.m:1-4).
 *  - Arrays are FP64, column-major per patch (x index fastest), packed back to
 *    back in patch order and level order 1..nlev:
 *      Q        (mx+2)*(my+2)  padded solution
 *      FTrue    mx*my          source
 *      DeW      (mx+1)*my      D on x-faces
 *      DnS      mx*(my+1)      D on y-faces
 *      CC       mx*my          reaction
 *      InvDiagS mx*my          1/(diagonal incl. the physical-boundary fold)
 *      g        my+my+mx+mx    boundary data W,E,S,N (zeros on internal faces)
 *  - The caller owns every host buffer; each call copies what it needs.
 *  - Every function returns 0 on success or a negative BSAM_E_* code;
 *    bsam_last_error() returns the thread's last message, prefixed with the
 *    reference's error identifier (e.g. "eb_build_ops:nesting").
 *  - Calls on distinct hierarchies are independent (each has its own HIP
 *    stream); calls on one hierarchy must not overlap.
 */
#ifndef BSAMGPU_H
#define BSAMGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define BSAM_OK 0
#define BSAM_E_ARG (-1)
#define BSAM_E_ALIGN (-2)      /* eb_build_ops:align        ebBuildOps.m:38-42 */
#define BSAM_E_NESTING (-3)    /* eb_build_ops:nesting      ebBuildOps.m:138-142 */
#define BSAM_E_RING (-4)       /* eb_build_ops:ring         ebBuildOps.m:233-236 */
#define BSAM_E_OVERHANG (-5)   /* eb_build_ops:overhang     ebBuildOps.m:240-242 */
#define BSAM_E_CORR (-6)       /* eb_build_ops:corr         ebBuildOps.m:269-271 */
#define BSAM_E_COVER (-7)      /* eb_gather_parent_coeffs:cover  ebGatherParentCoeffs.m:53-56 */
#define BSAM_E_BOX (-8)        /* eb_make_level:box         ebMakeLevel.m:17-20 */
#define BSAM_E_D (-9)          /* eb_make_level:D           ebMakeLevel.m:43-45 */
#define BSAM_E_NONFINITE (-10) /* eb_make_level:eval        ebMakeLevel.m:99-104 */
#define BSAM_E_SINGULAR (-11)  /* coarsest matrix singular */
#define BSAM_E_HIP (-20)       /* HIP runtime failure */
#define BSAM_E_NOMEM (-21)

enum bsam_field { BSAM_Q = 0, BSAM_F = 1, BSAM_FTRUE = 2, BSAM_QC = 3, BSAM_TAGMAP = 4 /* size query only */ };
enum bsam_transfer_kind { BSAM_RESTRICT = 0, BSAM_RING = 1, BSAM_CORRECT = 2, BSAM_FILLDOWN = 3 };
enum bsam_precision { BSAM_FP64 = 0, BSAM_MIXED = 1 };

typedef struct bsam_ctx bsam_ctx;
typedef struct bsam_hier bsam_hier;

/* The hierarchy value H handed to ebSolveCycles (EllipticBSAM/src/ebSetup.m:10-68,
 * ebMakeLevel.m:78-143).  Descriptor tables (ghost maps, ring/window gathers,
 * scatters, reflux targets, covered masks, the coarsest factorisation) are
 * re-derived from the boxes inside the library. */
typedef struct {
  int nlev;
  int root_idx;
  double domain[4];
  const int *n;        /* [2*nlev]   cells per level (nx, ny) */
  const int *np;       /* [nlev]     patches per level */
  const int *boxes;    /* [4*sum np] */
  const int *bc_code;  /* [4*sum np] W E S N: 0 internal, 1 Dirichlet, 2 Neumann */
  const double *Q;
  const double *FTrue;
  const double *DeW;
  const double *DnS;
  const double *CC;
  const double *InvDiagS;
  const double *g;
} bsam_hier_desc;

/* Solver options (EllipticBSAM/src/ebOptions.m:44-66) plus device choices. */
typedef struct {
  int npre;            /* NPre  (3) */
  int npost;           /* NPost (3) */
  double omega;        /* Omega (1.0) */
  double rel_tol;      /* RelTol */
  int max_vcycles;     /* MaxVCycles */
  int ghost_interp;    /* 0 = 'qif', 1 = 'lif' */
  int mass_correction; /* MassCorrection (1) */
  int precision;       /* BSAM_FP64 or BSAM_MIXED (FP32 inner V-cycles, FP64 defect correction) */
  int inner_cycles;    /* FP32 V-cycles per defect-correction step (mixed only, >= 1) */
  int use_graph;       /* replay each cycle as a captured HIP graph */
  int coarse_cells;    /* levels 1..kc with at most this many cells each run their
                          V-cycle inside one workgroup (0: default 1100, -1: off) */
  int smoother;        /* 0: red-black Gauss-Seidel exactly as ebRelaxLevel, bit-identical
                          to the reference sequence: on levels with region tiles a
                          whole relax call in one launch (LDS canvas per tile with
                          the halo its passes depend on, coefficients in registers,
                          a barrier between colour passes), elsewhere one launch per
                          colour;
                          1: LDS-tiled -- every colour pass of a relax call in one
                          launch, one patch per workgroup, sibling ghosts frozen
                          for the call (same converged solution);
                          2: as 0, bit-identical, but on levels whose patches split
                          into at most 3 x CU-count tiles of <= 32x32 cells, each
                          pair of red+black sweeps runs as fused sweep launches
                          (red half in LDS, black half from it), alternating Q
                          and a second copy of Q;
                          3: colour passes on every level;
                          4: as 0 without region launches (colour passes).
                          Under 0, 2 and 4, a level that is one patch of even width
                          with physical sides only and >= 2^20 cells runs pairs of
                          sweeps as streaming sweep launches (one pass over the
                          rows per red+black sweep), also bit-identical */
  int stream_rows;     /* row-block height of a streaming sweep launch (one 64-lane
                          wave per block of 124 columns x stream_rows rows):
                          0 = automatic (about 32 waves per CU in FP64 and 8 in
                          FP32, at most 64 and 32 rows), else 8, 16, 32 or 64;
                          the result is the same for every value */
} bsam_solve_opts;

/* One solve stage (EllipticBSAM/src/ebSolveCycles.m:58-70) plus device timing. */
typedef struct {
  int amr_levels;
  int nlev;
  int vcycles;
  double final_rel;
  long long visible_cells;
  long long total_cells;
  double avg_factor;
  double *res_hist;    /* caller buffer for |r|_inf per cycle (may be NULL) */
  int res_hist_cap;    /* its capacity; res_hist_len <= cap entries are written */
  int res_hist_len;
  double gpu_ms;       /* device time of the cycle loop (HIP events) */
} bsam_stage_stats;

/* EllipticBSAM/src/ebCompResidual.m:80-84 */
typedef struct {
  double linf, l2, fnorm, rel;
  long long ncell;
} bsam_residual;

void bsam_default_opts(bsam_solve_opts *o);
const char *bsam_last_error(void);
const char *bsam_version(void);

int bsam_ctx_create(int hip_device, bsam_ctx **out);
void bsam_ctx_destroy(bsam_ctx *ctx);

/* Upload H and derive its tables: ebBuildOps.m:17-199, ebGatherParentCoeffs.m:12-63,
 * ebCoarsest.m:14-77 ('assemble'). */
int bsam_hier_create(bsam_ctx *ctx, const bsam_hier_desc *desc, const bsam_solve_opts *opts,
                     bsam_hier **out);
void bsam_hier_destroy(bsam_hier *h);

/* == H = ebSolveCycles(H)   EllipticBSAM/src/ebSolveCycles.m:1-80 */
int bsam_solve_cycles(bsam_hier *h, const bsam_solve_opts *opts, bsam_stage_stats *out);
/* == [res,H] = ebCompResidual(H)   EllipticBSAM/src/ebCompResidual.m:1-85 */
int bsam_comp_residual(bsam_hier *h, bsam_residual *out);
/* == H = ebVcycle(H, nlev)   EllipticBSAM/src/ebVcycle.m:1-36 */
int bsam_vcycle(bsam_hier *h);
/* == H = ebFillGhosts(H, k)   EllipticBSAM/src/ebFillGhosts.m:1-95 */
int bsam_fill_ghosts(bsam_hier *h, int k);
/* == H = ebRelaxLevel(H, k, nsweeps)   EllipticBSAM/src/ebRelaxLevel.m:1-39 */
int bsam_relax_level(bsam_hier *h, int k, int nsweeps);
/* == H = ebTransfer(H, k, what)   EllipticBSAM/src/ebTransfer.m:1-84 */
int bsam_transfer(bsam_hier *h, int k, int what);
/* == H = ebCoarseLoad(H, k)   EllipticBSAM/src/ebCoarseLoad.m:1-47 */
int bsam_coarse_load(bsam_hier *h, int k);
/* == H = ebCoarsest(H, 'solve')   EllipticBSAM/src/ebCoarsest.m:79-91 */
int bsam_coarsest_solve(bsam_hier *h);
/* F = FTrue on every level (ebSolveCycles.m:30-34) */
int bsam_reset_force(bsam_hier *h);
/* ebAddLevel.m:25-32: fill level k-1 ghosts, prolong its window into level k, fill k */
int bsam_init_level_prolong(bsam_hier *h, int k);
/* == [m, area] = ebCompositeMean(H)   EllipticBSAM/src/ebCompositeMean.m:1-24 */
int bsam_composite_mean(bsam_hier *h, double *mean, double *area);
/* Tagging of level k on its device-resident solution: the indicator, buffer
   and proper-nesting clip of ebTagCluster (EllipticBSAM/src/ebTagCluster.m:27-90;
   'rte': EllipticBSAM/src/ebRte.m:14-42).  indicator: BSAM_TAG_ULAP or
   BSAM_TAG_RTE; a cell is tagged when its indicator exceeds threshold; tags
   are dilated by `buffer` cells ((2b+1)^2 square, b <= 0: none) and kept only
   where the cell and its in-domain 8-neighbours are covered by level k.
   tags / raw (either may be NULL) receive n1*n2 bytes (0 or 1), column-major
   in the level's index space: the final tags, and the indicator tags before
   buffer and clip (bsam_level_field_size(h, k, BSAM_TAGMAP) = n1*n2).
   *ntag (may be NULL): number of final tags.  Clustering
   of the tags into boxes stays with the caller.  'rte' needs even patch
   widths (BSAM_E_ARG otherwise). */
#define BSAM_TAG_ULAP 0
#define BSAM_TAG_RTE 1
int bsam_tag_level(bsam_hier *h, int k, int indicator, double threshold, int buffer, unsigned char *tags,
                   unsigned char *raw, long long *ntag);

/* bsam_tag_level without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as bsam_tag_level, except that the tag
 * maps and the patch-corner table live in one library-owned device array of
 * 12.125 MiB per device (held until the process ends): a level map may have
 * at most 2^22 cells (n1*n2, e.g. 2048 x 2048) and a level at most 16384
 * patches, else BSAM_E_ARG.  Every call on a device shares that device's
 * array, so calls that overlap on one device (from several threads or
 * streams) race; each call clears what must start at zero. */
int bsam_tag_level_static(bsam_hier *h, int k, int indicator, double threshold, int buffer, unsigned char *tags,
                          unsigned char *raw, long long *ntag);

/* Clustering of a tag map into patch boxes (EllipticBSAM/src/ebTagCluster.m:92-380):
   Berger-Rigoutsos signature splitting, bisection of boxes that leave the
   allowed (proper-nesting) region, growth to min_width (MinBoxWidth) and the
   merge or drop of slivers thinner than 2 cells, with MATLAB's tie-breaking.
   tags, allowed: n1*n2 bytes, column-major (bsam_tag_level's output and the
   level's erode8 coverage).  fill_ratio: MinFillRatio.  boxes: room for cap
   boxes of 4 ints, 1-based inclusive [ilo ihi jlo jhi] in the level's index
   space; *nboxes: the number of boxes (BSAM_E_ARG when above cap);
   *ndropped (may be NULL): slivers dropped.  Host code: needs no device. */
int bsam_cluster_tags(const unsigned char *tags, const unsigned char *allowed, int n1, int n2, int min_width,
                      double fill_ratio, int *boxes, int cap, int *nboxes, int *ndropped);
/* The tree variant's clustering (EllipticBSAMTree/src/ebtTagCluster.m:93-128):
 * the same steps as bsam_cluster_tags, run inside every parent patch on that
 * parent's own tags and allowed region, so every box lies in exactly one
 * parent.  tags/allowed: the level-wide maps (n1 x n2, column-major, after the
 * global buffer and nesting clip); parents: np boxes [ilo ihi jlo jhi] (1-based,
 * level index space).  Writes min(count, cap) rows [parent ilo ihi jlo jhi]
 * (parent 1-based, box in level coordinates), parent by parent; *nboxes =
 * count, *ndropped = slivers dropped over all parents. */
int bsam_cluster_tags_per_parent(const unsigned char *tags, const unsigned char *allowed, int n1, int n2,
                                 const int *parents, int np, int min_width, double fill_ratio, int *boxes,
                                 int cap, int *nboxes, int *ndropped);

/* Native hierarchy construction (host code, needs no device): the sub-root
 * ladder and root of EllipticBSAM/src/ebSetup.m:10-68, level sampling
 * (ebMakeLevel.m:1-105: D at the x- and y-face midpoints, C and f at the cell
 * centres, the boundary data at the face midpoints of the domain sides), the
 * coefficient fill-down (ebCoeffFilldown.m:1-62), the folded inverse
 * diagonal (ebKernels.m:28-34, ebBcFold.m:12-42), the source projection of
 * a singular all-Neumann, C = 0 problem (ebCompatProject.m:1-111) and the
 * refinement of parent-space boxes into a new finest level
 * (ebAddLevel.m:10-24).  bsam_build_desc gives the hierarchy value that
 * bsam_hier_create takes (arrays owned by the builder, valid until the next
 * call on it); the new level's prolonged initial guess is
 * bsam_init_level_prolong on the device hierarchy (ebAddLevel.m:25-32), after
 * bsam_hier_inherit from the previous one.  The coefficient callbacks receive
 * n points (x[i], y[i]) and write n values; non-finite values give
 * BSAM_E_NONFINITE, D <= 0 gives BSAM_E_D, a bad box BSAM_E_BOX.  Only the
 * midpoint source (SourceQuad 1) is built natively. */
typedef void (*bsam_field_fn)(void *user, const double *x, const double *y, long long n, double *out);
typedef struct {
  double domain[4];    /* [xmin xmax ymin ymax] */
  bsam_field_fn D, C, f;
  int bc_type[4];      /* W E S N: 1 Dirichlet, 2 Neumann */
  bsam_field_fn g[4];  /* boundary data per side (NULL: 0) */
  bsam_field_fn u0;    /* InitialGuess at the cell centres of the ladder and root (NULL: 0) */
  void *user;          /* passed to every callback */
} bsam_problem;
typedef struct {
  int base_cells[2];   /* BaseCells */
  int coarsest_cells;  /* CoarsestCells (16) */
  int source_quad;     /* SourceQuad: 1 (midpoint) */
} bsam_build_opts;
typedef struct bsam_build bsam_build;
int bsam_build_create(const bsam_problem *prob, const bsam_build_opts *opts, bsam_build **out);
void bsam_build_destroy(bsam_build *b);
/* boxes: nboxes rows [ilo ihi jlo jhi] in the finest level's index space (1-based) */
int bsam_build_add_level(bsam_build *b, const int *boxes, int nboxes);
int bsam_build_desc(bsam_build *b, bsam_hier_desc *out);
int bsam_build_info(const bsam_build *b, int *nlev, int *root_idx, int *singular);
/* the finest level's allowed map for clustering (erode8 of its coverage,
 * EllipticBSAM/src/ebTagCluster.m:130-141): n1*n2 bytes, column-major */
int bsam_build_allowed(const bsam_build *b, unsigned char *out);

/* Diffuse-domain geometry on device `device`: for npts points (x, y), the
   signed distance d (positive inside) to the closed, radially ordered polygon
   of nv vertices (vx, vy; the last repeats the first), as
   geom.signedDistance of DiffuseDomainBSAM/src/ddStarfishGeometry.m:35-72, and
   where phi / delta are not NULL the transmission fields
   phi = (1 + tanh(d/eps))/2 and delta = sech(d/eps)^2/(2 eps) of
   DiffuseDomainBSAM/src/ddTransmissionProblem.m:14-15.  Host buffers; no
   hierarchy is needed. */
int bsam_polygon_fields(int device, const double *vx, const double *vy, int nv, const double *x, const double *y,
                        long long npts, double eps, double *d, double *phi, double *delta);

/* Restore the uploaded initial guess on every level (Q), F = FTrue, QC = 0:
 * restarts a solve from the descriptor's state without host traffic. */
int bsam_hier_reset(bsam_hier *h);

/* Measurement helpers.  bsam_time_smoother launches `nlaunch` smoother
 * launches of the form the solve uses on level k (fused red+black sweeps on
 * levels that have sweep tiles, else single-colour passes alternating
 * red/black), with no ghost refresh, on the hierarchy's stream and returns
 * the average launch time from HIP events; precision BSAM_MIXED times the
 * FP32 kernel.  bsam_smoother_bytes returns the algorithmic bytes of one such
 * launch:  B = b*sum_patches[(mx+2)(my+2) + 3 mx my + (mx+1) my + mx (my+1)]
 * (read Q, F, InvDiagS, DeW, DnS, write Q per red+black sweep) for a sweep,
 * B/2 for a colour pass. */
int bsam_time_smoother(bsam_hier *h, int k, int precision, int nlaunch, double *ms_per_launch);
double bsam_smoother_bytes(const bsam_hier *h, int k, int elem_bytes);
/* 1 when level k runs its sweeps as streaming sweep launches under smoothers
 * 0 and 2 (one patch of even width, physical sides only, >= 2^20 cells), so
 * that the stream_rows option applies to it; 0 otherwise; -1 on a bad level.
 * For option searches (no counterpart in the reference). */
int bsam_level_streams(const bsam_hier *h, int k);

/* Packed per-level field access (layout as in the descriptor; QC is
 * (mx/2+2)*(my/2+2) per patch, levels >= 2). */
long long bsam_level_field_size(const bsam_hier *h, int k, int field);
int bsam_get_level(bsam_hier *h, int k, int field, double *out);
int bsam_set_level(bsam_hier *h, int k, int field, const double *in);

/* Table export for parity checks against the reference descriptors.
 * what = 0: ghost classes per patch [W(my) E(my) S(mx) N(mx) SW SE NW NE]
 *           (1 physical, 2 same-level, 3 coarse-fine: ebBuildOps.m:47-113 land/cornerClass)
 * what = 1: per QC cell of level k the parent Q element index (patch-local
 *           padded linear index + 1e6*parent patch), ebBuildOps.m:115-142
 * what = 2: reflux entries of level k in order: (target parent cell index
 *           in the packed level-(k-1) cell layout, face 0..3) pairs,
 *           ebBuildOps.m:158-166
 * what = 3: one value, 1 when every face ghost record of level k has the
 *           standard interpolation stencil and sibling push target, so that
 *           smoother passes after the first assemble ghosts from neighbours
 * what = 4: one value, the number of fused red+black sweep tiles of level k
 *           (0: the level's relax calls run single-colour passes)
 * what = 5: the region tiles of level k (whole relax calls per launch):
 *           [ntile, passes covered, then per tile: canvas x0 y0 width height,
 *           patches np, depth-0 cells, colour-1 cells, colour-2 cells,
 *           (patch ox oy) x np, the cell entries: patch slot (bits 0-3) |
 *           i (4-15) | j (16-27) | depth (28-31)]; [0, 0] when none
 * what = 6: k = 1 only: the bit patterns of the coarsest inverse (M x M,
 *           column-major; the device copy on a device context) */
long long bsam_export_size(const bsam_hier *h, int k, int what);
int bsam_export(const bsam_hier *h, int k, int what, long long *out);

/* == the refresh of ebpSetTime (ParabolicBSAM/src/ebpSetTime.m:19-51): replace
 * CC, DeW, DnS, InvDiagS and g on every level and re-derive what depends on
 * them (physical ghost data, the reflux interface diffusivities of
 * ebGatherParentCoeffs.m:47-61, the coarsest factorisation of ebCoarsest.m:14-77).
 * The caller evaluated the coefficients at the new time with the implicit
 * shift C + alpha/dt and ran the coefficient fill-down (ebCoeffFilldown.m:1-62).
 * desc must repeat the hierarchy's n, np, boxes and bc codes; its Q and FTrue
 * are not read.  Nothing is changed when an argument is rejected. */
int bsam_update_coeffs(bsam_hier *h, const bsam_hier_desc *desc);

/* Device state of a hierarchy rebuilt with more levels (ebAddLevel.m:10-40
   keeps the existing levels): for every level k whose boxes equal those of
   src, Q (with ghosts) and the parabolic history Uo/Uoo are copied device to
   device from src.  Both on the same device. */
int bsam_hier_inherit(bsam_hier *h, const bsam_hier *src);
/* transferLevel of ParabolicBSAM/src/ebpRegrid.m:62-151 for level k of the new
   hierarchy h (k = root: the root copied from old, ebpRegrid.m:20-25):
   conservative-linear prolongation of the parent window into Q's interior
   and, when old carries Uoo, of the parent footprint's history into Uoo,
   both overwritten by old's level-k values wherever an old patch covers the
   cell.  Levels below k must already be transferred; ghosts are left to
   bsam_fill_ghosts. */
int bsam_regrid_level(bsam_hier *h, int k, const bsam_hier *old);
/* H.par.Uo = interiors(H) after a regrid (ebpRegrid.m:48-55); Uoo keeps the
   transferred history and is zero below the root. */
int bsam_parabolic_uo_from_q(bsam_hier *h);
/* Parabolic stepping (ParabolicBSAM/src/ebpSolve.m:23-78).  The history
 * H.par.Uo = u^n, H.par.Uoo = u^{n-1} and the source f(t1) are device-resident
 * arrays packed like FTrue over all levels 1..nlev. */
enum bsam_scheme { BSAM_BE = 0, BSAM_BDF2 = 1, BSAM_CN = 2 };
/* Uo = interior(Q), no u^{n-1} yet (ebpInit.m:32-35) */
int bsam_parabolic_init(bsam_hier *h);
/* Upload Uo and/or Uoo (NULL keeps the resident array); after a regrid
 * (ebpRegrid.m:80-91) */
int bsam_parabolic_set_history(bsam_hier *h, const double *Uo, const double *Uoo);
/* which = 0: Uo, 1: Uoo */
int bsam_parabolic_get_history(bsam_hier *h, int which, double *out);
/* f(x,y,t1) at cell centres, packed like FTrue (sub-root entries unused);
 * kept resident until replaced; NULL means f = 0 (ebpRhs.m:31-36) */
int bsam_parabolic_set_source(bsam_hier *h, const double *f);
/* The shift aDt = alpha/Dt that the resident CC carries (C_eff = C + aDt),
 * H.par.aDt of ebpSetTime.m:54; only the cn right side reads it. */
int bsam_parabolic_set_shift(bsam_hier *h, double aDt);
/* FTrue on levels >= root (ParabolicBSAM/src/ebpRhs.m:15-67):
 *   BSAM_BE:   f + Uo/dt
 *   BSAM_BDF2: f + (2 Uo - Uoo/2)/dt
 *   BSAM_CN:   2 f + (2/dt + aDt) Uo - L(u^n), with f taken at t1 - dt/2, L
 *              the operator in place (old coefficients: call before the
 *              coefficients move to t1) after a fill-down and ghost fill of
 *              u^n, plus the old state's coarse-fine flux corrections added
 *              to the parent cells when mass correction is on. */
int bsam_parabolic_rhs(bsam_hier *h, double dt, int scheme);
/* Uoo = Uo, Uo = interior(Q) after an accepted step (ebpSolve.m:55-57) */
int bsam_parabolic_advance(bsam_hier *h);
/* One accepted step of ebpSolve.m:40-58 without host traffic beyond the
 * per-cycle residual: rhs, ebSolveCycles (opts as in bsam_solve_cycles),
 * history rotation and compMass = composite mean * area (ebpSolve.m:103-106;
 * mass may be NULL).  be and bdf2 read no operator in the rhs, so the
 * coefficients of t1 are set before the call; cn reads the operator in place
 * for both halves, which is the step only when D, C and the boundary data do
 * not depend on t (else: bsam_parabolic_rhs, coefficients of t1,
 * bsam_solve_cycles, bsam_parabolic_advance). */
int bsam_parabolic_step(bsam_hier *h, double dt, int scheme, const bsam_solve_opts *opts,
                        bsam_stage_stats *out, double *mass);

/* Batched stage steps for serving many independent hierarchies
 * (no counterpart in the reference, which solves one hierarchy at a time).
 * One cooperative launch advances n hierarchies of the batch's device; each
 * runs on its own group of workgroups (one per 8192 cells of its largest
 * level, at most 16; groups that do not fit the device's resident capacity
 * go to a following launch) with exactly the arithmetic of its own
 * bsam_solve_cycles, so results do not depend on what else is in the batch.
 * A group barrier that waits too long fails the call with BSAM_E_HIP instead
 * of hanging.  phase[i] = 0 runs the
 * stage preamble of ebSolveCycles.m:11-19 (fill every level, baseline
 * residual), 1 runs one cycle of ebSolveCycles.m:28-41 (reset F, V-cycle,
 * filldown, residual; with opts->precision == BSAM_MIXED the FP32-inner
 * defect-correction step).  out[i] receives item i's composite residual; the
 * caller applies the stopping rule (rel <= RelTol or the cycle cap).
 * opts->coarse_cells, smoother and use_graph are not used. */
typedef struct bsam_batch bsam_batch;
int bsam_batch_create(bsam_ctx *ctx, int capacity, bsam_batch **out);
void bsam_batch_destroy(bsam_batch *b);
int bsam_batch_step(bsam_batch *b, bsam_hier *const *hs, const int *phase, int n,
                    const bsam_solve_opts *opts, bsam_residual *out);

#ifdef __cplusplus
}
#endif
#endif /* BSAMGPU_H */
