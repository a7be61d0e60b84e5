/* bsamgpu_mex -- MATLAB gateway to libbsamgpu.
 *
 * Lets the reference's MATLAB entry points (ebSolve, ebtSolve, ebpSolve,
 * ddSolve) hand their hierarchy value H to the device library at the
 * ebSolveCycles seam (EllipticBSAM/src/ebSolve.m:38,41,52,
 * ParabolicBSAM/src/ebpSolve.m:50).  Build inside MATLAB with
 *
 *   mex -R2018a -I../../include bsamgpu_mex.c -L../csrc -lbsamgpu
 *
 * Commands (h is a uint64 handle):
 *   h = bsamgpu_mex('create', H [, device])      upload H.lev{k} (ebSetup.m:10-68 layout)
 *   st = bsamgpu_mex('solve', h, opts)           == ebSolveCycles; st as H.stats.stages{end}
 *   Q = bsamgpu_mex('get_q', h)                  Q{k}{p}, padded (mx+2) x (my+2)
 *   bsamgpu_mex('set_q', h, H)                   upload H.lev{k}.Q{p}
 *   bsamgpu_mex('update_coeffs', h, H)           refresh after ebpSetTime (CC/DeW/DnS/InvDiagS/bc)
 *   bsamgpu_mex('parabolic_init', h)             Uo = interior(Q)
 *   bsamgpu_mex('set_history', h, Uo, Uoo)       U{k}{p} mx x my ([] keeps)
 *   bsamgpu_mex('set_source', h, F)              F{k}{p} mx x my, [] for f = 0
 *   [st, mass] = bsamgpu_mex('parabolic_step', h, dt, scheme, opts)   scheme 'be' | 'bdf2' | 'cn'
 *   bsamgpu_mex('parabolic_rhs', h, dt, scheme [, aDt])   ebpRhs alone (cn: aDt = H.par.aDt,
 *                                                         before ebpSetTime moves to t1)
 *   mass = bsamgpu_mex('parabolic_advance', h)   history rotation and compMass after a 'solve'
 *   U = bsamgpu_mex('get_history', h, which)     0: Uo, 1: Uoo
 *   boxes = bsamgpu_mex('cluster_per_parent', T, allowed, parentBoxes, minw, fill)   tree variant
 *   bsamgpu_mex('destroy', h)
 *
 * Library errors are raised as MATLAB errors with identifier
 * "bsamgpu:<reference identifier>" (e.g. bsamgpu:eb_build_ops:align).
 */
#include <stdint.h>
#include <stdlib.h>
#include <ctype.h>
#include <string.h>

#include "mex.h"
#include "bsamgpu.h"

#define MAXDEV 16

/* 'be' | 'bdf2' | 'cn' -> enum bsam_scheme (-1: the library rejects it) */
static int scheme_code(const mxArray* a) {
  char sch[8] = {0};
  mxGetString(a, sch, sizeof sch);
  return !strcmp(sch, "be") ? BSAM_BE : !strcmp(sch, "bdf2") ? BSAM_BDF2 : !strcmp(sch, "cn") ? BSAM_CN : -1;
}
static bsam_ctx* g_ctx[MAXDEV];

static void at_exit(void) {
  for (int i = 0; i < MAXDEV; ++i)
    if (g_ctx[i]) { bsam_ctx_destroy(g_ctx[i]); g_ctx[i] = NULL; }
}

static void lib_error(void) {
  const char* m = bsam_last_error();
  char id[160] = "bsamgpu:";
  const char* sep = strstr(m, ": ");
  size_t n = sep ? (size_t)(sep - m) : 0;
  if (n > sizeof id - 9) n = sizeof id - 9;
  memcpy(id + 8, m, n);
  id[8 + n] = 0;
  if (!n) strcpy(id, "bsamgpu:error");
  mexErrMsgIdAndTxt(id, "%s", m);
}

static void chk(int rc) {
  if (rc) lib_error();
}

static const mxArray* field(const mxArray* s, const char* name) {
  const mxArray* f = mxGetField(s, 0, name);
  if (!f) mexErrMsgIdAndTxt("bsamgpu:arg", "missing field %s", name);
  return f;
}

static double scalar(const mxArray* s, const char* name) { return mxGetScalar(field(s, name)); }

static bsam_hier* handle(const mxArray* a) {
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("bsamgpu:arg", "expected a uint64 handle");
  uint64_t v = *(const uint64_t*)mxGetData(a);
  if (!v) mexErrMsgIdAndTxt("bsamgpu:arg", "null handle");
  return (bsam_hier*)(uintptr_t)v;
}

/* growable packed buffers of the descriptor */
typedef struct {
  double* a;
  size_t n, cap;
} dbuf;

static void push(dbuf* b, const double* v, size_t n) {
  if (b->n + n > b->cap) {
    size_t c = b->cap ? b->cap : 1024;
    while (c < b->n + n) c *= 2;
    b->a = (double*)mxRealloc(b->a, c * sizeof(double));
    b->cap = c;
  }
  if (n) memcpy(b->a + b->n, v, n * sizeof(double));
  b->n += n;
}

static void push_zeros(dbuf* b, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    double z = 0;
    push(b, &z, 1);
  }
}

/* per-patch array of a level field, checked against (m, n) */
static const double* patch_array(const mxArray* lev, const char* name, int p, size_t m, size_t n) {
  const mxArray* c = mxGetCell(field(lev, name), p);
  if (!c || !mxIsDouble(c) || mxIsComplex(c) || mxGetM(c) != m || mxGetN(c) != n)
    mexErrMsgIdAndTxt("bsamgpu:arg", "lev.%s{%d} must be a real %zu x %zu double array", name, p + 1, m, n);
  return mxGetPr(c);
}

typedef struct {
  bsam_hier_desc d;
  int *n, *np, *boxes, *bc;
  dbuf Q, FT, DeW, DnS, CC, ID, g;
} packed;

static void pack(const mxArray* H, packed* pk) {
  memset(pk, 0, sizeof *pk);
  const int nlev = (int)scalar(H, "nlev");
  const mxArray* levs = field(H, "lev");
  const mxArray* dom = field(field(H, "prob"), "domain");
  if (nlev < 1 || mxGetNumberOfElements(dom) != 4) mexErrMsgIdAndTxt("bsamgpu:arg", "bad hierarchy");
  pk->d.nlev = nlev;
  pk->d.root_idx = (int)scalar(H, "rootIdx");
  for (int i = 0; i < 4; ++i) pk->d.domain[i] = mxGetPr(dom)[i];
  pk->n = (int*)mxCalloc(2 * nlev, sizeof(int));
  pk->np = (int*)mxCalloc(nlev, sizeof(int));
  int P = 0;
  for (int k = 0; k < nlev; ++k) P += (int)scalar(mxGetCell(levs, k), "np");
  pk->boxes = (int*)mxCalloc(4 * (size_t)P, sizeof(int));
  pk->bc = (int*)mxCalloc(4 * (size_t)P, sizeof(int));
  int pb = 0;
  for (int k = 0; k < nlev; ++k) {
    const mxArray* lev = mxGetCell(levs, k);
    const double* n = mxGetPr(field(lev, "n"));
    const int np = (int)scalar(lev, "np");
    const mxArray* box = field(lev, "box");
    if (mxGetM(box) != (size_t)np || mxGetN(box) != 4) mexErrMsgIdAndTxt("bsamgpu:arg", "lev.box must be np x 4");
    pk->n[2 * k] = (int)n[0];
    pk->n[2 * k + 1] = (int)n[1];
    pk->np[k] = np;
    for (int p = 0; p < np; ++p, ++pb) {
      int b[4];
      for (int c = 0; c < 4; ++c) b[c] = pk->boxes[4 * pb + c] = (int)mxGetPr(box)[p + c * np];
      const size_t mx = (size_t)(b[1] - b[0] + 1), my = (size_t)(b[3] - b[2] + 1);
      push(&pk->Q, patch_array(lev, "Q", p, mx + 2, my + 2), (mx + 2) * (my + 2));
      push(&pk->FT, patch_array(lev, "FTrue", p, mx, my), mx * my);
      push(&pk->DeW, patch_array(lev, "DeW", p, mx + 1, my), (mx + 1) * my);
      push(&pk->DnS, patch_array(lev, "DnS", p, mx, my + 1), mx * (my + 1));
      push(&pk->CC, patch_array(lev, "CC", p, mx, my), mx * my);
      push(&pk->ID, patch_array(lev, "InvDiagS", p, mx, my), mx * my);
      const mxArray* bcs = mxGetCell(field(lev, "bc"), p);
      for (int s = 0; s < 4; ++s) {
        const size_t len = s < 2 ? my : mx;
        const mxArray* code = mxGetField(bcs, s, "code");
        const mxArray* g = mxGetField(bcs, s, "g");
        pk->bc[4 * pb + s] = code ? (int)mxGetScalar(code) : 0;
        if (pk->bc[4 * pb + s] != 0) {
          if (!g || mxGetNumberOfElements(g) != len)
            mexErrMsgIdAndTxt("bsamgpu:arg", "lev.bc{%d}(%d).g must have %zu entries", p + 1, s + 1, len);
          push(&pk->g, mxGetPr(g), len);
        } else {
          push_zeros(&pk->g, len);
        }
      }
    }
  }
  pk->d.n = pk->n; pk->d.np = pk->np; pk->d.boxes = pk->boxes; pk->d.bc_code = pk->bc;
  pk->d.Q = pk->Q.a; pk->d.FTrue = pk->FT.a; pk->d.DeW = pk->DeW.a; pk->d.DnS = pk->DnS.a;
  pk->d.CC = pk->CC.a; pk->d.InvDiagS = pk->ID.a; pk->d.g = pk->g.a;
}

static void unpack_free(packed* pk) {
  mxFree(pk->n); mxFree(pk->np); mxFree(pk->boxes); mxFree(pk->bc);
  mxFree(pk->Q.a); mxFree(pk->FT.a); mxFree(pk->DeW.a); mxFree(pk->DnS.a);
  mxFree(pk->CC.a); mxFree(pk->ID.a); mxFree(pk->g.a);
}

/* ebOptions fields -> bsam_solve_opts (GhostInterp is fixed at create) */
static void opts_from(const mxArray* o, bsam_solve_opts* so) {
  bsam_default_opts(so);
  if (!o || !mxIsStruct(o)) return;
  const mxArray* f;
  if ((f = mxGetField(o, 0, "NPre"))) so->npre = (int)mxGetScalar(f);
  if ((f = mxGetField(o, 0, "NPost"))) so->npost = (int)mxGetScalar(f);
  if ((f = mxGetField(o, 0, "Omega"))) so->omega = mxGetScalar(f);
  if ((f = mxGetField(o, 0, "RelTol"))) so->rel_tol = mxGetScalar(f);
  if ((f = mxGetField(o, 0, "MaxVCycles"))) so->max_vcycles = (int)mxGetScalar(f);
  if ((f = mxGetField(o, 0, "MassCorrection"))) so->mass_correction = mxGetScalar(f) != 0;
  if ((f = mxGetField(o, 0, "StreamRows"))) so->stream_rows = (int)mxGetScalar(f);
  if ((f = mxGetField(o, 0, "GhostInterp")) && mxIsChar(f)) {
    char s[8] = {0};
    mxGetString(f, s, sizeof s);
    so->ghost_interp = !strcmp(s, "lif") ? 1 : 0;
  }
  if ((f = mxGetField(o, 0, "Precision")) && mxIsChar(f)) {
    char s[8] = {0};
    mxGetString(f, s, sizeof s);
    so->precision = !strcmp(s, "mixed") ? BSAM_MIXED : BSAM_FP64;
  }
}

static mxArray* stage_struct(const bsam_stage_stats* st, const double* hist) {
  static const char* names[] = {"amrLevels", "nlev", "vcycles", "resHist", "finalRel",
                                "visibleCells", "totalCells", "avgFactor", "gpuMs"};
  mxArray* s = mxCreateStructMatrix(1, 1, 9, names);
  mxSetField(s, 0, "amrLevels", mxCreateDoubleScalar(st->amr_levels));
  mxSetField(s, 0, "nlev", mxCreateDoubleScalar(st->nlev));
  mxSetField(s, 0, "vcycles", mxCreateDoubleScalar(st->vcycles));
  mxArray* h = mxCreateDoubleMatrix(1, (size_t)st->res_hist_len, mxREAL);
  if (st->res_hist_len) memcpy(mxGetPr(h), hist, (size_t)st->res_hist_len * sizeof(double));
  mxSetField(s, 0, "resHist", h);
  mxSetField(s, 0, "finalRel", mxCreateDoubleScalar(st->final_rel));
  mxSetField(s, 0, "visibleCells", mxCreateDoubleScalar((double)st->visible_cells));
  mxSetField(s, 0, "totalCells", mxCreateDoubleScalar((double)st->total_cells));
  mxSetField(s, 0, "avgFactor", mxCreateDoubleScalar(st->avg_factor));
  mxSetField(s, 0, "gpuMs", mxCreateDoubleScalar(st->gpu_ms));
  return s;
}

/* Q{k}{p} (padded) or U{k}{p} (interior) cells from a packed buffer */
static mxArray* level_cells(bsam_hier* h, const double* buf, const mxArray* boxes_src, int padded) {
  (void)h;
  const int nlev = (int)mxGetNumberOfElements(boxes_src);
  mxArray* out = mxCreateCellMatrix(1, nlev);
  size_t off = 0;
  for (int k = 0; k < nlev; ++k) {
    const mxArray* box = mxGetCell(boxes_src, k);
    const size_t np = mxGetM(box);
    mxArray* lc = mxCreateCellMatrix(1, np);
    for (size_t p = 0; p < np; ++p) {
      const double* b = mxGetPr(box);
      const size_t mx = (size_t)(b[p + np] - b[p] + 1) + (padded ? 2 : 0);
      const size_t my = (size_t)(b[p + 3 * np] - b[p + 2 * np] + 1) + (padded ? 2 : 0);
      mxArray* a = mxCreateDoubleMatrix(mx, my, mxREAL);
      memcpy(mxGetPr(a), buf + off, mx * my * sizeof(double));
      off += mx * my;
      mxSetCell(lc, p, a);
    }
    mxSetCell(out, k, lc);
  }
  return out;
}

/* the handle keeps the level boxes for shaping outputs */
typedef struct {
  bsam_hier* h;
  mxArray* boxes; /* persistent cell {nlev} of np x 4 */
  long long ncell;
  long long nq;
} entry;

#define MAXH 4096
static entry g_h[MAXH];

static entry* find(bsam_hier* h) {
  for (int i = 0; i < MAXH; ++i)
    if (g_h[i].h == h) return &g_h[i];
  mexErrMsgIdAndTxt("bsamgpu:arg", "unknown handle");
  return NULL;
}

static double* pack_cells(const mxArray* U, const entry* e) {
  if (mxIsEmpty(U)) return NULL;
  double* buf = (double*)mxCalloc((size_t)e->ncell, sizeof(double));
  size_t off = 0;
  const int nlev = (int)mxGetNumberOfElements(e->boxes);
  for (int k = 0; k < nlev; ++k) {
    const mxArray* box = mxGetCell(e->boxes, k);
    const size_t np = mxGetM(box);
    const mxArray* lc = mxGetCell(U, k);
    for (size_t p = 0; p < np; ++p) {
      const double* b = mxGetPr(box);
      const size_t mx = (size_t)(b[p + np] - b[p] + 1), my = (size_t)(b[p + 3 * np] - b[p + 2 * np] + 1);
      const mxArray* a = lc ? mxGetCell(lc, p) : NULL;
      if (!a || mxGetM(a) != mx || mxGetN(a) != my)
        mexErrMsgIdAndTxt("bsamgpu:arg", "U{%d}{%zu} must be %zu x %zu", k + 1, p + 1, mx, my);
      memcpy(buf + off, mxGetPr(a), mx * my * sizeof(double));
      off += mx * my;
    }
  }
  return buf;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  mexAtExit(at_exit);
  if (nrhs < 1 || !mxIsChar(prhs[0])) mexErrMsgIdAndTxt("bsamgpu:arg", "first argument is a command string");
  char cmd[32];
  mxGetString(prhs[0], cmd, sizeof cmd);

  if (!strcmp(cmd, "create")) {
    if (nrhs < 2 || !mxIsStruct(prhs[1])) mexErrMsgIdAndTxt("bsamgpu:arg", "create needs the hierarchy H");
    const int dev = nrhs > 2 ? (int)mxGetScalar(prhs[2]) : 0;
    if (dev < 0 || dev >= MAXDEV) mexErrMsgIdAndTxt("bsamgpu:arg", "bad device");
    if (!g_ctx[dev]) chk(bsam_ctx_create(dev, &g_ctx[dev]));
    packed pk;
    pack(prhs[1], &pk);
    bsam_solve_opts so;
    opts_from(mxGetField(prhs[1], 0, "opts"), &so);
    bsam_hier* h = NULL;
    const int rc = bsam_hier_create(g_ctx[dev], &pk.d, &so, &h);
    entry* e = NULL;
    for (int i = 0; i < MAXH && !e; ++i)
      if (!g_h[i].h) e = &g_h[i];
    if (rc == 0 && !e) {
      bsam_hier_destroy(h);
      unpack_free(&pk);
      mexErrMsgIdAndTxt("bsamgpu:arg", "too many open hierarchies");
    }
    if (rc) {
      unpack_free(&pk);
      lib_error();
    }
    e->h = h;
    e->ncell = (long long)pk.FT.n;
    e->nq = (long long)pk.Q.n;
    e->boxes = mxCreateCellMatrix(1, pk.d.nlev);
    const mxArray* levs = field(prhs[1], "lev");
    for (int k = 0; k < pk.d.nlev; ++k) mxSetCell(e->boxes, k, mxDuplicateArray(field(mxGetCell(levs, k), "box")));
    mexMakeArrayPersistent(e->boxes);
    unpack_free(&pk);
    plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
    *(uint64_t*)mxGetData(plhs[0]) = (uint64_t)(uintptr_t)h;
    return;
  }
  if (!strcmp(cmd, "cluster")) {
    /* boxes = bsamgpu_mex('cluster', T, allowed, MinBoxWidth, MinFillRatio):
       the clustering half of ebTagCluster.m:92-110 (host code, no handle) */
    if (nrhs < 5 || !mxIsLogical(prhs[1]) || !mxIsLogical(prhs[2]))
      mexErrMsgIdAndTxt("bsamgpu:arg", "cluster(T, allowed, minw, fill) with logical T and allowed");
    const int n1 = (int)mxGetM(prhs[1]), n2 = (int)mxGetN(prhs[1]);
    if ((int)mxGetM(prhs[2]) != n1 || (int)mxGetN(prhs[2]) != n2)
      mexErrMsgIdAndTxt("bsamgpu:arg", "cluster: T and allowed differ in size");
    int cap = 64, nb = 0, nd = 0, rc;
    int* buf = NULL;
    for (;;) {
      buf = (int*)mxCalloc((size_t)cap * 4, sizeof(int));
      rc = bsam_cluster_tags((const unsigned char*)mxGetLogicals(prhs[1]), (const unsigned char*)mxGetLogicals(prhs[2]),
                             n1, n2, (int)mxGetScalar(prhs[3]), mxGetScalar(prhs[4]), buf, cap, &nb, &nd);
      if (rc == 0 || nb <= cap) break;
      mxFree(buf);
      cap = nb;
    }
    chk(rc);
    if (nd > 0) mexWarnMsgIdAndTxt("eb_tag_cluster:dropthin", "dropping %d ungrowable sliver cluster box(es)", nd);
    plhs[0] = mxCreateDoubleMatrix((mwSize)nb, 4, mxREAL);
    double* out = mxGetPr(plhs[0]);
    for (int b = 0; b < nb; ++b)
      for (int c = 0; c < 4; ++c) out[b + (size_t)c * nb] = buf[4 * b + c];
    mxFree(buf);
    return;
  }
  if (!strcmp(cmd, "cluster_per_parent")) {
    /* boxes = bsamgpu_mex('cluster_per_parent', T, allowed, parentBoxes, MinBoxWidth, MinFillRatio):
       the per-parent clustering of EllipticBSAMTree/src/ebtTagCluster.m:93-128;
       parentBoxes np x 4 [ilo ihi jlo jhi], result nb x 5 [parentIdx ilo ihi jlo jhi] */
    if (nrhs < 6 || !mxIsLogical(prhs[1]) || !mxIsLogical(prhs[2]) || !mxIsDouble(prhs[3]) ||
        (mxGetNumberOfElements(prhs[3]) > 0 && mxGetN(prhs[3]) != 4))
      mexErrMsgIdAndTxt("bsamgpu:arg", "cluster_per_parent(T, allowed, parentBoxes (np x 4), minw, fill)");
    const int n1 = (int)mxGetM(prhs[1]), n2 = (int)mxGetN(prhs[1]);
    if ((int)mxGetM(prhs[2]) != n1 || (int)mxGetN(prhs[2]) != n2)
      mexErrMsgIdAndTxt("bsamgpu:arg", "cluster_per_parent: T and allowed differ in size");
    const int np = (int)mxGetM(prhs[3]);
    const double* pb = mxGetPr(prhs[3]);
    int* par = (int*)mxCalloc((size_t)(np > 0 ? np : 1) * 4, sizeof(int));
    for (int q = 0; q < np; ++q)
      for (int c = 0; c < 4; ++c) par[4 * q + c] = (int)pb[q + (size_t)c * np];
    int cap = 64, nb = 0, nd = 0, rc;
    int* buf = NULL;
    for (;;) {
      buf = (int*)mxCalloc((size_t)cap * 5, sizeof(int));
      rc = bsam_cluster_tags_per_parent((const unsigned char*)mxGetLogicals(prhs[1]),
                                        (const unsigned char*)mxGetLogicals(prhs[2]), n1, n2, par, np,
                                        (int)mxGetScalar(prhs[4]), mxGetScalar(prhs[5]), buf, cap, &nb, &nd);
      if (rc == 0 || nb <= cap) break;
      mxFree(buf);
      cap = nb;
    }
    mxFree(par);
    chk(rc);
    if (nd > 0) mexWarnMsgIdAndTxt("ebt_tag_cluster:dropthin", "dropping %d ungrowable sliver cluster box(es)", nd);
    plhs[0] = mxCreateDoubleMatrix((mwSize)nb, 5, mxREAL);
    double* out = mxGetPr(plhs[0]);
    for (int b = 0; b < nb; ++b)
      for (int c = 0; c < 5; ++c) out[b + (size_t)c * nb] = buf[5 * b + c];
    mxFree(buf);
    return;
  }
  if (nrhs < 2) mexErrMsgIdAndTxt("bsamgpu:arg", "%s needs a handle", cmd);
  bsam_hier* h = handle(prhs[1]);
  entry* e = find(h);

  if (!strcmp(cmd, "destroy")) {
    bsam_hier_destroy(h);
    mxDestroyArray(e->boxes);
    memset(e, 0, sizeof *e);
  } else if (!strcmp(cmd, "solve") || !strcmp(cmd, "parabolic_step")) {
    const int step = cmd[0] == 'p';
    const mxArray* o = nrhs > (step ? 4 : 2) ? prhs[step ? 4 : 2] : NULL;
    bsam_solve_opts so;
    opts_from(o, &so);
    double* hist = (double*)mxCalloc((size_t)so.max_vcycles + 1, sizeof(double));
    bsam_stage_stats st;
    memset(&st, 0, sizeof st);
    st.res_hist = hist;
    st.res_hist_cap = so.max_vcycles + 1;
    double mass = 0;
    if (step) {
      if (nrhs < 4 || !mxIsChar(prhs[3])) mexErrMsgIdAndTxt("bsamgpu:arg", "parabolic_step(h, dt, scheme, opts)");
      chk(bsam_parabolic_step(h, mxGetScalar(prhs[2]), scheme_code(prhs[3]), &so, &st, &mass));
    } else {
      chk(bsam_solve_cycles(h, &so, &st));
    }
    plhs[0] = stage_struct(&st, hist);
    if (step && nlhs > 1) plhs[1] = mxCreateDoubleScalar(mass);
    mxFree(hist);
  } else if (!strcmp(cmd, "get_q")) {
    double* buf = (double*)mxCalloc((size_t)e->nq, sizeof(double));
    size_t off = 0;
    for (int k = 1; k <= (int)mxGetNumberOfElements(e->boxes); ++k) {
      chk(bsam_get_level(h, k, BSAM_Q, buf + off));
      off += (size_t)bsam_level_field_size(h, k, BSAM_Q);
    }
    plhs[0] = level_cells(h, buf, e->boxes, 1);
    mxFree(buf);
  } else if (!strcmp(cmd, "set_q") || !strcmp(cmd, "update_coeffs")) {
    if (nrhs < 3 || !mxIsStruct(prhs[2])) mexErrMsgIdAndTxt("bsamgpu:arg", "%s needs the hierarchy H", cmd);
    packed pk;
    pack(prhs[2], &pk);
    int rc = 0;
    if (cmd[0] == 's') {
      size_t off = 0;
      for (int k = 1; k <= pk.d.nlev && !rc; ++k) {
        rc = bsam_set_level(h, k, BSAM_Q, pk.Q.a + off);
        off += (size_t)bsam_level_field_size(h, k, BSAM_Q);
      }
    } else {
      rc = bsam_update_coeffs(h, &pk.d);
    }
    unpack_free(&pk);
    chk(rc);
  } else if (!strcmp(cmd, "parabolic_init")) {
    chk(bsam_parabolic_init(h));
  } else if (!strcmp(cmd, "parabolic_rhs")) {
    if (nrhs < 4 || !mxIsChar(prhs[3])) mexErrMsgIdAndTxt("bsamgpu:arg", "parabolic_rhs(h, dt, scheme [, aDt])");
    if (nrhs > 4) chk(bsam_parabolic_set_shift(h, mxGetScalar(prhs[4])));
    chk(bsam_parabolic_rhs(h, mxGetScalar(prhs[2]), scheme_code(prhs[3])));
  } else if (!strcmp(cmd, "parabolic_advance")) {
    double mbar = 0, area = 0;
    chk(bsam_parabolic_advance(h));
    chk(bsam_composite_mean(h, &mbar, &area));
    plhs[0] = mxCreateDoubleScalar(mbar * area);
  } else if (!strcmp(cmd, "set_history")) {
    double* a = nrhs > 2 ? pack_cells(prhs[2], e) : NULL;
    double* b = nrhs > 3 ? pack_cells(prhs[3], e) : NULL;
    const int rc = bsam_parabolic_set_history(h, a, b);
    if (a) mxFree(a);
    if (b) mxFree(b);
    chk(rc);
  } else if (!strcmp(cmd, "set_source")) {
    double* a = nrhs > 2 ? pack_cells(prhs[2], e) : NULL;
    const int rc = bsam_parabolic_set_source(h, a);
    if (a) mxFree(a);
    chk(rc);
  } else if (!strcmp(cmd, "get_history")) {
    const int which = nrhs > 2 ? (int)mxGetScalar(prhs[2]) : 0;
    double* buf = (double*)mxCalloc((size_t)e->ncell, sizeof(double));
    chk(bsam_parabolic_get_history(h, which, buf));
    plhs[0] = level_cells(h, buf, e->boxes, 0);
    mxFree(buf);
  } else if (!strcmp(cmd, "tag")) {
    /* [T, Traw] = bsamgpu_mex('tag', h, k, lev.n, Indicator, threshold, TagBuffer):
       the tag loop, buffer and nesting clip of ebTagCluster.m:27-90 */
    if (nrhs < 7 || !mxIsChar(prhs[4]))
      mexErrMsgIdAndTxt("bsamgpu:arg", "tag(h, k, n, indicator, threshold, buffer)");
    const int k = (int)mxGetScalar(prhs[2]);
    char ind[8] = {0};
    mxGetString(prhs[4], ind, sizeof ind);
    for (char* c = ind; *c; ++c) *c = (char)tolower((unsigned char)*c);
    const int code = !strcmp(ind, "ulap") ? BSAM_TAG_ULAP : !strcmp(ind, "rte") ? BSAM_TAG_RTE : -1;
    const double* n = mxGetPr(prhs[3]);
    if (mxGetNumberOfElements(prhs[3]) != 2 ||
        bsam_level_field_size(h, k, BSAM_TAGMAP) != (long long)n[0] * (long long)n[1])
      mexErrMsgIdAndTxt("bsamgpu:arg", "tag: n does not match level %d", k);
    plhs[0] = mxCreateLogicalMatrix((mwSize)n[0], (mwSize)n[1]);
    mxArray* raw = mxCreateLogicalMatrix((mwSize)n[0], (mwSize)n[1]);
    chk(bsam_tag_level(h, k, code, mxGetScalar(prhs[5]), (int)mxGetScalar(prhs[6]),
                       (unsigned char*)mxGetLogicals(plhs[0]), (unsigned char*)mxGetLogicals(raw), NULL));
    if (nlhs > 1)
      plhs[1] = raw;
    else
      mxDestroyArray(raw);
  } else {
    mexErrMsgIdAndTxt("bsamgpu:arg", "unknown command %s", cmd);
  }
}
