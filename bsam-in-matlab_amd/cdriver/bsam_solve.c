/* bsam_solve -- the staged adaptive solve of EllipticBSAM/src/ebSolve.m:22-77
 * in plain C over the library's C ABI, with no MATLAB and no Python:
 *
 *   native hierarchy (bsam_build_create: ladder and root, ebSetup.m)
 *   stage 1: bsam_hier_create + bsam_solve_cycles (ebSolveCycles.m)
 *   while AMR levels < MaxLevels:
 *     tag the finest level on the device (bsam_tag_level, ebTagCluster.m:27-90)
 *     cluster the tags (bsam_cluster_tags, ebTagCluster.m:92-380)
 *     no boxes: stop
 *     add the refined level (bsam_build_add_level, ebAddLevel.m:10-24),
 *     create its device hierarchy, carry the existing levels' Q over
 *     (bsam_hier_inherit) and prolong the parent window into the new level
 *     (bsam_init_level_prolong, ebAddLevel.m:25-32)
 *     solve stage
 *
 * The problem is the README elliptic example: -Lap u + u = f on [0,1]^2,
 * u = sin(2 pi x) sin(2 pi y), f = (1 + 8 pi^2) u, homogeneous Dirichlet
 * sides, 'ulap' tagging.  Prints one JSON line per stage and a final line
 * with the boxes of every level; --out writes Q of every level packed like
 * the descriptor.
 *
 *   bsam_solve [--base N] [--levels L] [--threshold T] [--rel-tol R]
 *              [--max-vcycles M] [--out Q.bin] [--device D]
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/bsamgpu.h"

static void f_plain(void* user, const double* x, const double* y, long long n, double* out) {
  (void)user;
  const double c = 1.0 + 8.0 * pow(M_PI, 2.0);
  for (long long i = 0; i < n; ++i) out[i] = c * (sin((2.0 * M_PI) * x[i]) * sin((2.0 * M_PI) * y[i]));
}
static void one(void* user, const double* x, const double* y, long long n, double* out) {
  (void)user; (void)x; (void)y;
  for (long long i = 0; i < n; ++i) out[i] = 1.0;
}

static int die(const char* what) {
  fprintf(stderr, "bsam_solve: %s: %s\n", what, bsam_last_error());
  return 1;
}

static int stage(bsam_hier* h, const bsam_solve_opts* o, int index) {
  double hist[1024];
  bsam_stage_stats st;
  memset(&st, 0, sizeof st);
  st.res_hist = hist;
  st.res_hist_cap = 1024;
  if (bsam_solve_cycles(h, o, &st)) return die("bsam_solve_cycles");
  printf("{\"stage\": %d, \"amr_levels\": %d, \"nlev\": %d, \"vcycles\": %d, \"final_rel\": %.17g, "
         "\"visible_cells\": %lld, \"total_cells\": %lld, \"avg_factor\": %.17g, \"res_hist\": [",
         index, st.amr_levels, st.nlev, st.vcycles, st.final_rel, st.visible_cells, st.total_cells, st.avg_factor);
  for (int i = 0; i < st.res_hist_len; ++i) printf(i ? ", %.17g" : "%.17g", hist[i]);
  printf("]}\n");
  fflush(stdout);
  return 0;
}

int main(int argc, char** argv) {
  int base = 32, levels = 3, device = 0, maxv = 60;
  double thr = 1e-2, rel_tol = 1e-9;
  const char* out = NULL;
  for (int i = 1; i < argc; ++i) {
    if (!strcmp(argv[i], "--base") && i + 1 < argc) base = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--levels") && i + 1 < argc) levels = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--threshold") && i + 1 < argc) thr = atof(argv[++i]);
    else if (!strcmp(argv[i], "--rel-tol") && i + 1 < argc) rel_tol = atof(argv[++i]);
    else if (!strcmp(argv[i], "--max-vcycles") && i + 1 < argc) maxv = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--out") && i + 1 < argc) out = argv[++i];
    else if (!strcmp(argv[i], "--device") && i + 1 < argc) device = atoi(argv[++i]);
    else { fprintf(stderr, "bsam_solve: unexpected argument %s\n", argv[i]); return 2; }
  }
  /* the problem (ebProblem) and options (ebOptions defaults: TagBuffer 2,
     MinBoxWidth 4, MinFillRatio 0.7, CoarsestCells 16) */
  bsam_problem pr;
  memset(&pr, 0, sizeof pr);
  pr.domain[0] = 0; pr.domain[1] = 1; pr.domain[2] = 0; pr.domain[3] = 1;
  pr.D = one; pr.C = one; pr.f = f_plain;
  for (int s = 0; s < 4; ++s) pr.bc_type[s] = 1;
  bsam_build_opts bo;
  memset(&bo, 0, sizeof bo);
  bo.base_cells[0] = bo.base_cells[1] = base;
  bo.coarsest_cells = 16;
  bo.source_quad = 1;
  const int tag_buffer = 2, min_width = 4;
  const double fill_ratio = 0.70;
  bsam_solve_opts o;
  bsam_default_opts(&o);
  o.rel_tol = rel_tol;
  o.max_vcycles = maxv;

  bsam_build* b = NULL;
  bsam_ctx* ctx = NULL;
  bsam_hier* h = NULL;
  bsam_hier_desc d;
  if (bsam_build_create(&pr, &bo, &b)) return die("bsam_build_create");
  if (bsam_ctx_create(device, &ctx)) return die("bsam_ctx_create");
  if (bsam_build_desc(b, &d) || bsam_hier_create(ctx, &d, &o, &h)) return die("bsam_hier_create");
  int nlev = 0, root = 0, sing = 0, nst = 1;
  if (bsam_build_info(b, &nlev, &root, &sing)) return die("bsam_build_info");
  if (stage(h, &o, nst)) return 1;
  while (nlev - root + 1 < levels) {
    const int n1 = d.n[2 * (nlev - 1)], n2 = d.n[2 * (nlev - 1) + 1];
    unsigned char* tags = malloc((size_t)n1 * n2);
    unsigned char* allowed = malloc((size_t)n1 * n2);
    if (!tags || !allowed) { fprintf(stderr, "bsam_solve: out of memory\n"); return 1; }
    long long ntag = 0;
    /* TagThreshold at AMR level nlev - root + 1 (constant here) */
    if (bsam_tag_level(h, nlev, BSAM_TAG_ULAP, thr, tag_buffer, tags, NULL, &ntag)) return die("bsam_tag_level");
    if (bsam_build_allowed(b, allowed)) return die("bsam_build_allowed");
    int cap = 256, nboxes = 0, ndrop = 0;
    int* boxes = NULL;
    if (ntag > 0)
      for (;;) {
        boxes = realloc(boxes, (size_t)cap * 4 * sizeof(int));
        if (!boxes) { fprintf(stderr, "bsam_solve: out of memory\n"); return 1; }
        if (bsam_cluster_tags(tags, allowed, n1, n2, min_width, fill_ratio, boxes, cap, &nboxes, &ndrop))
          return die("bsam_cluster_tags");
        if (nboxes <= cap) break;
        cap = nboxes;
      }
    free(tags);
    free(allowed);
    if (nboxes == 0) {  /* ebSolve.m:45-50: nothing tagged */
      free(boxes);
      break;
    }
    if (bsam_build_add_level(b, boxes, nboxes)) return die("bsam_build_add_level");
    free(boxes);
    if (bsam_build_info(b, &nlev, &root, &sing)) return die("bsam_build_info");
    bsam_hier* h2 = NULL;
    if (bsam_build_desc(b, &d) || bsam_hier_create(ctx, &d, &o, &h2)) return die("bsam_hier_create");
    if (bsam_hier_inherit(h2, h)) return die("bsam_hier_inherit");
    bsam_hier_destroy(h);
    h = h2;
    if (bsam_init_level_prolong(h, nlev)) return die("bsam_init_level_prolong");
    if (stage(h, &o, ++nst)) return 1;
  }
  /* boxes of every level */
  printf("{\"nlev\": %d, \"root_idx\": %d, \"singular\": %d, \"boxes\": [", nlev, root, sing);
  int off = 0;
  for (int k = 0; k < nlev; ++k) {
    printf(k ? ", [" : "[");
    for (int p = 0; p < d.np[k]; ++p, off += 4)
      printf(p ? ", [%d, %d, %d, %d]" : "[%d, %d, %d, %d]", d.boxes[off], d.boxes[off + 1], d.boxes[off + 2],
             d.boxes[off + 3]);
    printf("]");
  }
  printf("]}\n");
  if (out) {
    FILE* fo = fopen(out, "wb");
    if (!fo) { perror(out); return 1; }
    for (int k = 1; k <= nlev; ++k) {
      long long m = bsam_level_field_size(h, k, BSAM_Q);
      double* buf = malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
      if (!buf || bsam_get_level(h, k, BSAM_Q, buf)) return die("bsam_get_level");
      if (fwrite(buf, sizeof(double), (size_t)m, fo) != (size_t)m) { perror(out); return 1; }
      free(buf);
    }
    fclose(fo);
  }
  bsam_hier_destroy(h);
  bsam_ctx_destroy(ctx);
  bsam_build_destroy(b);
  return 0;
}
