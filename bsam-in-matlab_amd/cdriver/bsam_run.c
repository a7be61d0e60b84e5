/* bsam_run -- plain C driver of libbsamgpu.
 *
 * Reads a hierarchy descriptor file (the bsam_hier_desc arrays plus the
 * solver options, written by bsamgpu.descfile.write_desc), runs one solve
 * stage (== ebSolveCycles, EllipticBSAM/src/ebSolveCycles.m:1-80) on a HIP
 * device through the C ABI and prints the stage record as one JSON line.
 * Optionally writes the solution Q of every level (packed like the
 * descriptor) and repeats the solve from the uploaded state to time it.
 *
 *   bsam_run DESC [--out Q.bin] [--repeat N] [--mixed] [--device D] [--no-graph]
 *
 * With --device -1 the descriptor is only parsed and its tables derived
 * (host-only hierarchy): a check that needs no GPU.
 *
 * File layout (little endian):
 *   char  magic[8] = "BSAMHD01"
 *   int32 nlev, root_idx;  float64 domain[4]
 *   int32 n[2*nlev], np[nlev], boxes[4*P], bc_code[4*P]        (P = sum np)
 *   int64 nq, ncell, nx, ny, ng
 *   float64 Q[nq], FTrue[ncell], DeW[nx], DnS[ny], CC[ncell], InvDiagS[ncell], g[ng]
 *   int32 npre, npost, max_vcycles, ghost_interp, mass_correction; float64 omega, rel_tol
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/bsamgpu.h"

static int rd(FILE* f, void* p, size_t sz, size_t n) { return fread(p, sz, n, f) == n ? 0 : -1; }

static void* rd_alloc(FILE* f, size_t sz, long long n, int* err) {
  void* p = malloc(n > 0 ? (size_t)n * sz : 1);
  if (!p || (n > 0 && rd(f, p, sz, (size_t)n))) *err = 1;
  return p;
}

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

int main(int argc, char** argv) {
  const char *path = NULL, *out = NULL;
  int repeat = 0, mixed = 0, device = 0, graph = 1;
  for (int i = 1; i < argc; ++i) {
    if (!strcmp(argv[i], "--out") && i + 1 < argc) out = argv[++i];
    else if (!strcmp(argv[i], "--repeat") && i + 1 < argc) repeat = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--device") && i + 1 < argc) device = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--mixed")) mixed = 1;
    else if (!strcmp(argv[i], "--no-graph")) graph = 0;
    else if (!path) path = argv[i];
    else { fprintf(stderr, "bsam_run: unexpected argument %s\n", argv[i]); return 2; }
  }
  if (!path) {
    fprintf(stderr, "usage: bsam_run DESC [--out Q.bin] [--repeat N] [--mixed] [--device D] [--no-graph]\n");
    return 2;
  }
  FILE* f = fopen(path, "rb");
  if (!f) { perror(path); return 1; }
  char magic[8];
  bsam_hier_desc d;
  memset(&d, 0, sizeof d);
  int err = rd(f, magic, 1, 8) || memcmp(magic, "BSAMHD01", 8);
  err = err || rd(f, &d.nlev, sizeof(int), 1) || rd(f, &d.root_idx, sizeof(int), 1) ||
        rd(f, d.domain, sizeof(double), 4);
  if (err || d.nlev < 1 || d.nlev > 64) { fprintf(stderr, "bsam_run: bad descriptor header\n"); return 1; }
  int* n = rd_alloc(f, sizeof(int), 2LL * d.nlev, &err);
  int* np = rd_alloc(f, sizeof(int), d.nlev, &err);
  long long P = 0;
  for (int k = 0; !err && k < d.nlev; ++k) P += np[k];
  int* boxes = rd_alloc(f, sizeof(int), 4 * P, &err);
  int* bc = rd_alloc(f, sizeof(int), 4 * P, &err);
  int64_t sz[5] = {0, 0, 0, 0, 0};
  err = err || rd(f, sz, sizeof(int64_t), 5);
  double* Q = rd_alloc(f, sizeof(double), sz[0], &err);
  double* FT = rd_alloc(f, sizeof(double), sz[1], &err);
  double* DeW = rd_alloc(f, sizeof(double), sz[2], &err);
  double* DnS = rd_alloc(f, sizeof(double), sz[3], &err);
  double* CC = rd_alloc(f, sizeof(double), sz[1], &err);
  double* ID = rd_alloc(f, sizeof(double), sz[1], &err);
  double* g = rd_alloc(f, sizeof(double), sz[4], &err);
  bsam_solve_opts o;
  bsam_default_opts(&o);
  int iop[5];
  double dop[2];
  err = err || rd(f, iop, sizeof(int), 5) || rd(f, dop, sizeof(double), 2);
  fclose(f);
  if (err) { fprintf(stderr, "bsam_run: truncated descriptor %s\n", path); return 1; }
  o.npre = iop[0]; o.npost = iop[1]; o.max_vcycles = iop[2]; o.ghost_interp = iop[3];
  o.mass_correction = iop[4]; o.omega = dop[0]; o.rel_tol = dop[1];
  o.precision = mixed ? BSAM_MIXED : BSAM_FP64;
  o.use_graph = graph;
  d.n = n; d.np = np; d.boxes = boxes; d.bc_code = bc;
  d.Q = Q; d.FTrue = FT; d.DeW = DeW; d.DnS = DnS; d.CC = CC; d.InvDiagS = ID; d.g = g;

  bsam_ctx* ctx = NULL;
  bsam_hier* h = NULL;
  if (bsam_ctx_create(device, &ctx) || bsam_hier_create(ctx, &d, &o, &h)) {
    fprintf(stderr, "bsam_run: %s\n", bsam_last_error());
    return 1;
  }
  if (device < 0) {  /* host-only: descriptor parsed and tables derived, no solve */
    printf("{\"created\": true, \"nlev\": %d, \"patches\": %lld}\n", d.nlev, P);
    bsam_hier_destroy(h);
    bsam_ctx_destroy(ctx);
    return 0;
  }
  double hist[1024];
  bsam_stage_stats st;
  memset(&st, 0, sizeof st);
  st.res_hist = hist;
  st.res_hist_cap = 1024;
  double t0 = now_s();
  if (bsam_solve_cycles(h, &o, &st)) { fprintf(stderr, "bsam_run: %s\n", bsam_last_error()); return 1; }
  double t1 = now_s();
  double rep_s = 0;
  if (repeat > 0) {  /* restart from the uploaded state each time */
    bsam_stage_stats s2;
    memset(&s2, 0, sizeof s2);
    double a = now_s();
    for (int r = 0; r < repeat; ++r)
      if (bsam_hier_reset(h) || bsam_solve_cycles(h, &o, &s2)) {
        fprintf(stderr, "bsam_run: %s\n", bsam_last_error());
        return 1;
      }
    rep_s = (now_s() - a) / repeat;
  }
  printf("{\"vcycles\": %d, \"final_rel\": %.17g, \"visible_cells\": %lld, \"total_cells\": %lld, "
         "\"avg_factor\": %.17g, \"gpu_ms\": %.6f, \"wall_s\": %.6f, \"repeat\": %d, \"s_per_solve\": %.6f, "
         "\"res_hist\": [",
         st.vcycles, st.final_rel, st.visible_cells, st.total_cells, st.avg_factor, st.gpu_ms, t1 - t0, repeat,
         rep_s);
  for (int i = 0; i < st.res_hist_len; ++i) printf(i ? ", %.17g" : "%.17g", hist[i]);
  printf("]}\n");
  if (out) {
    FILE* fo = fopen(out, "wb");
    if (!fo) { perror(out); return 1; }
    for (int k = 1; k <= d.nlev; ++k) {
      long long m = bsam_level_field_size(h, k, BSAM_Q);
      double* buf = malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
      if (!buf || bsam_get_level(h, k, BSAM_Q, buf)) { fprintf(stderr, "bsam_run: %s\n", bsam_last_error()); return 1; }
      if (fwrite(buf, sizeof(double), (size_t)m, fo) != (size_t)m) { perror(out); return 1; }
      free(buf);
    }
    fclose(fo);
  }
  bsam_hier_destroy(h);
  bsam_ctx_destroy(ctx);
  free(n); free(np); free(boxes); free(bc); free(Q); free(FT); free(DeW); free(DnS); free(CC); free(ID); free(g);
  return 0;
}
