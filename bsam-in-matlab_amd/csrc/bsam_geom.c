/* Host-side geometry helper: signed distance to a closed, radially ordered
 * polygon (positive inside), used to sample the diffuse-domain starfish
 * coefficients that the solve path consumes as inputs.
 *
 * Follows DiffuseDomainBSAM/src/ddStarfishGeometry.m:23-72: exact
 * point-to-segment distance, pruned by a bounding-box tree over segment
 * ranges (leaves of at most 8 segments), and the inside/outside sign taken
 * from the ray through the point's polar angle and its enclosing segment.
 * The pruning never discards a segment that could lower the minimum, so the
 * distance equals the brute-force minimum over all segments.
 *
 * Built with -ffp-contract=off so every expression rounds as written.
 */
#include <math.h>
#include <stdlib.h>

typedef struct {
  double box[4]; /* xmin ymin xmax ymax */
  int left, right, first, last;
} gnode;

typedef struct {
  const double *ax, *ay, *ex, *ey, *len2;
  gnode *nodes;
  int nn;
} gtree;

static int build(gtree *t, int first, int last) {
  int id = t->nn++;
  gnode *n = &t->nodes[id];
  double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
  for (int j = first; j <= last; ++j) {
    double xa = t->ax[j], ya = t->ay[j];
    double xb = t->ax[j] + t->ex[j], yb = t->ay[j] + t->ey[j];
    (void)xb;
    if (xa < x0) x0 = xa; if (xa > x1) x1 = xa;
    if (ya < y0) y0 = ya; if (ya > y1) y1 = ya;
  }
  n->first = first; n->last = last; n->left = -1; n->right = -1;
  n->box[0] = x0; n->box[1] = y0; n->box[2] = x1; n->box[3] = y1;
  if (last - first > 7) {
    int mid = (first + last) / 2;
    int l = build(t, first, mid);
    int r = build(t, mid + 1, last);
    t->nodes[id].left = l;
    t->nodes[id].right = r;
  }
  return id;
}

/* vx, vy: nv vertices with v[nv-1] == v[0] (closed polygon, ns = nv-1
 * segments).  bx, by: segment end points (the next vertex) so that the box
 * covers both ends exactly as [a; b] in the reference. */
int bsam_polygon_signed_distance(const double *vx, const double *vy, int nv,
                                 const double *x, const double *y, long npts,
                                 double *d) {
  int ns = nv - 1;
  if (ns < 3) return -1;
  double *ax = (double *)malloc(sizeof(double) * ns * 5);
  if (!ax) return -2;
  double *ay = ax + ns, *ex = ax + 2 * ns, *ey = ax + 3 * ns, *len2 = ax + 4 * ns;
  for (int j = 0; j < ns; ++j) {
    ax[j] = vx[j]; ay[j] = vy[j];
    ex[j] = vx[j + 1] - vx[j];
    ey[j] = vy[j + 1] - vy[j];
    len2[j] = ex[j] * ex[j] + ey[j] * ey[j];
  }
  gtree t;
  t.ax = ax; t.ay = ay; t.ex = ex; t.ey = ey; t.len2 = len2;
  t.nodes = (gnode *)malloc(sizeof(gnode) * (2 * ns + 8));
  t.nn = 0;
  if (!t.nodes) { free(ax); return -2; }
  /* node boxes must include both end points of each segment */
  build(&t, 0, ns - 1);
  for (int id = 0; id < t.nn; ++id) {
    gnode *n = &t.nodes[id];
    for (int j = n->first; j <= n->last; ++j) {
      double xb = vx[j + 1], yb = vy[j + 1];
      if (xb < n->box[0]) n->box[0] = xb; if (xb > n->box[2]) n->box[2] = xb;
      if (yb < n->box[1]) n->box[1] = yb; if (yb > n->box[3]) n->box[3] = yb;
    }
  }
  const double twopi = 2.0 * M_PI;
  const double slack = 64.0 * 2.220446049250313e-16;
  long errs = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : errs)
  for (long i = 0; i < npts; ++i) {
    double xi = x[i], yi = y[i];
    double ang = atan2(yi, xi);
    ang = ang - floor(ang / twopi) * twopi;
    int seg = (int)floor(ang * ns / twopi) + 1;
    if (seg > ns) seg = ns;
    seg -= 1; /* 0-based */
    double dx0 = xi - ax[seg], dy0 = yi - ay[seg];
    double best = dx0 * dx0 + dy0 * dy0;
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const gnode *n = &t.nodes[stack[--sp]];
      double bx = fmax(fmax(n->box[0] - xi, 0.0), xi - n->box[2]);
      double by = fmax(fmax(n->box[1] - yi, 0.0), yi - n->box[3]);
      if (bx * bx + by * by > best + slack) continue;
      if (n->left >= 0) {
        if (sp + 2 > 128) { errs++; break; }
        stack[sp++] = n->left;
        stack[sp++] = n->right;
      } else {
        for (int j = n->first; j <= n->last; ++j) {
          double tt = ((xi - ax[j]) * ex[j] + (yi - ay[j]) * ey[j]) / len2[j];
          tt = fmax(0.0, fmin(1.0, tt));
          double px = xi - ax[j] - tt * ex[j];
          double py = yi - ay[j] - tt * ey[j];
          double dd = px * px + py * py;
          if (dd < best) best = dd;
        }
      }
    }
    double ux = cos(ang), uy = sin(ang);
    double radius = (ax[seg] * ey[seg] - ay[seg] * ex[seg]) / (ux * ey[seg] - uy * ex[seg]);
    double s = (hypot(xi, yi) > radius) ? -1.0 : 1.0;
    d[i] = s * sqrt(best);
  }
  free(t.nodes);
  free(ax);
  return errs ? -3 : 0;
}
