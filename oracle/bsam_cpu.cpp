// CPU restatement of the solve stage in C++ -- TEST AND BASELINE
// INFRASTRUCTURE ONLY (the CPU baseline that bench.py times, and a checker
// that tests compare against the NumPy oracle).  Nothing in the product
// links or loads it.
//
// It runs ebSolveCycles (EllipticBSAM/src/ebSolveCycles.m:8-79) on the
// descriptors the NumPy oracle derives (oracle/bsam_oracle.py, build_ops and
// gather_parent_coeffs, passed in flattened by oracle/bsam_cpu.py), with the
// same per-cell expressions in the same order, so that it reproduces the
// NumPy oracle's residual history and solution bit for bit.  Patches of a
// level are processed in parallel with OpenMP where the reference's
// per-patch steps are independent; the flux-balance additions into parent
// cells run in the reference's (patch, face, entry) order.  Build:
// oracle/Makefile (-O3 -fopenmp -ffp-contract=off).
//
// Functions and the reference lines they restate:
//   op2d / relax2d / restrict2 / prolong_bilin   ebKernels.m:13-123
//   interp_cf                                    ebInterpCf.m:36-154
//   fill_ghosts                                  ebFillGhosts.m:1-95
//   relax_level                                  ebRelaxLevel.m:28-38
//   gather_window, transfer                      ebGatherWindow.m:7-19, ebTransfer.m:10-79
//   face_corrections, coarse_load                ebFaceCorrections.m:29-55, ebCoarseLoad.m:17-46
//   coarsest_solve                               ebCoarsest.m:79-91 (explicit inverse, ascending dots)
//   comp_residual                                ebCompResidual.m:23-84
//   vcycle, solve_cycles                         ebVcycle.m:15-35, ebSolveCycles.m:8-79
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

const double A1 = -1.0 / 5.0, A2 = 2.0 / 3.0, A3 = 8.0 / 15.0;
const double B1 = 5.0 / 32.0, B2 = 15.0 / 16.0, B3 = -3.0 / 32.0;
const double C83 = 8.0 / 3.0, C13 = 1.0 / 3.0, C23 = 2.0 / 3.0;

// a column-major 2-D view (i along x)
struct A2d {
  double* a = nullptr;
  int n0 = 0, n1 = 0;
  double& operator()(int i, int j) const { return a[i + (long long)j * n0]; }
};

struct Rect {
  int q, s0, s1, s2, s3, d0, d1, d2, d3;
};
struct Pt {
  int q, si, sj, di, dj;
};
struct Corr {
  int q, src, di, dj;
};

struct Patch {
  int mx, my, par0;
  int code[4];
  int fint[4];
  int ccl[4];
  A2d Q, F, FT, DeW, DnS, CC, ID, QC;
  double* g[4];
  const unsigned char* cov;
  std::vector<Pt> same, ring;
  std::vector<Rect> winrect, usc, fsc, pcc, pdew, pdns;
  std::vector<Corr> corr[4];
  double* cfD[4];
};

struct Level {
  int n[2];
  double h[2];
  std::vector<Patch> P;
};

struct Hier {
  std::vector<Level> L;  // index 0 = level 1
  int root;
  int npre, npost, mass, qif;
  double omega;
  const double* Ainv;
  const double* bcRhs;
  int N, M, singular;
  Level& lev(int k) { return L[k - 1]; }
};

// ---------------------------------------------------------------- kernels
// -(feW(i+1)-feW(i))/hx^2 - (fnS(j+1)-fnS(j))/hy^2 + C u at interior (i, j) (0-based)
inline double op2d_at(const A2d& Q, const A2d& DeW, const A2d& DnS, const A2d& CC, int i, int j, double hx,
                      double hy) {
  const double fe = DeW(i + 1, j) * (Q(i + 2, j + 1) - Q(i + 1, j + 1));
  const double fw = DeW(i, j) * (Q(i + 1, j + 1) - Q(i, j + 1));
  const double fn = DnS(i, j + 1) * (Q(i + 1, j + 2) - Q(i + 1, j + 1));
  const double fs = DnS(i, j) * (Q(i + 1, j + 1) - Q(i + 1, j));
  return (-(fe - fw) / (hx * hx) - (fn - fs) / (hy * hy)) + CC(i, j) * Q(i + 1, j + 1);
}

void relax2d(Patch& p, int color, double hx, double hy, double omega) {
  A2d& Q = p.Q;
  const int mx = p.mx, my = p.my;
  for (int j = 1; j <= my; ++j) {
    if (p.code[0] == 1) Q(0, j) = C83 * p.g[0][j - 1] + C13 * Q(2, j);
    else if (p.code[0] == 2) Q(0, j) = hx * p.g[0][j - 1];
    if (p.code[1] == 1) Q(mx + 1, j) = C83 * p.g[1][j - 1] + C13 * Q(mx - 1, j);
    else if (p.code[1] == 2) Q(mx + 1, j) = hx * p.g[1][j - 1];
  }
  for (int i = 1; i <= mx; ++i) {
    if (p.code[2] == 1) Q(i, 0) = C83 * p.g[2][i - 1] + C13 * Q(i, 2);
    else if (p.code[2] == 2) Q(i, 0) = hy * p.g[2][i - 1];
    if (p.code[3] == 1) Q(i, my + 1) = C83 * p.g[3][i - 1] + C13 * Q(i, my - 1);
    else if (p.code[3] == 2) Q(i, my + 1) = hy * p.g[3][i - 1];
  }
  const double hx2 = hx * hx, hy2 = hy * hy;
  for (int j = 0; j < my; ++j)
    for (int i = 0; i < mx; ++i) {
      if ((((i + 1) + (j + 1) + p.par0) % 2) != color - 1) continue;
      // the cell's colour-mates do not read it, so the in-place update is the
      // vectorised Jacobi candidate of ebKernels.m:75-84
      const double tmp = ((((p.F(i, j) + Q(i + 2, j + 1) * p.DeW(i + 1, j) / hx2) + Q(i, j + 1) * p.DeW(i, j) / hx2) +
                           Q(i + 1, j + 2) * p.DnS(i, j + 1) / hy2) +
                          Q(i + 1, j) * p.DnS(i, j) / hy2) *
                         p.ID(i, j);
      Q(i + 1, j + 1) = omega == 1 ? tmp : omega * tmp + (1 - omega) * Q(i + 1, j + 1);
    }
}

void interp_cf(Patch& p, bool qif) {
  A2d& Q = p.Q;
  const A2d& QC = p.QC;
  const int mx = p.mx, my = p.my, cmx = mx / 2, cmy = my / 2;
  if (p.fint[0]) {  // WEST
    if (qif) {
      for (int jj = 2; jj <= my - 2; jj += 2) Q(0, jj) = A1 * Q(2, jj + 2) + A2 * Q(1, jj + 1) + A3 * QC(0, jj / 2);
      for (int jj = 3; jj <= my - 1; jj += 2)
        Q(0, jj) = A1 * Q(2, jj - 2) + A2 * Q(1, jj - 1) + A3 * QC(0, (jj + 1) / 2);
      Q(0, my) = A1 * Q(2, my) + A2 * Q(1, my) + A3 * (B1 * QC(0, cmy + 1) + B2 * QC(0, cmy) + B3 * QC(0, cmy - 1));
      Q(0, 1) = A1 * Q(2, 1) + A2 * Q(1, 1) + A3 * (B1 * QC(0, 0) + B2 * QC(0, 1) + B3 * QC(0, 2));
    } else {
      for (int jj = 2; jj <= my; jj += 2) Q(0, jj) = C23 * QC(0, jj / 2) + C23 * Q(1, jj) - C13 * Q(1, jj - 1);
      for (int jj = 1; jj <= my - 1; jj += 2)
        Q(0, jj) = C23 * QC(0, (jj + 1) / 2) + C23 * Q(1, jj) - C13 * Q(1, jj + 1);
    }
  }
  if (p.fint[1]) {  // EAST
    if (qif) {
      for (int jj = 2; jj <= my - 2; jj += 2)
        Q(mx + 1, jj) = A1 * Q(mx - 1, jj + 2) + A2 * Q(mx, jj + 1) + A3 * QC(cmx + 1, jj / 2);
      for (int jj = 3; jj <= my - 1; jj += 2)
        Q(mx + 1, jj) = A1 * Q(mx - 1, jj - 2) + A2 * Q(mx, jj - 1) + A3 * QC(cmx + 1, (jj + 1) / 2);
      Q(mx + 1, my) = A1 * Q(mx - 1, my) + A2 * Q(mx, my) +
                      A3 * (B1 * QC(cmx + 1, cmy + 1) + B2 * QC(cmx + 1, cmy) + B3 * QC(cmx + 1, cmy - 1));
      Q(mx + 1, 1) = A1 * Q(mx - 1, 1) + A2 * Q(mx, 1) +
                     A3 * (B1 * QC(cmx + 1, 0) + B2 * QC(cmx + 1, 1) + B3 * QC(cmx + 1, 2));
    } else {
      for (int jj = 2; jj <= my; jj += 2)
        Q(mx + 1, jj) = C23 * QC(cmx + 1, jj / 2) + C23 * Q(mx, jj) - C13 * Q(mx, jj - 1);
      for (int jj = 1; jj <= my - 1; jj += 2)
        Q(mx + 1, jj) = C23 * QC(cmx + 1, (jj + 1) / 2) + C23 * Q(mx, jj) - C13 * Q(mx, jj + 1);
    }
  }
  if (p.fint[2]) {  // SOUTH
    if (qif) {
      for (int ii = 2; ii <= mx - 2; ii += 2) Q(ii, 0) = A1 * Q(ii + 2, 2) + A2 * Q(ii + 1, 1) + A3 * QC(ii / 2, 0);
      for (int ii = 3; ii <= mx - 1; ii += 2)
        Q(ii, 0) = A1 * Q(ii - 2, 2) + A2 * Q(ii - 1, 1) + A3 * QC((ii + 1) / 2, 0);
      Q(mx, 0) = A1 * Q(mx, 2) + A2 * Q(mx, 1) + A3 * (B1 * QC(cmx + 1, 0) + B2 * QC(cmx, 0) + B3 * QC(cmx - 1, 0));
      Q(1, 0) = A1 * Q(1, 2) + A2 * Q(1, 1) + A3 * (B1 * QC(0, 0) + B2 * QC(1, 0) + B3 * QC(2, 0));
    } else {
      for (int ii = 2; ii <= mx; ii += 2) Q(ii, 0) = C23 * QC(ii / 2, 0) + C23 * Q(ii, 1) - C13 * Q(ii - 1, 1);
      for (int ii = 1; ii <= mx - 1; ii += 2)
        Q(ii, 0) = C23 * QC((ii + 1) / 2, 0) + C23 * Q(ii, 1) - C13 * Q(ii + 1, 1);
    }
  }
  if (p.fint[3]) {  // NORTH
    if (qif) {
      for (int ii = 2; ii <= mx - 2; ii += 2)
        Q(ii, my + 1) = A1 * Q(ii + 2, my - 1) + A2 * Q(ii + 1, my) + A3 * QC(ii / 2, cmy + 1);
      for (int ii = 3; ii <= mx - 1; ii += 2)
        Q(ii, my + 1) = A1 * Q(ii - 2, my - 1) + A2 * Q(ii - 1, my) + A3 * QC((ii + 1) / 2, cmy + 1);
      Q(mx, my + 1) = A1 * Q(mx, my - 1) + A2 * Q(mx, my) +
                      A3 * (B1 * QC(cmx + 1, cmy + 1) + B2 * QC(cmx, cmy + 1) + B3 * QC(cmx - 1, cmy + 1));
      Q(1, my + 1) = A1 * Q(1, my - 1) + A2 * Q(1, my) +
                     A3 * (B1 * QC(0, cmy + 1) + B2 * QC(1, cmy + 1) + B3 * QC(2, cmy + 1));
    } else {
      for (int ii = 2; ii <= mx; ii += 2)
        Q(ii, my + 1) = C23 * QC(ii / 2, cmy + 1) + C23 * Q(ii, my) - C13 * Q(ii - 1, my);
      for (int ii = 1; ii <= mx - 1; ii += 2)
        Q(ii, my + 1) = C23 * QC((ii + 1) / 2, cmy + 1) + C23 * Q(ii, my) - C13 * Q(ii + 1, my);
    }
  }
  if (p.ccl[0] == 3)
    Q(0, 0) = qif ? (A1 * Q(2, 2) + A2 * Q(1, 1) + A3 * QC(0, 0)) : (C23 * QC(0, 0) + C13 * Q(1, 1));
  if (p.ccl[1] == 3)
    Q(mx + 1, 0) = qif ? (A1 * Q(mx - 1, 2) + A2 * Q(mx, 1) + A3 * QC(cmx + 1, 0))
                       : (C23 * QC(cmx + 1, 0) + C13 * Q(mx, 1));
  if (p.ccl[2] == 3)
    Q(0, my + 1) = qif ? (A1 * Q(2, my - 1) + A2 * Q(1, my) + A3 * QC(0, cmy + 1))
                       : (C23 * QC(0, cmy + 1) + C13 * Q(1, my));
  if (p.ccl[3] == 3)
    Q(mx + 1, my + 1) = qif ? (A1 * Q(mx - 1, my - 1) + A2 * Q(mx, my) + A3 * QC(cmx + 1, cmy + 1))
                            : (C23 * QC(cmx + 1, cmy + 1) + C13 * Q(mx, my));
}

void fill_ghosts(Hier& H, int k) {
  Level& lev = H.lev(k);
  const double hx = lev.h[0], hy = lev.h[1];
  const int np = (int)lev.P.size();
  if (k > 1) {
    Level& par = H.lev(k - 1);
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < np; ++p) {
      Patch& P = lev.P[p];
      const bool any = P.fint[0] || P.fint[1] || P.fint[2] || P.fint[3] || P.ccl[0] == 3 || P.ccl[1] == 3 ||
                       P.ccl[2] == 3 || P.ccl[3] == 3;
      if (!any) continue;
      for (const Pt& r : P.ring) P.QC(r.di, r.dj) = par.P[r.q].Q(r.si, r.sj);
      interp_cf(P, H.qif);
    }
  }
#pragma omp parallel for schedule(dynamic)
  for (int p = 0; p < np; ++p)
    for (const Pt& s : lev.P[p].same) lev.P[p].Q(s.di, s.dj) = lev.P[s.q].Q(s.si, s.sj);
#pragma omp parallel for schedule(dynamic)
  for (int p = 0; p < np; ++p) {
    Patch& P = lev.P[p];
    const int* c = P.code;
    const int* cc = P.ccl;
    if (c[0] == 0 && c[1] == 0 && c[2] == 0 && c[3] == 0 && cc[0] != 1 && cc[1] != 1 && cc[2] != 1 && cc[3] != 1)
      continue;
    A2d& Q = P.Q;
    const int mx = P.mx, my = P.my;
    for (int j = 1; j <= my; ++j) {
      if (c[0] == 1) Q(0, j) = C83 * P.g[0][j - 1] - 2 * Q(1, j) + C13 * Q(2, j);
      else if (c[0] == 2) Q(0, j) = Q(1, j) + hx * P.g[0][j - 1];
    }
    for (int j = 1; j <= my; ++j) {
      if (c[1] == 1) Q(mx + 1, j) = C83 * P.g[1][j - 1] - 2 * Q(mx, j) + C13 * Q(mx - 1, j);
      else if (c[1] == 2) Q(mx + 1, j) = Q(mx, j) + hx * P.g[1][j - 1];
    }
    for (int i = 1; i <= mx; ++i) {
      if (c[2] == 1) Q(i, 0) = C83 * P.g[2][i - 1] - 2 * Q(i, 1) + C13 * Q(i, 2);
      else if (c[2] == 2) Q(i, 0) = Q(i, 1) + hy * P.g[2][i - 1];
    }
    for (int i = 1; i <= mx; ++i) {
      if (c[3] == 1) Q(i, my + 1) = C83 * P.g[3][i - 1] - 2 * Q(i, my) + C13 * Q(i, my - 1);
      else if (c[3] == 2) Q(i, my + 1) = Q(i, my) + hy * P.g[3][i - 1];
    }
    if (cc[0] == 1) Q(0, 0) = Q(1, 0) + Q(0, 1) - Q(1, 1);
    if (cc[1] == 1) Q(mx + 1, 0) = Q(mx, 0) + Q(mx + 1, 1) - Q(mx, 1);
    if (cc[2] == 1) Q(0, my + 1) = Q(1, my + 1) + Q(0, my) - Q(1, my);
    if (cc[3] == 1) Q(mx + 1, my + 1) = Q(mx, my + 1) + Q(mx + 1, my) - Q(mx, my);
  }
}

void relax_level(Hier& H, int k, int nsweeps) {
  Level& lev = H.lev(k);
  const int np = (int)lev.P.size();
  for (int s = 0; s < nsweeps; ++s)
    for (int color = 1; color <= 2; ++color) {
#pragma omp parallel for schedule(dynamic)
      for (int p = 0; p < np; ++p) relax2d(lev.P[p], color, lev.h[0], lev.h[1], H.omega);
      fill_ghosts(H, k);
    }
}

inline double restrict_at(const A2d& Q, int ci, int cj) {  // restrict2 of the interior at coarse (ci, cj)
  const int i = 2 * ci + 1, j = 2 * cj + 1;
  return 0.25 * (((Q(i, j) + Q(i + 1, j)) + Q(i, j + 1)) + Q(i + 1, j + 1));
}

void restrict_level(Hier& H, int k) {
  Level& lev = H.lev(k);
  Level& par = H.lev(k - 1);
  const int np = (int)lev.P.size();
#pragma omp parallel for schedule(dynamic)
  for (int p = 0; p < np; ++p) {
    Patch& P = lev.P[p];
    const int cmx = P.mx / 2, cmy = P.my / 2;
    for (int cj = 0; cj < cmy; ++cj)
      for (int ci = 0; ci < cmx; ++ci) P.QC(ci + 1, cj + 1) = restrict_at(P.Q, ci, cj);
    for (const Rect& r : P.usc)
      for (int j = r.d2; j <= r.d3; ++j)
        for (int i = r.d0; i <= r.d1; ++i) par.P[r.q].Q(i, j) = P.QC(r.s0 + (i - r.d0) + 1, r.s2 + (j - r.d2) + 1);
  }
}

void ring_level(Hier& H, int k) {
  Level& lev = H.lev(k);
  Level& par = H.lev(k - 1);
  const int np = (int)lev.P.size();
#pragma omp parallel for schedule(dynamic)
  for (int p = 0; p < np; ++p)
    for (const Pt& r : lev.P[p].ring) lev.P[p].QC(r.di, r.dj) = par.P[r.q].Q(r.si, r.sj);
}

void correct_level(Hier& H, int k) {
  Level& lev = H.lev(k);
  Level& par = H.lev(k - 1);
  const int np = (int)lev.P.size();
#pragma omp parallel for schedule(dynamic)
  for (int p = 0; p < np; ++p) {
    Patch& P = lev.P[p];
    const int cm = P.mx / 2, cn = P.my / 2;
    std::vector<double> Wv((size_t)(cm + 2) * (cn + 2), 0.0);
    A2d W{Wv.data(), cm + 2, cn + 2};
    for (const Pt& r : P.ring) W(r.di, r.dj) = par.P[r.q].Q(r.si, r.sj);
    for (const Rect& r : P.winrect)
      for (int j = r.d2; j <= r.d3; ++j)
        for (int i = r.d0; i <= r.d1; ++i) W(i, j) = par.P[r.q].Q(r.s0 + (i - r.d0), r.s2 + (j - r.d2));
    for (size_t t = 0; t < Wv.size(); ++t) Wv[t] = Wv[t] - P.QC.a[t];  // E = W - QC
    for (int cj = 0; cj < cn; ++cj)
      for (int ci = 0; ci < cm; ++ci) {
        const double C = W(ci + 1, cj + 1), Wm = W(ci, cj + 1), Wp = W(ci + 2, cj + 1), Sm = W(ci + 1, cj),
                     Sp = W(ci + 1, cj + 2);
        const double Dmm = W(ci, cj), Dpm = W(ci + 2, cj), Dmp = W(ci, cj + 2), Dpp = W(ci + 2, cj + 2);
        const double f00 = (9 * C + 3 * Wm + 3 * Sm + Dmm) / 16, f10 = (9 * C + 3 * Wp + 3 * Sm + Dpm) / 16;
        const double f01 = (9 * C + 3 * Wm + 3 * Sp + Dmp) / 16, f11 = (9 * C + 3 * Wp + 3 * Sp + Dpp) / 16;
        A2d& Q = P.Q;
        Q(2 * ci + 1, 2 * cj + 1) = Q(2 * ci + 1, 2 * cj + 1) + f00;
        Q(2 * ci + 2, 2 * cj + 1) = Q(2 * ci + 2, 2 * cj + 1) + f10;
        Q(2 * ci + 1, 2 * cj + 2) = Q(2 * ci + 1, 2 * cj + 2) + f01;
        Q(2 * ci + 2, 2 * cj + 2) = Q(2 * ci + 2, 2 * cj + 2) + f11;
      }
  }
}

void filldown(Hier& H) {
  for (int kk = (int)H.L.size(); kk >= 2; --kk) {
    restrict_level(H, kk);
    fill_ghosts(H, kk - 1);
  }
}

// ebFaceCorrections.m:29-55: vals[f][t] for the coarse faces of patch p
void face_corrections(Hier& H, int k, Patch& P, std::vector<double> vals[4]) {
  const A2d& Q = P.Q;
  const A2d& QC = P.QC;
  const int mx = P.mx, my = P.my, cmx = mx / 2, cmy = my / 2;
  const double hcx2 = std::pow(H.lev(k - 1).h[0], 2.0), hcy2 = std::pow(H.lev(k - 1).h[1], 2.0);
  vals[0].resize(cmy); vals[1].resize(cmy); vals[2].resize(cmx); vals[3].resize(cmx);
  for (int j = 1; j <= cmy; ++j) {
    vals[0][j - 1] = (P.DeW(0, 2 * j - 2) * (Q(1, 2 * j - 1) - Q(0, 2 * j - 1)) +
                      P.DeW(0, 2 * j - 1) * (Q(1, 2 * j) - Q(0, 2 * j)) + P.cfD[0][j - 1] * (QC(0, j) - QC(1, j))) /
                     hcx2;
    vals[1][j - 1] = (P.DeW(mx, 2 * j - 2) * (Q(mx, 2 * j - 1) - Q(mx + 1, 2 * j - 1)) +
                      P.DeW(mx, 2 * j - 1) * (Q(mx, 2 * j) - Q(mx + 1, 2 * j)) +
                      P.cfD[1][j - 1] * (QC(cmx + 1, j) - QC(cmx, j))) /
                     hcx2;
  }
  for (int i = 1; i <= cmx; ++i) {
    vals[2][i - 1] = (P.DnS(2 * i - 2, 0) * (Q(2 * i - 1, 1) - Q(2 * i - 1, 0)) +
                      P.DnS(2 * i - 1, 0) * (Q(2 * i, 1) - Q(2 * i, 0)) + P.cfD[2][i - 1] * (QC(i, 0) - QC(i, 1))) /
                     hcy2;
    vals[3][i - 1] = (P.DnS(2 * i - 2, my) * (Q(2 * i - 1, my) - Q(2 * i - 1, my + 1)) +
                      P.DnS(2 * i - 1, my) * (Q(2 * i, my) - Q(2 * i, my + 1)) +
                      P.cfD[3][i - 1] * (QC(i, cmy + 1) - QC(i, cmy))) /
                     hcy2;
  }
}

void coarse_load(Hier& H, int k) {
  Level& lev = H.lev(k);
  Level& par = H.lev(k - 1);
  const double hx = lev.h[0], hy = lev.h[1], hcx = par.h[0], hcy = par.h[1];
  const int np = (int)lev.P.size();
  std::vector<std::vector<double>> vals((size_t)np * 4);
#pragma omp parallel for schedule(dynamic)
  for (int p = 0; p < np; ++p) {
    Patch& P = lev.P[p];
    const int mx = P.mx, my = P.my, cmx = mx / 2, cmy = my / 2;
    // parent coefficients over the footprint (ebParentCoeffs.m:16-30)
    std::vector<double> pdw((size_t)(cmx + 1) * cmy), pds((size_t)cmx * (cmy + 1)), pcc((size_t)cmx * cmy);
    A2d PDeW{pdw.data(), cmx + 1, cmy}, PDnS{pds.data(), cmx, cmy + 1}, PCC{pcc.data(), cmx, cmy};
    for (const Rect& r : P.pdew)
      for (int j = r.d2; j <= r.d3; ++j)
        for (int i = r.d0; i <= r.d1; ++i) PDeW(i, j) = par.P[r.q].DeW(r.s0 + (i - r.d0), r.s2 + (j - r.d2));
    for (const Rect& r : P.pdns)
      for (int j = r.d2; j <= r.d3; ++j)
        for (int i = r.d0; i <= r.d1; ++i) PDnS(i, j) = par.P[r.q].DnS(r.s0 + (i - r.d0), r.s2 + (j - r.d2));
    for (const Rect& r : P.pcc)
      for (int j = r.d2; j <= r.d3; ++j)
        for (int i = r.d0; i <= r.d1; ++i) PCC(i, j) = par.P[r.q].CC(r.s0 + (i - r.d0), r.s2 + (j - r.d2));
    // loadC = restrict2(F - L_h u) + L_2h(QC)
    std::vector<double> lc((size_t)cmx * cmy);
    for (int cj = 0; cj < cmy; ++cj)
      for (int ci = 0; ci < cmx; ++ci) {
        const int i = 2 * ci, j = 2 * cj;
        const double r00 = P.F(i, j) - op2d_at(P.Q, P.DeW, P.DnS, P.CC, i, j, hx, hy);
        const double r10 = P.F(i + 1, j) - op2d_at(P.Q, P.DeW, P.DnS, P.CC, i + 1, j, hx, hy);
        const double r01 = P.F(i, j + 1) - op2d_at(P.Q, P.DeW, P.DnS, P.CC, i, j + 1, hx, hy);
        const double r11 = P.F(i + 1, j + 1) - op2d_at(P.Q, P.DeW, P.DnS, P.CC, i + 1, j + 1, hx, hy);
        const double R = 0.25 * (((r00 + r10) + r01) + r11);
        lc[ci + (size_t)cj * cmx] = R + op2d_at(P.QC, PDeW, PDnS, PCC, ci, cj, hcx, hcy);
      }
    for (const Rect& r : P.fsc)
      for (int j = r.d2; j <= r.d3; ++j)
        for (int i = r.d0; i <= r.d1; ++i) par.P[r.q].F(i, j) = lc[(r.s0 + (i - r.d0)) + (size_t)(r.s2 + (j - r.d2)) * cmx];
    if (H.mass) face_corrections(H, k, P, &vals[(size_t)p * 4]);
  }
  if (!H.mass) return;
  for (int p = 0; p < np; ++p)  // the reference's (patch, face, entry) order
    for (int f = 0; f < 4; ++f)
      for (const Corr& c : lev.P[p].corr[f])
        par.P[c.q].F(c.di, c.dj) = par.P[c.q].F(c.di, c.dj) + vals[(size_t)p * 4 + f][c.src];
}

void coarsest_solve(Hier& H) {
  Patch& P = H.lev(1).P[0];
  const int N = H.N, M = H.M;
  std::vector<double> rhs(M, 0.0);
  for (int t = 0; t < N; ++t) rhs[t] = P.F.a[t] + H.bcRhs[t];
  // u_i = sum_j Ainv(i, j) rhs_j accumulated in ascending j (column-major Ainv)
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) {
    double acc = 0.0;
    for (int j = 0; j < M; ++j) acc = acc + H.Ainv[i + (size_t)j * M] * rhs[j];
    P.Q(i % P.mx + 1, i / P.mx + 1) = acc;
  }
  fill_ghosts(H, 1);
}

void vcycle(Hier& H, int k) {
  if (k == 1) {
    coarsest_solve(H);
    return;
  }
  relax_level(H, k, H.npre);
  restrict_level(H, k);
  fill_ghosts(H, k - 1);
  ring_level(H, k);
  coarse_load(H, k);
  vcycle(H, k - 1);
  correct_level(H, k);
  fill_ghosts(H, k);
  relax_level(H, k, H.npost);
}

struct Res {
  double linf, l2sq, fnorm;
  long long ncell;
};

Res comp_residual(Hier& H) {
  const int r0 = H.root, nlev = (int)H.L.size();
  // chunks: per target level and patch, (di, dj, value) in the reference's order
  std::vector<std::vector<std::vector<Corr>>> tq(nlev + 1);
  std::vector<std::vector<std::vector<double>>> tv(nlev + 1);
  for (int k = r0; k <= nlev; ++k) {
    tq[k].assign(H.lev(k).P.size(), {});
    tv[k].assign(H.lev(k).P.size(), {});
  }
  if (H.mass) {
    for (int k = std::max(2, r0 + 1); k <= nlev; ++k) {
      ring_level(H, k);
      Level& lev = H.lev(k);
#pragma omp parallel for schedule(dynamic)
      for (int p = 0; p < (int)lev.P.size(); ++p) {
        Patch& P = lev.P[p];
        for (int cj = 0; cj < P.my / 2; ++cj)
          for (int ci = 0; ci < P.mx / 2; ++ci) P.QC(ci + 1, cj + 1) = restrict_at(P.Q, ci, cj);
      }
    }
    for (int k = r0 + 1; k <= nlev; ++k) {
      Level& lev = H.lev(k);
      for (size_t p = 0; p < lev.P.size(); ++p) {
        std::vector<double> vals[4];
        face_corrections(H, k, lev.P[p], vals);
        for (int f = 0; f < 4; ++f)
          for (const Corr& c : lev.P[p].corr[f]) {
            tq[k - 1][c.q].push_back(c);
            tv[k - 1][c.q].push_back(vals[f][c.src]);
          }
      }
    }
  }
  Res out{0.0, 0.0, 0.0, 0};
  for (int k = r0; k <= nlev; ++k) {
    Level& lev = H.lev(k);
    const double area = lev.h[0] * lev.h[1];
    const int np = (int)lev.P.size();
    std::vector<Res> part(np);
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < np; ++p) {
      Patch& P = lev.P[p];
      const int mx = P.mx, my = P.my;
      std::vector<double> r((size_t)mx * my);
      for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i)
          r[i + (size_t)j * mx] = P.FT(i, j) - op2d_at(P.Q, P.DeW, P.DnS, P.CC, i, j, lev.h[0], lev.h[1]);
      for (size_t t = 0; t < tq[k][p].size(); ++t) {
        const Corr& c = tq[k][p][t];
        r[c.di + (size_t)c.dj * mx] = r[c.di + (size_t)c.dj * mx] + tv[k][p][t];
      }
      Res s{0.0, 0.0, 0.0, 0};
      for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i)
          if (!P.cov[i + (size_t)j * mx]) {
            const double e = r[i + (size_t)j * mx];
            s.linf = std::max(s.linf, std::fabs(e));
            s.l2sq += e * e;
            s.fnorm = std::max(s.fnorm, std::fabs(P.FT(i, j)));
            ++s.ncell;
          }
      part[p] = s;
    }
    for (int p = 0; p < np; ++p)
      if (part[p].ncell) {
        out.linf = std::max(out.linf, part[p].linf);
        out.l2sq = out.l2sq + part[p].l2sq * area;
        out.fnorm = std::max(out.fnorm, part[p].fnorm);
        out.ncell += part[p].ncell;
      }
  }
  out.fnorm = std::max(out.fnorm, 2.2250738585072014e-308);
  return out;
}

}  // namespace

// ---------------------------------------------------------------- C interface
// The flattened hierarchy (oracle/bsam_cpu.py builds it from an
// OracleHierarchy).  Arrays are column-major (i along x).  Per patch, in
// level order:
//   pint  15 ints:  mx my par0 code[4] faceInternal[4] cornerClass[4]
//   poff  17 offsets into data (cov: into cov): Q F FTrue DeW DnS CC InvDiagS
//                   QC g[4] cfD[4] covered
//   pcnt  12 counts: same ring winrect usc fsc pcc pdew pdns corr[4]
//   desc  the descriptor entries in that order: same/ring 5 ints
//         (q, si, sj, di, dj), rectangles 9 ints (q, src i0 i1 j0 j1,
//         dst i0 i1 j0 j1), corr 4 ints (q, src, di, dj)
extern "C" {

struct CpuDesc {
  int nlev, root, npre, npost, mass, qif, N, M, singular;
  double omega;
  const int* lev_n;        // [2 nlev]
  const double* lev_h;     // [2 nlev]
  const int* lev_np;       // [nlev]
  const int* pint;
  const long long* poff;
  const int* pcnt;
  const long long* pdesc;  // per patch: start of its entries in desc
  const int* desc;
  double* data;            // every array; Q, F and QC are updated in place
  const unsigned char* cov;
  const double* Ainv;
  const double* bcRhs;
};

int bsam_cpu_solve(const CpuDesc* d, int max_cycles, double rel_tol, int threads, double* hist, int hist_cap,
                   int* ncycles, double* final_rel) {
  if (threads > 0) omp_set_num_threads(threads);
  Hier H;
  H.root = d->root; H.npre = d->npre; H.npost = d->npost; H.mass = d->mass; H.qif = d->qif; H.omega = d->omega;
  H.Ainv = d->Ainv; H.bcRhs = d->bcRhs; H.N = d->N; H.M = d->M; H.singular = d->singular;
  H.L.resize(d->nlev);
  long long pi = 0;
  for (int k = 0; k < d->nlev; ++k) {
    Level& L = H.L[k];
    L.n[0] = d->lev_n[2 * k]; L.n[1] = d->lev_n[2 * k + 1];
    L.h[0] = d->lev_h[2 * k]; L.h[1] = d->lev_h[2 * k + 1];
    L.P.resize(d->lev_np[k]);
    for (int p = 0; p < d->lev_np[k]; ++p, ++pi) {
      Patch& P = L.P[p];
      const int* ip = d->pint + pi * 15;
      P.mx = ip[0]; P.my = ip[1]; P.par0 = ip[2];
      for (int s = 0; s < 4; ++s) { P.code[s] = ip[3 + s]; P.fint[s] = ip[7 + s]; P.ccl[s] = ip[11 + s]; }
      const long long* of = d->poff + pi * 17;
      const int mx = P.mx, my = P.my;
      P.Q = A2d{d->data + of[0], mx + 2, my + 2};
      P.F = A2d{d->data + of[1], mx, my};
      P.FT = A2d{d->data + of[2], mx, my};
      P.DeW = A2d{d->data + of[3], mx + 1, my};
      P.DnS = A2d{d->data + of[4], mx, my + 1};
      P.CC = A2d{d->data + of[5], mx, my};
      P.ID = A2d{d->data + of[6], mx, my};
      P.QC = A2d{d->data + of[7], mx / 2 + 2, my / 2 + 2};
      for (int s = 0; s < 4; ++s) P.g[s] = d->data + of[8 + s];
      for (int s = 0; s < 4; ++s) P.cfD[s] = d->data + of[12 + s];
      P.cov = d->cov + of[16];
      const int* cnt = d->pcnt + pi * 12;
      const int* ds = d->desc + d->pdesc[pi];
      auto pts = [&](int n, std::vector<Pt>& v) {
        for (int t = 0; t < n; ++t, ds += 5) v.push_back(Pt{ds[0], ds[1], ds[2], ds[3], ds[4]});
      };
      auto rects = [&](int n, std::vector<Rect>& v) {
        for (int t = 0; t < n; ++t, ds += 9) v.push_back(Rect{ds[0], ds[1], ds[2], ds[3], ds[4], ds[5], ds[6], ds[7], ds[8]});
      };
      pts(cnt[0], P.same);
      pts(cnt[1], P.ring);
      rects(cnt[2], P.winrect);
      rects(cnt[3], P.usc);
      rects(cnt[4], P.fsc);
      rects(cnt[5], P.pcc);
      rects(cnt[6], P.pdew);
      rects(cnt[7], P.pdns);
      for (int f = 0; f < 4; ++f)
        for (int t = 0; t < cnt[8 + f]; ++t, ds += 4) P.corr[f].push_back(Corr{ds[0], ds[1], ds[2], ds[3]});
    }
  }
  // ebSolveCycles.m:11-19
  for (int k = 1; k <= d->nlev; ++k) fill_ghosts(H, k);
  Res r = comp_residual(H);
  int it = 0, nh = 0;
  double rel = r.linf / r.fnorm;
  if (hist && nh < hist_cap) hist[nh++] = r.linf;
  while (it < max_cycles && rel > rel_tol) {
    ++it;
    for (Level& L : H.L)
      for (Patch& P : L.P) std::memcpy(P.F.a, P.FT.a, sizeof(double) * (size_t)P.mx * P.my);
    vcycle(H, d->nlev);
    filldown(H);
    r = comp_residual(H);
    rel = r.linf / r.fnorm;
    if (hist && nh < hist_cap) hist[nh++] = r.linf;
  }
  if (ncycles) *ncycles = it;
  if (final_rel) *final_rel = rel;
  return 0;
}

}  // extern "C"
