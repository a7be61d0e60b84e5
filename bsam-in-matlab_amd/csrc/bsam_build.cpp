// Host-side hierarchy construction of libbsamgpu: the sub-root ladder and
// root (EllipticBSAM/src/ebSetup.m:10-68), level sampling from the caller's
// coefficient callbacks (ebMakeLevel.m:1-105), the coefficient fill-down
// (ebCoeffFilldown.m:1-62), the folded inverse diagonal (ebKernels.m:28-34,
// ebBcFold.m:12-42), covered masks, the source projection of singular
// problems (ebCompatProject.m:1-111) and the addition of a refined level
// (ebAddLevel.m:10-24).  It produces the bsam_hier_desc that bsam_hier_create
// takes, so a caller without MATLAB or Python (the plain C driver, a serving
// loop) builds and refines hierarchies natively.  Host code: needs no device.
//
// Every array is computed with the expressions, and in the order, of the
// reference (and of the Python host layer that restates it), so the same
// coefficient values give the same descriptor bit for bit.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bsamgpu.h"

// the thread's last error message (bsam_lib.cpp)
int bsam_set_error(int code, const char* msg);

namespace {

struct BBox {
  int ilo, ihi, jlo, jhi;
};

struct BLevel {
  int n[2] = {0, 0};
  double h[2] = {0, 0};
  std::vector<BBox> box;
  // per patch, column-major (x index fastest)
  std::vector<std::vector<double>> Q, FT, DeW, DnS, CC, ID;
  std::vector<std::array<int, 4>> code;                  // W E S N: 0 internal, 1 Dirichlet, 2 Neumann
  std::vector<std::array<std::vector<double>, 4>> g;     // boundary data (zeros on internal faces)
  std::vector<std::vector<unsigned char>> covered;
  double shift = 0.0;                                    // sourceShift (ebCompatProject)
  int mx(int p) const { return box[p].ihi - box[p].ilo + 1; }
  int my(int p) const { return box[p].jhi - box[p].jlo + 1; }
};

int err(int code, const std::string& msg) { return bsam_set_error(code, msg.c_str()); }

// ebMakeLevel.m:95-105 (evalgrid): the callback's values at n points, finite
int eval(bsam_field_fn fn, void* user, const std::vector<double>& x, const std::vector<double>& y,
         std::vector<double>& out) {
  out.assign(x.size(), 0.0);
  if (x.empty()) return 0;
  fn(user, x.data(), y.data(), (long long)x.size(), out.data());
  for (double v : out)
    if (!std::isfinite(v))
      return err(BSAM_E_NONFINITE, "eb_make_level:eval: coefficient function returned non-finite values.");
  return 0;
}

}  // namespace

struct bsam_build {
  bsam_problem prob{};
  bsam_build_opts opts{};
  std::vector<BLevel> L;  // index 0 = level 1
  int root = 0;
  bool singular = false;
  // the packed descriptor of the last bsam_build_desc call
  std::vector<int> dn, dnp, dbox, dbc;
  std::vector<double> dQ, dFT, dDeW, dDnS, dCC, dID, dG;
};

namespace {

// invdiag2d (ebKernels.m:28-34) and the physical-boundary fold
// (ebBcFold.m:12-42): InvDiagS = 1 ./ (1 ./ ID + fold)
void refresh_invdiag(BLevel& lv, int p) {
  const int mx = lv.mx(p), my = lv.my(p);
  const double hx = lv.h[0], hy = lv.h[1];
  const double hx2 = std::pow(hx, 2.0), hy2 = std::pow(hy, 2.0);
  const std::vector<double>& DeW = lv.DeW[p];
  const std::vector<double>& DnS = lv.DnS[p];
  const std::array<int, 4>& c = lv.code[p];
  std::vector<double>& ID = lv.ID[p];
  ID.assign((size_t)mx * my, 0.0);
  for (int j = 0; j < my; ++j)
    for (int i = 0; i < mx; ++i) {
      const double den = (DeW[(i + 1) + j * (mx + 1)] + DeW[i + j * (mx + 1)]) / (hx * hx) +
                         (DnS[i + (j + 1) * mx] + DnS[i + j * mx]) / (hy * hy) + lv.CC[p][i + j * mx];
      const double id = 1.0 / den;
      double add = 0.0;
      if (i == 0) {
        if (c[0] == 1) add = add + 2 * DeW[0 + j * (mx + 1)] / hx2;
        else if (c[0] == 2) add = add - DeW[0 + j * (mx + 1)] / hx2;
      }
      if (i == mx - 1) {
        if (c[1] == 1) add = add + 2 * DeW[mx + j * (mx + 1)] / hx2;
        else if (c[1] == 2) add = add - DeW[mx + j * (mx + 1)] / hx2;
      }
      if (j == 0) {
        if (c[2] == 1) add = add + 2 * DnS[i] / hy2;
        else if (c[2] == 2) add = add - DnS[i] / hy2;
      }
      if (j == my - 1) {
        if (c[3] == 1) add = add + 2 * DnS[i + my * mx] / hy2;
        else if (c[3] == 2) add = add - DnS[i + my * mx] / hy2;
      }
      ID[i + j * mx] = 1.0 / (1.0 / id + add);
    }
}

// ebMakeLevel.m:1-75 for level of n cells with the given boxes
int make_level(const bsam_problem& pr, const int n[2], const std::vector<BBox>& boxes, BLevel& lv) {
  const double* dom = pr.domain;
  lv = BLevel();
  lv.n[0] = n[0];
  lv.n[1] = n[1];
  lv.h[0] = (dom[1] - dom[0]) / n[0];
  lv.h[1] = (dom[3] - dom[2]) / n[1];
  const double hx = lv.h[0], hy = lv.h[1];
  for (const BBox& b : boxes) {
    if (b.ilo < 1 || b.ilo > n[0] || b.ihi < 1 || b.ihi > n[0] || b.jlo < 1 || b.jlo > n[1] || b.jhi < 1 ||
        b.jhi > n[1] || b.ilo > b.ihi || b.jlo > b.jhi)
      return err(BSAM_E_BOX, "eb_make_level:box: invalid patch box.");
  }
  lv.box = boxes;
  const int np = (int)boxes.size();
  lv.Q.resize(np); lv.FT.resize(np); lv.DeW.resize(np); lv.DnS.resize(np); lv.CC.resize(np); lv.ID.resize(np);
  lv.code.resize(np); lv.g.resize(np); lv.covered.resize(np);
  for (int p = 0; p < np; ++p) {
    const BBox& b = boxes[p];
    const int mx = b.ihi - b.ilo + 1, my = b.jhi - b.jlo + 1;
    lv.Q[p].assign((size_t)(mx + 2) * (my + 2), 0.0);
    std::vector<double> xf(mx + 1), yc(my), xc(mx), yf(my + 1);
    for (int t = 0; t <= mx; ++t) xf[t] = dom[0] + (double)(b.ilo - 1 + t) * hx;
    for (int t = 0; t < my; ++t) yc[t] = dom[2] + ((double)(b.jlo + t) - 0.5) * hy;
    for (int t = 0; t < mx; ++t) xc[t] = dom[0] + ((double)(b.ilo + t) - 0.5) * hx;
    for (int t = 0; t <= my; ++t) yf[t] = dom[2] + (double)(b.jlo - 1 + t) * hy;
    std::vector<double> X, Y;
    auto grid = [&](const std::vector<double>& xs, const std::vector<double>& ys) {
      X.clear(); Y.clear();
      for (double yv : ys)
        for (double xv : xs) { X.push_back(xv); Y.push_back(yv); }
    };
    int rc;
    grid(xf, yc);  // D at x-faces (ndgrid(xf, yc))
    if ((rc = eval(pr.D, pr.user, X, Y, lv.DeW[p]))) return rc;
    grid(xc, yf);  // D at y-faces
    if ((rc = eval(pr.D, pr.user, X, Y, lv.DnS[p]))) return rc;
    grid(xc, yc);  // C and f at cell centres (midpoint source)
    if ((rc = eval(pr.C, pr.user, X, Y, lv.CC[p]))) return rc;
    if ((rc = eval(pr.f, pr.user, X, Y, lv.FT[p]))) return rc;
    for (double v : lv.DeW[p])
      if (v <= 0) return err(BSAM_E_D, "eb_make_level:D: D(x,y) must be strictly positive at faces.");
    for (double v : lv.DnS[p])
      if (v <= 0) return err(BSAM_E_D, "eb_make_level:D: D(x,y) must be strictly positive at faces.");
    // physical faces and their boundary data (ebMakeLevel.m:50-57, bcSide :78-92)
    const bool fint[4] = {b.ilo > 1, b.ihi < n[0], b.jlo > 1, b.jhi < n[1]};
    for (int s = 0; s < 4; ++s) {
      const int len = s < 2 ? my : mx;
      lv.g[p][s].assign(len, 0.0);
      lv.code[p][s] = 0;
      if (fint[s]) continue;
      lv.code[p][s] = pr.bc_type[s] == 1 ? 1 : 2;
      if (!pr.g[s]) continue;
      std::vector<double> gx(len), gy(len);
      for (int t = 0; t < len; ++t) {
        gx[t] = s == 0 ? dom[0] : s == 1 ? dom[1] : xc[t];
        gy[t] = s == 2 ? dom[2] : s == 3 ? dom[3] : yc[t];
      }
      if ((rc = eval(pr.g[s], pr.user, gx, gy, lv.g[p][s]))) return rc;
    }
    lv.covered[p].assign((size_t)mx * my, 0);
    refresh_invdiag(lv, p);
  }
  return 0;
}

bool isect(const BBox& a, const BBox& b, BBox& r) {
  r = BBox{std::max(a.ilo, b.ilo), std::min(a.ihi, b.ihi), std::max(a.jlo, b.jlo), std::min(a.jhi, b.jhi)};
  return r.ilo <= r.ihi && r.jlo <= r.jhi;
}

// ebCoeffFilldown.m:1-62: fine face / cell averages onto the covered parent
// faces and cells, finest level first, then every level's inverse diagonal
void coeff_filldown(bsam_build& B) {
  for (int k = (int)B.L.size(); k >= 2; --k) {
    BLevel& F = B.L[k - 1];
    BLevel& C = B.L[k - 2];
    for (size_t p = 0; p < F.box.size(); ++p) {
      const BBox& b = F.box[p];
      const int Il = (b.ilo + 1) / 2, Ih = b.ihi / 2, Jl = (b.jlo + 1) / 2, Jh = b.jhi / 2;
      const int mx = F.mx((int)p);
      const int cmx = Ih - Il + 1, cmy = Jh - Jl + 1;
      const std::vector<double>& DeW = F.DeW[p];
      const std::vector<double>& DnS = F.DnS[p];
      const std::vector<double>& CC = F.CC[p];
      auto cDeW = [&](int ii, int jj) {  // coarse x-face ii of coarse row jj
        return 0.5 * (DeW[2 * ii + (2 * jj) * (mx + 1)] + DeW[2 * ii + (2 * jj + 1) * (mx + 1)]);
      };
      auto cDnS = [&](int ii, int jj) {  // coarse y-face jj of coarse column ii
        return 0.5 * (DnS[2 * ii + (2 * jj) * mx] + DnS[2 * ii + 1 + (2 * jj) * mx]);
      };
      auto cCC = [&](int ii, int jj) {  // restrict2
        const int a = 2 * ii + 2 * jj * mx;
        return 0.25 * (((CC[a] + CC[a + 1]) + CC[a + mx]) + CC[a + mx + 1]);
      };
      (void)cmx; (void)cmy;
      for (size_t q = 0; q < C.box.size(); ++q) {
        const BBox& qb = C.box[q];
        const int qmx = C.mx((int)q);
        BBox r;
        if (isect(BBox{Il, Ih, Jl, Jh}, qb, r))
          for (int J = r.jlo; J <= r.jhi; ++J)
            for (int I = r.ilo; I <= r.ihi; ++I) C.CC[q][(I - qb.ilo) + (J - qb.jlo) * qmx] = cCC(I - Il, J - Jl);
        if (isect(BBox{Il, Ih + 1, Jl, Jh}, BBox{qb.ilo, qb.ihi + 1, qb.jlo, qb.jhi}, r))
          for (int J = r.jlo; J <= r.jhi; ++J)
            for (int I = r.ilo; I <= r.ihi; ++I)
              C.DeW[q][(I - qb.ilo) + (J - qb.jlo) * (qmx + 1)] = cDeW(I - Il, J - Jl);
        if (isect(BBox{Il, Ih, Jl, Jh + 1}, BBox{qb.ilo, qb.ihi, qb.jlo, qb.jhi + 1}, r))
          for (int J = r.jlo; J <= r.jhi; ++J)
            for (int I = r.ilo; I <= r.ihi; ++I) C.DnS[q][(I - qb.ilo) + (J - qb.jlo) * qmx] = cDnS(I - Il, J - Jl);
      }
    }
  }
  for (BLevel& lv : B.L)
    for (size_t p = 0; p < lv.box.size(); ++p) refresh_invdiag(lv, (int)p);
}

// covered masks of level k-1 from the boxes of level k (ebBuildOps.m:171-189)
void covered_masks(bsam_build& B, int k) {
  BLevel& F = B.L[k - 1];
  BLevel& C = B.L[k - 2];
  for (size_t q = 0; q < C.box.size(); ++q) {
    const BBox& qb = C.box[q];
    const int qmx = C.mx((int)q);
    std::vector<unsigned char>& cov = C.covered[q];
    std::fill(cov.begin(), cov.end(), 0);
    for (const BBox& b : F.box) {
      BBox r;
      if (isect(BBox{(b.ilo + 1) / 2, b.ihi / 2, (b.jlo + 1) / 2, b.jhi / 2}, qb, r))
        for (int J = r.jlo; J <= r.jhi; ++J)
          for (int I = r.ilo; I <= r.ihi; ++I) cov[(I - qb.ilo) + (J - qb.jlo) * qmx] = 1;
    }
  }
}

int check_alignment(const bsam_build& B, int k) {
  for (const BBox& b : B.L[k - 1].box)
    if (b.ilo % 2 != 1 || b.ihi % 2 != 0 || b.jlo % 2 != 1 || b.jhi % 2 != 0)
      return err(BSAM_E_ALIGN, "eb_build_ops:align: patch not aligned to parent cells.");
  return 0;
}

// ebCoarsest.m:58-66: all-Neumann level 1 and C == 0 everywhere
bool is_singular(const bsam_build& B) {
  bool allneu = true;
  for (int s = 0; s < 4; ++s) allneu = allneu && B.L[0].code[0][s] != 1;
  double cmax = 0;
  for (const BLevel& lv : B.L)
    for (const std::vector<double>& cc : lv.CC)
      for (double v : cc) cmax = std::max(cmax, std::fabs(v));
  return allneu && cmax == 0;
}

// ebCompatProject.m:1-111.  Non-singular problems undo a previous shift; the
// singular all-Neumann problem shifts FTrue so that the visible source plus
// the boundary flux integrates to zero.  The sums run cell by cell in column
// order (the Python host layer uses NumPy's pairwise sums, so for singular
// problems the two agree to rounding, not bit for bit).
void compat_project(bsam_build& B) {
  if (!B.singular) {
    for (BLevel& lv : B.L) {
      if (lv.shift == 0) continue;
      for (std::vector<double>& ft : lv.FT)
        for (double& v : ft) v = v + lv.shift;
      lv.shift = 0.0;
    }
    return;
  }
  double S = 0.0, Av = 0.0, R = 0.0;
  for (int k = B.root; k <= (int)B.L.size(); ++k) {
    BLevel& lv = B.L[k - 1];
    const double old = lv.shift, hx = lv.h[0], hy = lv.h[1], da = hx * hy;
    for (size_t p = 0; p < lv.box.size(); ++p) {
      const int mx = lv.mx((int)p), my = lv.my((int)p);
      const std::vector<unsigned char>& cov = lv.covered[p];
      double s = 0.0;
      long long nvis = 0;
      for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i)
          if (!cov[i + j * mx]) { s += lv.FT[p][i + j * mx]; ++nvis; }
      S = S + (s + old * (double)nvis) * da;
      Av = Av + (double)nvis * da;
      const std::array<int, 4>& c = lv.code[p];
      const std::vector<double>& DeW = lv.DeW[p];
      const std::vector<double>& DnS = lv.DnS[p];
      double fl = 0.0;
      if (c[0] == 2) { fl = 0; for (int j = 0; j < my; ++j) if (!cov[0 + j * mx]) fl += DeW[0 + j * (mx + 1)] * lv.g[p][0][j]; R = R + fl * hy; }
      if (c[1] == 2) { fl = 0; for (int j = 0; j < my; ++j) if (!cov[mx - 1 + j * mx]) fl += DeW[mx + j * (mx + 1)] * lv.g[p][1][j]; R = R + fl * hy; }
      if (c[2] == 2) { fl = 0; for (int i = 0; i < mx; ++i) if (!cov[i]) fl += DnS[i] * lv.g[p][2][i]; R = R + fl * hx; }
      if (c[3] == 2) { fl = 0; for (int i = 0; i < mx; ++i) if (!cov[i + (my - 1) * mx]) fl += DnS[i + my * mx] * lv.g[p][3][i]; R = R + fl * hx; }
    }
  }
  const double cshift = (S + R) / std::max(Av, 2.2250738585072014e-308);
  for (BLevel& lv : B.L) {
    const double old = lv.shift;
    for (std::vector<double>& ft : lv.FT)
      for (double& v : ft) v = v - (cshift - old);
    lv.shift = cshift;
  }
}

// ebSetup.m:14-25: the sizes of levels 1..root (sub-root ladder, coarsest first)
std::vector<std::array<int, 2>> ladder(const bsam_build_opts& o) {
  std::vector<std::array<int, 2>> s{{o.base_cells[0], o.base_cells[1]}};
  std::array<int, 2> ns{o.base_cells[0], o.base_cells[1]};
  while (ns[0] % 2 == 0 && ns[1] % 2 == 0 && std::min(ns[0], ns[1]) / 2.0 >= 4 &&
         std::min(ns[0], ns[1]) > o.coarsest_cells) {
    ns[0] /= 2;
    ns[1] /= 2;
    s.insert(s.begin(), ns);
  }
  return s;
}

}  // namespace

extern "C" {

int bsam_build_create(const bsam_problem* prob, const bsam_build_opts* opts, bsam_build** out) {
  if (!prob || !opts || !out || !prob->D || !prob->C || !prob->f)
    return err(BSAM_E_ARG, "bsamgpu:arg: bsam_build_create needs a problem with D, C and f, options and an output");
  *out = nullptr;
  if (opts->base_cells[0] < 1 || opts->base_cells[1] < 1 || opts->coarsest_cells < 1)
    return err(BSAM_E_ARG, "bsamgpu:arg: BaseCells and CoarsestCells must be positive");
  if (opts->source_quad > 1)
    return err(BSAM_E_ARG, "bsamgpu:arg: only the midpoint source (SourceQuad <= 1) is built natively");
  for (int s = 0; s < 4; ++s)
    if (prob->bc_type[s] != 1 && prob->bc_type[s] != 2)
      return err(BSAM_E_ARG, "eb_problem:bc: bc type must be dirichlet (1) or neumann (2)");
  bsam_build* B = new bsam_build();
  B->prob = *prob;
  B->opts = *opts;
  const std::vector<std::array<int, 2>> sizes = ladder(*opts);
  B->root = (int)sizes.size();
  B->L.resize(sizes.size());
  for (size_t k = 0; k < sizes.size(); ++k) {
    const int n[2] = {sizes[k][0], sizes[k][1]};
    if (int rc = make_level(B->prob, n, {BBox{1, n[0], 1, n[1]}}, B->L[k])) {
      delete B;
      return rc;
    }
  }
  coeff_filldown(*B);
  for (int k = 2; k <= (int)B->L.size(); ++k) {
    if (int rc = check_alignment(*B, k)) {
      delete B;
      return rc;
    }
    covered_masks(*B, k);
  }
  if (prob->u0)  // InitialGuess at the cell centres of the ladder and root (ebSetup.m)
    for (BLevel& lv : B->L)
      for (size_t p = 0; p < lv.box.size(); ++p) {
        const BBox& b = lv.box[p];
        const int mx = lv.mx((int)p), my = lv.my((int)p);
        std::vector<double> X, Y, u;
        for (int j = 0; j < my; ++j)
          for (int i = 0; i < mx; ++i) {
            X.push_back(prob->domain[0] + ((double)(b.ilo + i) - 0.5) * lv.h[0]);
            Y.push_back(prob->domain[2] + ((double)(b.jlo + j) - 0.5) * lv.h[1]);
          }
        if (int rc = eval(prob->u0, prob->user, X, Y, u)) {
          delete B;
          return rc;
        }
        for (int j = 0; j < my; ++j)
          for (int i = 0; i < mx; ++i) lv.Q[p][(i + 1) + (j + 1) * (mx + 2)] = u[i + j * mx];
      }
  B->singular = is_singular(*B);
  compat_project(*B);
  *out = B;
  return 0;
}

void bsam_build_destroy(bsam_build* b) { delete b; }

int bsam_build_add_level(bsam_build* B, const int* boxes, int nboxes) {
  if (!B || !boxes || nboxes < 1) return err(BSAM_E_ARG, "bsamgpu:arg: bsam_build_add_level needs boxes");
  const BLevel& top = B->L.back();
  const int n[2] = {top.n[0] * 2, top.n[1] * 2};
  std::vector<BBox> fine;
  for (int i = 0; i < nboxes; ++i) {
    const int* b = boxes + 4 * i;
    fine.push_back(BBox{2 * b[0] - 1, 2 * b[1], 2 * b[2] - 1, 2 * b[3]});
  }
  BLevel lv;
  if (int rc = make_level(B->prob, n, fine, lv)) return rc;
  B->L.push_back(std::move(lv));
  const int k = (int)B->L.size();
  coeff_filldown(*B);
  if (int rc = check_alignment(*B, k)) {
    B->L.pop_back();
    coeff_filldown(*B);
    return rc;
  }
  covered_masks(*B, k);
  B->singular = is_singular(*B);
  compat_project(*B);
  return 0;
}

int bsam_build_info(const bsam_build* B, int* nlev, int* root_idx, int* singular) {
  if (!B) return err(BSAM_E_ARG, "bsamgpu:arg: null builder");
  if (nlev) *nlev = (int)B->L.size();
  if (root_idx) *root_idx = B->root;
  if (singular) *singular = B->singular ? 1 : 0;
  return 0;
}

int bsam_build_desc(bsam_build* B, bsam_hier_desc* d) {
  if (!B || !d) return err(BSAM_E_ARG, "bsamgpu:arg: null builder or descriptor");
  B->dn.clear(); B->dnp.clear(); B->dbox.clear(); B->dbc.clear();
  B->dQ.clear(); B->dFT.clear(); B->dDeW.clear(); B->dDnS.clear(); B->dCC.clear(); B->dID.clear(); B->dG.clear();
  for (const BLevel& lv : B->L) {
    B->dn.push_back(lv.n[0]);
    B->dn.push_back(lv.n[1]);
    B->dnp.push_back((int)lv.box.size());
    for (size_t p = 0; p < lv.box.size(); ++p) {
      const BBox& b = lv.box[p];
      B->dbox.insert(B->dbox.end(), {b.ilo, b.ihi, b.jlo, b.jhi});
      B->dbc.insert(B->dbc.end(), lv.code[p].begin(), lv.code[p].end());
      B->dQ.insert(B->dQ.end(), lv.Q[p].begin(), lv.Q[p].end());
      B->dFT.insert(B->dFT.end(), lv.FT[p].begin(), lv.FT[p].end());
      B->dDeW.insert(B->dDeW.end(), lv.DeW[p].begin(), lv.DeW[p].end());
      B->dDnS.insert(B->dDnS.end(), lv.DnS[p].begin(), lv.DnS[p].end());
      B->dCC.insert(B->dCC.end(), lv.CC[p].begin(), lv.CC[p].end());
      B->dID.insert(B->dID.end(), lv.ID[p].begin(), lv.ID[p].end());
      for (int s = 0; s < 4; ++s) B->dG.insert(B->dG.end(), lv.g[p][s].begin(), lv.g[p][s].end());
    }
  }
  std::memset(d, 0, sizeof *d);
  d->nlev = (int)B->L.size();
  d->root_idx = B->root;
  for (int i = 0; i < 4; ++i) d->domain[i] = B->prob.domain[i];
  d->n = B->dn.data(); d->np = B->dnp.data(); d->boxes = B->dbox.data(); d->bc_code = B->dbc.data();
  d->Q = B->dQ.data(); d->FTrue = B->dFT.data(); d->DeW = B->dDeW.data(); d->DnS = B->dDnS.data();
  d->CC = B->dCC.data(); d->InvDiagS = B->dID.data(); d->g = B->dG.data();
  return 0;
}

int bsam_build_allowed(const bsam_build* B, unsigned char* out) {
  if (!B || !out) return err(BSAM_E_ARG, "bsamgpu:arg: null builder or output");
  const BLevel& lv = B->L.back();
  const int n1 = lv.n[0], n2 = lv.n[1];
  std::vector<unsigned char> cov((size_t)n1 * n2, 0);
  for (const BBox& b : lv.box)
    for (int j = b.jlo; j <= b.jhi; ++j)
      for (int i = b.ilo; i <= b.ihi; ++i) cov[(size_t)(i - 1) + (size_t)(j - 1) * n1] = 1;
  // erode8 (ebTagCluster.m:130-141): the cell and every in-domain 8-neighbour covered
  for (int J = 0; J < n2; ++J)
    for (int I = 0; I < n1; ++I) {
      unsigned char a = 1;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int x = I + dx, y = J + dy;
          if (x >= 0 && x < n1 && y >= 0 && y < n2) a = a && cov[(size_t)x + (size_t)y * n1];
        }
      out[(size_t)I + (size_t)J * n1] = a;
    }
  return 0;
}

}  // extern "C"
