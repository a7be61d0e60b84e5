// Host clustering of a level's tag map into patch boxes: Berger-Rigoutsos
// signature splitting, the proper-nesting clip, minimum-width growth and the
// sliver merge of EllipticBSAM/src/ebTagCluster.m:92-380, with MATLAB's
// tie-breaking (first index on min/max, LIFO stack popped from the end,
// column-major scans).  Boxes are 1-based inclusive [ilo ihi jlo jhi] in the
// level's index space.  Host code only, built by the host compiler.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstddef>
#include <limits>
#include <numeric>
#include <vector>

namespace {

struct Box {
  int ilo, ihi, jlo, jhi;
  int w0() const { return ihi - ilo + 1; }
  int w1() const { return jhi - jlo + 1; }
};

struct Map {
  const unsigned char* m;
  int n1, n2;
  bool at(int i, int j) const { return m[(size_t)(i - 1) + (size_t)(j - 1) * n1] != 0; }
  long long count(const Box& b) const {
    long long c = 0;
    for (int j = b.jlo; j <= b.jhi; ++j)
      for (int i = b.ilo; i <= b.ihi; ++i) c += at(i, j);
    return c;
  }
  bool all(const Box& b) const {
    for (int j = b.jlo; j <= b.jhi; ++j)
      for (int i = b.ilo; i <= b.ihi; ++i)
        if (!at(i, j)) return false;
    return true;
  }
};

// ebTagCluster.m:201-208: bounding box of the tags inside bb
bool tight_bbox(const Map& T, const Box& bb, Box& out) {
  int rlo = INT_MAX, rhi = -1, clo = INT_MAX, chi = -1;
  for (int j = bb.jlo; j <= bb.jhi; ++j)
    for (int i = bb.ilo; i <= bb.ihi; ++i)
      if (T.at(i, j)) {
        rlo = std::min(rlo, i);
        rhi = std::max(rhi, i);
        clo = std::min(clo, j);
        chi = std::max(chi, j);
      }
  if (rhi < 0) return false;
  out = Box{rlo, rhi, clo, chi};
  return true;
}

// ebTagCluster.m:211-236: cut position (1-based, the first half's width) and
// its quality, 2 = a zero of the signature, 1 = the strongest inflection
void pick_split(const std::vector<double>& sig, int minW, int& s, int& quality) {
  const int n = (int)sig.size();
  s = 0;
  quality = 0;
  int hole = 0;
  double hd = std::numeric_limits<double>::infinity();
  for (int h = 1; h <= n; ++h)
    if (sig[h - 1] == 0 && h > minW && h <= n - minW) {
      const double d = std::fabs(h - n / 2.0);
      if (d < hd) {
        hd = d;
        hole = h;
      }
    }
  if (hole) {
    s = hole;
    quality = 2;
    return;
  }
  if (n < 2 * minW) return;
  std::vector<double> d2(std::max(0, n - 2));
  for (int m = 0; m + 2 < n; ++m) d2[m] = sig[m + 2] - 2 * sig[m + 1] + sig[m];
  double best = -1.0;
  for (int i = 2; i <= n - 2; ++i) {
    const double j = std::fabs(d2[i - 1] - d2[i - 2]);
    if (i >= minW && n - i >= minW && j > best) {
      best = j;
      s = i;
    }
  }
  if (s > 0) quality = 1;
}

void halves(const Box& bb, bool alongx, int s, std::vector<Box>& stack) {
  if (alongx) {
    stack.push_back(Box{bb.ilo, bb.ilo + s - 1, bb.jlo, bb.jhi});
    stack.push_back(Box{bb.ilo + s, bb.ihi, bb.jlo, bb.jhi});
  } else {
    stack.push_back(Box{bb.ilo, bb.ihi, bb.jlo, bb.jlo + s - 1});
    stack.push_back(Box{bb.ilo, bb.ihi, bb.jlo + s, bb.jhi});
  }
}

// ebTagCluster.m:144-198
std::vector<Box> br_cluster(const Map& T, int minW, double eff) {
  std::vector<Box> out, stack;
  Box all{1, T.n1, 1, T.n2}, bb;
  if (!tight_bbox(T, all, bb)) return out;
  stack.push_back(bb);
  std::vector<double> sigx, sigy;
  while (!stack.empty()) {
    const Box cand = stack.back();
    stack.pop_back();
    if (!tight_bbox(T, cand, bb)) continue;
    const int w0 = bb.w0(), w1 = bb.w1();
    sigx.assign(w0, 0.0);
    sigy.assign(w1, 0.0);
    long long cnt = 0;
    for (int j = bb.jlo; j <= bb.jhi; ++j)
      for (int i = bb.ilo; i <= bb.ihi; ++i)
        if (T.at(i, j)) {
          sigx[i - bb.ilo] += 1;
          sigy[j - bb.jlo] += 1;
          ++cnt;
        }
    const double fill = (double)cnt / ((double)w0 * (double)w1);
    if (fill >= eff || (w0 <= minW && w1 <= minW)) {
      out.push_back(bb);
      continue;
    }
    int sx, qx, sy, qy;
    pick_split(sigx, minW, sx, qx);
    pick_split(sigy, minW, sy, qy);
    if (qx == 0 && qy == 0) {
      if (w0 >= 2 * minW || w1 >= 2 * minW)
        halves(bb, w0 >= w1, w0 >= w1 ? w0 / 2 : w1 / 2, stack);
      else
        out.push_back(bb);
      continue;
    }
    bool usex = qx > qy || (qx == qy && w0 >= w1);
    if (qx == 0) usex = false;
    if (qy == 0) usex = true;
    halves(bb, usex, usex ? sx : sy, stack);
  }
  return out;
}

// ebTagCluster.m:239-264: bisect boxes that leave the allowed region, drop tagless ones
std::vector<Box> clip_allowed(const std::vector<Box>& boxes, const Map& T, const Map& A) {
  std::vector<Box> out, stack(boxes);
  while (!stack.empty()) {
    const Box bb = stack.back();
    stack.pop_back();
    if (T.count(bb) == 0) continue;
    if (A.all(bb)) {
      out.push_back(bb);
      continue;
    }
    const bool alongx = bb.w0() >= bb.w1();
    halves(bb, alongx, alongx ? bb.w0() / 2 : bb.w1() / 2, stack);
  }
  return out;
}

bool overlaps(const Box& a, const Box& b) {
  return a.ilo <= b.ihi && a.ihi >= b.ilo && a.jlo <= b.jhi && a.jhi >= b.jlo;
}

// ebTagCluster.m:359-380: the strip at pos along dim is allowed and free
bool strip_ok(const std::vector<Box>& boxes, const Map& A, int b, int dim, int pos) {
  const Box& bb = boxes[b];
  const Box s = dim == 1 ? Box{pos, pos, bb.jlo, bb.jhi} : Box{bb.ilo, bb.ihi, pos, pos};
  if (!A.all(s)) return false;
  for (int o = 0; o < (int)boxes.size(); ++o)
    if (o != b && overlaps(s, boxes[o])) return false;
  return true;
}

// ebTagCluster.m:267-297: widen boxes below the minimum width a strip at a time
void grow_min(std::vector<Box>& boxes, const Map& A, int minW) {
  for (int b = 0; b < (int)boxes.size(); ++b)
    for (int dim = 1; dim <= 2; ++dim) {
      int lo = dim == 1 ? boxes[b].ilo : boxes[b].jlo;
      int hi = dim == 1 ? boxes[b].ihi : boxes[b].jhi;
      const int nd = dim == 1 ? A.n1 : A.n2;
      int side = 0;
      while (hi - lo + 1 < minW) {
        bool grown = false;
        for (int t = 0; t < 2 && !grown; ++t) {
          const int tryside = t == 0 ? side : 1 - side;
          if (tryside == 0 && lo > 1) {
            if (strip_ok(boxes, A, b, dim, lo - 1)) {
              --lo;
              grown = true;
            }
          } else if (tryside == 1 && hi < nd) {
            if (strip_ok(boxes, A, b, dim, hi + 1)) {
              ++hi;
              grown = true;
            }
          }
        }
        if (!grown) break;
        side = 1 - side;
      }
      if (dim == 1) {
        boxes[b].ilo = lo;
        boxes[b].ihi = hi;
      } else {
        boxes[b].jlo = lo;
        boxes[b].jhi = hi;
      }
    }
}

// ebTagCluster.m:299-356: merge boxes thinner than 2 cells into a touching
// neighbour when the union stays allowed, free and compact; else drop them
int merge_or_drop_thin(std::vector<Box>& boxes, const Map& A) {
  auto thin = [](const Box& b) { return b.w0() < 2 || b.w1() < 2; };
  auto area = [](const Box& b) { return (long long)b.w0() * b.w1(); };
  auto join = [](const Box& a, const Box& b) {
    return Box{std::min(a.ilo, b.ilo), std::max(a.ihi, b.ihi), std::min(a.jlo, b.jlo), std::max(a.jhi, b.jhi)};
  };
  int dropped = 0;
  size_t b = 0;
  while (b < boxes.size()) {
    if (!thin(boxes[b])) {
      ++b;
      continue;
    }
    const size_t nb = boxes.size();
    const Box bx = boxes[b];
    const double areab = (double)area(bx);
    std::vector<double> ua(nb, std::numeric_limits<double>::infinity());
    for (size_t o = 0; o < nb; ++o)
      if (o != b) ua[o] = (double)area(join(bx, boxes[o]));
    std::vector<size_t> order(nb);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return ua[x] < ua[y]; });
    bool merged = false;
    for (size_t o : order) {
      if (o == b || !std::isfinite(ua[o])) continue;
      const Box sb = boxes[b], ob = boxes[o];
      if (!(sb.ilo - 1 <= ob.ihi && sb.ihi + 1 >= ob.ilo && sb.jlo - 1 <= ob.jhi && sb.jhi + 1 >= ob.jlo)) continue;
      const Box u = join(sb, ob);
      if (ua[o] > 1.5 * (areab + (double)area(ob))) continue;
      if (!A.all(u)) continue;
      bool clash = false;
      for (size_t t = 0; t < nb && !clash; ++t)
        if (t != b && t != o && overlaps(u, boxes[t])) clash = true;
      if (clash) continue;
      boxes[o] = u;
      boxes.erase(boxes.begin() + (long)b);
      merged = true;
      break;
    }
    if (!merged) {
      ++dropped;
      boxes.erase(boxes.begin() + (long)b);
    }
  }
  return dropped;
}

}  // namespace

// Clusters the tags; writes min(count, cap) boxes and returns the count.
int bsam_cluster_impl(const unsigned char* tags, const unsigned char* allowed, int n1, int n2, int min_width,
                      double fill_ratio, int* boxes, int cap, int* ndropped) {
  const Map T{tags, n1, n2}, A{allowed, n1, n2};
  std::vector<Box> out = br_cluster(T, min_width, fill_ratio);
  out = clip_allowed(out, T, A);
  grow_min(out, A, min_width);
  const int dropped = merge_or_drop_thin(out, A);
  if (ndropped) *ndropped = dropped;
  for (size_t b = 0; b < out.size() && (int)b < cap; ++b) {
    boxes[4 * b + 0] = out[b].ilo;
    boxes[4 * b + 1] = out[b].ihi;
    boxes[4 * b + 2] = out[b].jlo;
    boxes[4 * b + 3] = out[b].jhi;
  }
  return (int)out.size();
}
