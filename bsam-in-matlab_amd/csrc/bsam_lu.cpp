// Host-side dense inverse of the coarsest-level matrix (ebCoarsest.m:72-87
// factorises it; here an explicit inverse so that the device solve is one
// matrix-vector product).  Compiled by the host compiler alone, with
// -ffp-contract=off so each product and difference rounds as written.
#include <cmath>
#include <cstddef>
#include <utility>
#include <vector>

// Explicit inverse of the row-major M x M matrix A (overwritten by its LU
// factors) by LU with partial pivoting, first maximum wins, then forward and
// backward substitution on the permuted identity; Ainv (M x M) is written
// column-major.
// Every entry sees the same operations in the same order on every build; the
// AVX2 clone only runs the independent row updates four at a time.
__attribute__((target_clones("avx2", "default"))) bool bsam_lu_inverse(double* __restrict__ A, int M,
                                                                        double* __restrict__ Ainv) {
  auto at = [&](int r, int c) -> double& { return A[(size_t)r * M + c]; };
  std::vector<int> perm(M);
  for (int i = 0; i < M; ++i) perm[i] = i;
  for (int k = 0; k < M; ++k) {
    int piv = k;
    double best = std::fabs(at(k, k));
    for (int i = k + 1; i < M; ++i)
      if (std::fabs(at(i, k)) > best) { best = std::fabs(at(i, k)); piv = i; }
    if (at(piv, k) == 0.0) return false;
    if (piv != k) {
      for (int c = 0; c < M; ++c) std::swap(at(k, c), at(piv, c));
      std::swap(perm[k], perm[piv]);
    }
    const double akk = at(k, k);
    for (int i = k + 1; i < M; ++i) at(i, k) = at(i, k) / akk;
    const double* __restrict__ rk = &at(k, 0);
    for (int i = k + 1; i < M; ++i) {
      const double l = at(i, k);
      double* __restrict__ ri = &at(i, 0);
      for (int c = k + 1; c < M; ++c) ri[c] = ri[c] - l * rk[c];
    }
  }
  // X = L^-1 P with its columns reordered (column c of X is column perm[c] of
  // L^-1 P), so that it starts as the identity and row k stays exactly +0
  // beyond column k during forward substitution: those terms (x - l*0 = x,
  // +0 - l*0 = +0) are skipped without changing any entry.
  std::vector<double> Xv((size_t)M * M, 0.0);  // row-major
  double* __restrict__ X = Xv.data();
  for (int i = 0; i < M; ++i) X[(size_t)i * M + i] = 1.0;
  for (int k = 0; k < M; ++k) {
    const double* __restrict__ xk = X + (size_t)k * M;
    for (int i = k + 1; i < M; ++i) {
      const double l = at(i, k);
      double* __restrict__ xi = X + (size_t)i * M;
      for (int c = 0; c <= k; ++c) xi[c] = xi[c] - l * xk[c];
    }
  }
  for (int k = M - 1; k >= 0; --k) {
    const double ukk = at(k, k);
    double* __restrict__ xk = X + (size_t)k * M;
    for (int c = 0; c < M; ++c) xk[c] = xk[c] / ukk;
    for (int i = 0; i < k; ++i) {
      const double u = at(i, k);
      double* __restrict__ xi = X + (size_t)i * M;
      for (int c = 0; c < M; ++c) xi[c] = xi[c] - u * xk[c];
    }
  }
  for (int i = 0; i < M; ++i)  // column-major for the device, columns back in place
    for (int c = 0; c < M; ++c) Ainv[i + (size_t)perm[c] * M] = X[(size_t)i * M + c];
  return true;
}
