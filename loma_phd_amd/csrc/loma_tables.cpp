// loma_tables.cpp -- host construction of the wavefront-scan tables of the
// fixed banded factors (first/second derivative and integration LU).
//
// The y-column solver splits a banded triangular solve over the 64 lanes of
// a wavefront, each lane owning M consecutive rows.  A lane's local sweep
// leaves a carry (the last P unknowns for the lower solve, the first P for
// the upper one) that depends affinely on the carry of its neighbour: c_l =
// c_l^0 + A_l c_{l-1} (lower) or c_l^0 + A_l c_{l+1} (upper).  For a fixed
// factor the maps A_l are data independent; this file tabulates the
// Hillis-Steele compositions W_k[l] = A_l A_{l-1} ... A_{l-2^k+1} (lower;
// mirrored for upper) for k = 0..5, lane-major, and the number of levels
// after which every longer product is below 1e-22 in max norm.
#include <cmath>
#include <vector>

#include "loma_internal.h"

namespace loma {

namespace {

using Mat = std::vector<double>;  // P x P row-major

Mat matmul(const Mat& a, const Mat& b, int P) {
  Mat c(P * P, 0.0);
  for (int i = 0; i < P; ++i)
    for (int k = 0; k < P; ++k)
      for (int j = 0; j < P; ++j) c[i * P + j] += a[i * P + k] * b[k * P + j];
  return c;
}

double maxabs(const Mat& a) {
  double m = 0.0;
  for (double v : a) m = std::fmax(m, std::fabs(v));
  return m;
}

}  // namespace

int scan_table(int my, int M, int P, bool upper, const double* lo, const double* up,
               const double* inv, double* H) {
  const int L = 64;
  // row accessors with zero padding; lower multipliers by destination row
  auto lod = [&](int r, int i) -> double {  // multiplier of row r-i into row r
    if (r >= my || r - i < 0) return 0.0;
    return lo[(size_t)(r - i) * P + i - 1];
  };
  auto upd = [&](int r, int d) -> double { return r < my ? up[(size_t)r * P + d - 1] : 0.0; };
  auto ivd = [&](int r) -> double { return r < my ? inv[r] : 0.0; };
  std::vector<Mat> A(L, Mat(P * P, 0.0));
  for (int l = 0; l < L; ++l) {
    const int r0 = l * M;
    for (int b = 0; b < P; ++b) {
      std::vector<double> g(M, 0.0);
      if (!upper) {
        for (int t = 0; t < M; ++t) {
          double acc = 0.0;
          for (int i = 1; i <= P; ++i) {
            const double pv = (t - i >= 0) ? g[t - i] : ((i - t - 1) == b ? 1.0 : 0.0);
            acc -= lod(r0 + t, i) * pv;
          }
          g[t] = acc;
        }
        for (int q = 0; q < P; ++q) A[l][q * P + b] = g[M - 1 - q];
      } else {
        for (int t = M - 1; t >= 0; --t) {
          double acc = 0.0;
          for (int d = 1; d <= P; ++d) {
            const double pv = (t + d < M) ? g[t + d] : ((t + d - M) == b ? 1.0 : 0.0);
            acc -= upd(r0 + t, d) * pv;
          }
          g[t] = acc * ivd(r0 + t);
        }
        for (int q = 0; q < P; ++q) A[l][q * P + b] = g[q];
      }
    }
  }
  // compositions per level
  std::vector<Mat> W = A;
  for (int k = 0; k < 6; ++k) {
    for (int l = 0; l < L; ++l)
      for (int q = 0; q < P; ++q)
        for (int b = 0; b < P; ++b)
          H[((size_t)(k * P + q) * P + b) * 64 + l] = W[l][q * P + b];
    const int d = 1 << k;
    std::vector<Mat> Wn = W;
    for (int l = 0; l < L; ++l) {
      const int m = upper ? l + d : l - d;
      if (m >= 0 && m < L) Wn[l] = matmul(W[l], W[m], P);
    }
    W.swap(Wn);
  }
  // levels: the smallest K such that every product of the maps of at least
  // 2^K consecutive lanes is below the threshold
  const double eps = 1e-22;
  for (int K = 0; K <= 6; ++K) {
    const int span = 1 << K;
    if (span >= L) return 6;
    bool ok = true;
    for (int l = 0; l < L && ok; ++l) {
      // products A_l ... A_m over lanes l, l-1, ..., m (lower) with l-m+1 >= span
      Mat prod = A[l];
      int len = 1;
      for (int m = upper ? l + 1 : l - 1; upper ? m < L : m >= 0; m += upper ? 1 : -1) {
        if (len >= span && maxabs(prod) >= eps) { ok = false; break; }
        prod = matmul(prod, A[m], P);
        ++len;
      }
      if (ok && len >= span && maxabs(prod) >= eps) ok = false;
    }
    if (ok) return K;
  }
  return 6;
}

}  // namespace loma
