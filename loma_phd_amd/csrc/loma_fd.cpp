// loma_fd.cpp -- host-side construction of the compact finite-difference
// operators in y (the product's own implementation; runs once per run).
//
// Reference behaviour reproduced:
//   first derivative: uniform-eta compact scheme with a metric factor,
//     trunk/lomacte3/src/cfdiff.f:24-182 (pred1m6o) and :912-937 (fmap)
//   second derivative: non-uniform compact scheme whose coefficients solve
//     local Taylor-matching systems, cfdiff.f:196-501 (prederiv2)
//   banded LU without pivoting, cfdiff.f:954-1137 (bandec/bandec7)
//
// Output layout (row-major, one row per grid point j):
//   b1[j][7], a1[j][7]  : RHS band B1 and LHS band A1 of the first derivative
//   b2[j][5], a2[j][5]  : RHS band B2 and LHS band A2 of the second derivative
//   fmap[j]             : 1/(dy/deta)
#include "loma_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace loma {

// Dense solve with partial pivoting; a is n x n row-major, b length n.
static bool dense_solve(std::vector<double>& a, std::vector<double>& b, int n) {
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r * n + c]) > std::fabs(a[piv * n + c])) piv = r;
    if (a[piv * n + c] == 0.0) return false;
    if (piv != c) {
      for (int k = 0; k < n; ++k) std::swap(a[c * n + k], a[piv * n + k]);
      std::swap(b[c], b[piv]);
    }
    for (int r = c + 1; r < n; ++r) {
      double f = a[r * n + c] / a[c * n + c];
      if (f == 0.0) continue;
      for (int k = c; k < n; ++k) a[r * n + k] -= f * a[c * n + k];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
    for (int k = r + 1; k < n; ++k) s -= a[r * n + k] * b[k];
    b[r] = s / a[r * n + r];
  }
  return true;
}

// Second-derivative compact stencil at point j on a non-uniform grid:
//   sum_{m in lhs} L_m d2u(j+m) = sum_{m in rhs} R_m u(j+m), L_0 = 1,
// with coefficients fixed by matching Taylor terms of orders 0..nl+nr-1.
static bool d2_stencil(const double* x, int j, const int* lo, int nl,
                       const int* ro, int nr, double* L, double* R) {
  const int n = nl + nr;  // unknowns: nl LHS (off-diagonal) + nr RHS
  std::vector<double> a(n * n, 0.0), b(n, 0.0);
  for (int q = 0; q < n; ++q) {
    // Taylor order q: sum R_m h_m^q / q! = sum_{m} L_m h_m^{q-2}/(q-2)! (+ the
    // diagonal L_0 = 1 term).  Multiply through by q! to keep integers.
    for (int c = 0; c < nl; ++c) {
      double h = x[j + lo[c]] - x[j];
      a[q * n + c] = (q >= 2) ? -double(q) * (q - 1) * std::pow(h, q - 2) : 0.0;
    }
    for (int c = 0; c < nr; ++c) {
      double h = x[j + ro[c]] - x[j];
      a[q * n + nl + c] = (ro[c] == 0) ? (q == 0 ? 1.0 : 0.0) : std::pow(h, q);
    }
    b[q] = (q == 2) ? 2.0 : 0.0;
  }
  if (!dense_solve(a, b, n)) return false;
  for (int c = 0; c < nl; ++c) L[c] = b[c];
  for (int c = 0; c < nr; ++c) R[c] = b[nl + c];
  return true;
}

// Banded LU without pivoting of a (n x nb) row-major band (half bandwidth p,
// entry d of row j is column j-p+d).  Outputs per row: inv pivot, upper p
// entries U(j,j+1..j+p), and p multipliers L(j,i) used for row j+i.
void band_lu(const double* band, int n, int p, double* inv, double* up,
             double* lo) {
  const int nb = 2 * p + 1;
  std::vector<double> w(band, band + (size_t)n * nb);
  for (int k = 0; k < n; ++k) {
    double piv = w[(size_t)k * nb + p];
    double iv = 1.0 / piv;
    inv[k] = iv;
    for (int d = 1; d <= p; ++d)
      up[(size_t)k * p + d - 1] = (k + d < n) ? w[(size_t)k * nb + p + d] : 0.0;
    for (int i = 1; i <= p; ++i) {
      if (k + i >= n) { lo[(size_t)k * p + i - 1] = 0.0; continue; }
      double* row = &w[(size_t)(k + i) * nb];
      // element (k+i, k) sits at offset p - i
      double l = row[p - i] * iv;
      lo[(size_t)k * p + i - 1] = l;
      row[p - i] = 0.0;
      for (int d = 1; d <= p; ++d) {
        int off = p - i + d;  // column k+d in row k+i
        if (off < nb) row[off] -= l * w[(size_t)k * nb + p + d];
      }
    }
  }
}

int fd_build(int my, const double* y, FdOps& op) {
  if (my < 8) return LOMA_ERR_ARG;
  op.my = my;
  op.a1.assign((size_t)my * 7, 0.0);
  op.b1.assign((size_t)my * 7, 0.0);
  op.a2.assign((size_t)my * 5, 0.0);
  op.b2.assign((size_t)my * 5, 0.0);
  op.fmap.assign(my, 0.0);
  // ---- first derivative in eta (uniform h = 2/(my-1)) ----
  {
    const double h = 2.0 / double(my - 1);
    // interior sixth-order-class coefficients (cfdiff.f:49-57)
    const double ca = 0.56926818603810, cb = 0.31997939759919,
                 cc = 0.03751544855390;
    const double al = -3.0 / 4.0 + (33.0 / 20.0) * cc + (239.0 / 180.0) * ca +
                      (82.0 / 45.0) * cb;
    const double be = 3.0 / 10.0 + (24.0 / 25.0) * cc - (88.0 / 225.0) * ca +
                      (34.0 / 225.0) * cb;
    const double de = -1.0 / 20.0 + (39.0 / 100.0) * cc + (19.0 / 300.0) * ca +
                      (2.0 / 75.0) * cb;
    const double Li[7] = {de, be, al, 1.0, al, be, de};
    const double Ri[7] = {-cc / h, -cb / h, -ca / h, 0.0, ca / h, cb / h, cc / h};
    // closures at rows 0,1,2 (cfdiff.f:83-130); mirrored at the top end
    const double L0[7] = {0, 0, 0, 1, 1, 0, 0};
    const double R0[7] = {0, 0, 0, -2.0 / h, 2.0 / h, 0, 0};
    const double L1[7] = {0, 0, 0.25, 1, 0.25, 0, 0};
    const double R1[7] = {0, 0, -0.75 / h, 0, 0.75 / h, 0, 0};
    const double L2[7] = {0, 1.0 / 36.0, 4.0 / 9.0, 1, 4.0 / 9.0, 1.0 / 36.0, 0};
    const double R2[7] = {0, -25.0 / 216.0 / h, -20.0 / 27.0 / h, 0,
                          20.0 / 27.0 / h, 25.0 / 216.0 / h, 0};
    const double* Lb[3] = {L0, L1, L2};
    const double* Rb[3] = {R0, R1, R2};
    for (int j = 0; j < my; ++j) {
      const double *L = Li, *R = Ri;
      double sgn = 1.0;
      bool mirror = false;
      if (j < 3) { L = Lb[j]; R = Rb[j]; }
      else if (j >= my - 3) { L = Lb[my - 1 - j]; R = Rb[my - 1 - j]; mirror = true; sgn = -1.0; }
      for (int d = 0; d < 7; ++d) {
        int s = mirror ? 6 - d : d;
        op.a1[(size_t)j * 7 + d] = L[s];
        op.b1[(size_t)j * 7 + d] = sgn * R[s];
      }
    }
  }
  // ---- second derivative on the stretched grid ----
  {
    for (int j = 0; j < my; ++j) {
      double L[8], R[8];
      if (j >= 2 && j <= my - 3) {
        const int lo[4] = {-2, -1, 1, 2}, ro[5] = {-2, -1, 0, 1, 2};
        if (!d2_stencil(y, j, lo, 4, ro, 5, L, R)) return LOMA_ERR_ARG;
        const double Lf[5] = {L[0], L[1], 1.0, L[2], L[3]};
        for (int d = 0; d < 5; ++d) {
          op.a2[(size_t)j * 5 + d] = Lf[d];
          op.b2[(size_t)j * 5 + d] = R[d];
        }
      } else if (j == 0) {
        const int lo[2] = {1, 2}, ro[3] = {0, 1, 2};
        if (!d2_stencil(y, j, lo, 2, ro, 3, L, R)) return LOMA_ERR_ARG;
        const double Lf[5] = {0, 0, 1.0, L[0], L[1]}, Rf[5] = {0, 0, R[0], R[1], R[2]};
        for (int d = 0; d < 5; ++d) { op.a2[d] = Lf[d]; op.b2[d] = Rf[d]; }
      } else if (j == 1) {
        const int lo[3] = {-1, 1, 2}, ro[4] = {-1, 0, 1, 2};
        if (!d2_stencil(y, j, lo, 3, ro, 4, L, R)) return LOMA_ERR_ARG;
        const double Lf[5] = {0, L[0], 1.0, L[1], L[2]}, Rf[5] = {0, R[0], R[1], R[2], R[3]};
        for (int d = 0; d < 5; ++d) { op.a2[5 + d] = Lf[d]; op.b2[5 + d] = Rf[d]; }
      } else if (j == my - 2) {
        const int lo[3] = {-2, -1, 1}, ro[4] = {-2, -1, 0, 1};
        if (!d2_stencil(y, j, lo, 3, ro, 4, L, R)) return LOMA_ERR_ARG;
        const double Lf[5] = {L[0], L[1], 1.0, L[2], 0}, Rf[5] = {R[0], R[1], R[2], R[3], 0};
        for (int d = 0; d < 5; ++d) { op.a2[(size_t)j * 5 + d] = Lf[d]; op.b2[(size_t)j * 5 + d] = Rf[d]; }
      } else {  // j == my-1
        const int lo[2] = {-2, -1}, ro[3] = {-2, -1, 0};
        if (!d2_stencil(y, j, lo, 2, ro, 3, L, R)) return LOMA_ERR_ARG;
        const double Lf[5] = {L[0], L[1], 1.0, 0, 0}, Rf[5] = {R[0], R[1], R[2], 0, 0};
        for (int d = 0; d < 5; ++d) { op.a2[(size_t)j * 5 + d] = Lf[d]; op.b2[(size_t)j * 5 + d] = Rf[d]; }
      }
    }
  }
  // ---- LU factors ----
  op.inv1.resize(my); op.up1.resize((size_t)my * 3); op.lo1.resize((size_t)my * 3);
  op.inv2.resize(my); op.up2.resize((size_t)my * 2); op.lo2.resize((size_t)my * 2);
  band_lu(op.a1.data(), my, 3, op.inv1.data(), op.up1.data(), op.lo1.data());
  band_lu(op.a2.data(), my, 2, op.inv2.data(), op.up2.data(), op.lo2.data());
  // ---- metric: fmap = 1 / (A1^{-1} B1 y) ----
  {
    std::vector<double> w(my);
    band_apply(op.b1.data(), my, 3, y, w.data());
    band_lu_solve(op.inv1.data(), op.up1.data(), op.lo1.data(), my, 3, w.data());
    for (int j = 0; j < my; ++j) op.fmap[j] = 1.0 / w[j];
  }
  // ---- integration operator (inty8, laps.f:484): B1 x = A1 f.  B1 has the
  // constant vector in its null space (its rows sum to zero); the reference
  // factors it anyway and removes the resulting huge constant by subtracting
  // x(1).  The same function is obtained, without that cancellation, by
  // replacing the last row of B1 with the gauge x(my) = 0.
  {
    std::vector<double> bi(op.b1);
    for (int d = 0; d < 7; ++d) bi[(size_t)(my - 1) * 7 + d] = 0.0;
    bi[(size_t)(my - 1) * 7 + 3] = 1.0;
    op.invi.resize(my); op.upi.resize((size_t)my * 3); op.loi.resize((size_t)my * 3);
    band_lu(bi.data(), my, 3, op.invi.data(), op.upi.data(), op.loi.data());
  }
  return LOMA_OK;
}

void band_apply(const double* band, int n, int p, const double* u, double* out) {
  const int nb = 2 * p + 1;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int d = 0; d < nb; ++d) {
      int c = j - p + d;
      if (c >= 0 && c < n) s += band[(size_t)j * nb + d] * u[c];
    }
    out[j] = s;
  }
}

void band_lu_solve(const double* inv, const double* up, const double* lo, int n,
                   int p, double* b) {
  for (int k = 0; k < n; ++k)
    for (int i = p; i >= 1; --i)
      if (k - i >= 0) b[k] -= lo[(size_t)(k - i) * p + i - 1] * b[k - i];
  for (int k = n - 1; k >= 0; --k) {
    double s = b[k];
    for (int d = 1; d <= p; ++d)
      if (k + d < n) s -= up[(size_t)k * p + d - 1] * b[k + d];
    b[k] = s * inv[k];
  }
}

}  // namespace loma
