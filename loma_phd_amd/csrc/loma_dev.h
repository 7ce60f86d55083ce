// loma_dev.h -- device building blocks for the y-direction operators.
//
// One 64-lane wavefront owns one complex y-column (one (kx,kz) mode).  Lane l
// holds the contiguous rows [l*M, l*M+M) in registers (rows >= my are zero
// padding whose coefficients are zero).  Banded triangular solves are split
// into a local sweep per lane, a log-step scan of the affine carry maps
// across the wavefront (Hillis-Steele over 64 lanes with __shfl_up/down), and
// a second local sweep from the exact incoming carry.  This gives the result
// of the reference's sequential banbks/banbks7 sweeps (cfdiff.f:954-1051)
// with all 64 lanes busy.
#pragma once

#include <hip/hip_runtime.h>

namespace loma {

constexpr int PADR = 4;  // leading zero rows in every per-row table

// element (row k, entry idx) of a row-major per-row table with `rs` doubles
// per row starting at column `off`; rows are shifted by PADR.
struct BandRef {
  const double* __restrict__ p;
  int rs;
  int off;
};
__device__ __forceinline__ double bget(const BandRef& b, int k, int idx) {
  return b.p[(k + PADR) * b.rs + b.off + idx];
}

// ---- wavefront scans of affine maps s_out = A s_in + c -----------------
template <int P, int NC>
__device__ __forceinline__ void scan_affine_up(double (&A)[P][P], double (&c)[NC][P],
                                               int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    double Ap[P][P], cp[NC][P];
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
      for (int b = 0; b < P; ++b) Ap[q][b] = __shfl_up(A[q][b], d);
#pragma unroll
    for (int n = 0; n < NC; ++n)
#pragma unroll
      for (int q = 0; q < P; ++q) cp[n][q] = __shfl_up(c[n][q], d);
    if (lane >= d) {
      double cn[NC][P], An[P][P];
#pragma unroll
      for (int n = 0; n < NC; ++n)
#pragma unroll
        for (int q = 0; q < P; ++q) {
          double s = c[n][q];
#pragma unroll
          for (int r = 0; r < P; ++r) s += A[q][r] * cp[n][r];
          cn[n][q] = s;
        }
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int b = 0; b < P; ++b) {
          double s = 0.0;
#pragma unroll
          for (int r = 0; r < P; ++r) s += A[q][r] * Ap[r][b];
          An[q][b] = s;
        }
#pragma unroll
      for (int n = 0; n < NC; ++n)
#pragma unroll
        for (int q = 0; q < P; ++q) c[n][q] = cn[n][q];
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int b = 0; b < P; ++b) A[q][b] = An[q][b];
    }
  }
}

template <int P, int NC>
__device__ __forceinline__ void scan_affine_down(double (&A)[P][P], double (&c)[NC][P],
                                                 int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    double Ap[P][P], cp[NC][P];
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
      for (int b = 0; b < P; ++b) Ap[q][b] = __shfl_down(A[q][b], d);
#pragma unroll
    for (int n = 0; n < NC; ++n)
#pragma unroll
      for (int q = 0; q < P; ++q) cp[n][q] = __shfl_down(c[n][q], d);
    if (lane + d < 64) {
      double cn[NC][P], An[P][P];
#pragma unroll
      for (int n = 0; n < NC; ++n)
#pragma unroll
        for (int q = 0; q < P; ++q) {
          double s = c[n][q];
#pragma unroll
          for (int r = 0; r < P; ++r) s += A[q][r] * cp[n][r];
          cn[n][q] = s;
        }
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int b = 0; b < P; ++b) {
          double s = 0.0;
#pragma unroll
          for (int r = 0; r < P; ++r) s += A[q][r] * Ap[r][b];
          An[q][b] = s;
        }
#pragma unroll
      for (int n = 0; n < NC; ++n)
#pragma unroll
        for (int q = 0; q < P; ++q) c[n][q] = cn[n][q];
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int b = 0; b < P; ++b) A[q][b] = An[q][b];
    }
  }
}

// ---- unit lower solve: w_r = x_r - sum_{i=P..1} L(r-i,i) w_{r-i} ---------
// lo(k, i-1) is the multiplier row k contributes to row k+i.
template <int M, int P, int NC>
__device__ __forceinline__ void lsolve(const BandRef lo, int r0, int lane,
                                       double (&x)[NC][M]) {
  static_assert(P <= M, "chunk shorter than bandwidth");
  // carry of this chunk with zero incoming state
  double c[NC][P];
  {
    double w[NC][M];
#pragma unroll
    for (int t = 0; t < M; ++t) {
#pragma unroll
      for (int n = 0; n < NC; ++n) w[n][t] = x[n][t];
#pragma unroll
      for (int i = P; i >= 1; --i) {
        if (t - i < 0) continue;
        double l = bget(lo, r0 + t - i, i - 1);
#pragma unroll
        for (int n = 0; n < NC; ++n) w[n][t] -= l * w[n][t - i];
      }
    }
#pragma unroll
    for (int n = 0; n < NC; ++n)
#pragma unroll
      for (int q = 1; q <= P; ++q) c[n][q - 1] = w[n][M - q];
  }
  // response of the carry to a unit incoming state s_b = w_{r0-b}
  double A[P][P];
#pragma unroll
  for (int b = 1; b <= P; ++b) {
    double g[M];
#pragma unroll
    for (int t = 0; t < M; ++t) {
      double acc = 0.0;
#pragma unroll
      for (int i = P; i >= 1; --i) {
        double pv = (t - i >= 0) ? g[t - i] : ((i - t) == b ? 1.0 : 0.0);
        acc -= bget(lo, r0 + t - i, i - 1) * pv;
      }
      g[t] = acc;
    }
#pragma unroll
    for (int q = 1; q <= P; ++q) A[q - 1][b - 1] = g[M - q];
  }
  scan_affine_up<P, NC>(A, c, lane);
  double s[NC][P];
#pragma unroll
  for (int n = 0; n < NC; ++n)
#pragma unroll
    for (int q = 0; q < P; ++q) {
      double v = __shfl_up(c[n][q], 1);
      s[n][q] = lane == 0 ? 0.0 : v;
    }
#pragma unroll
  for (int t = 0; t < M; ++t) {
#pragma unroll
    for (int i = P; i >= 1; --i) {
      double l = bget(lo, r0 + t - i, i - 1);
#pragma unroll
      for (int n = 0; n < NC; ++n) {
        double pv = (t - i >= 0) ? x[n][t - i] : s[n][i - t - 1];
        x[n][t] -= l * pv;
      }
    }
  }
}

// ---- upper solve: x_r = (w_r - sum_{d=1..P} U(r,d) x_{r+d}) * inv_r ------
template <int M, int P, int NC>
__device__ __forceinline__ void usolve(const BandRef up, const BandRef inv, int r0,
                                       int lane, double (&x)[NC][M]) {
  double c[NC][P];
  {
    double w[NC][M];
#pragma unroll
    for (int t = M - 1; t >= 0; --t) {
#pragma unroll
      for (int n = 0; n < NC; ++n) w[n][t] = x[n][t];
#pragma unroll
      for (int d = 1; d <= P; ++d) {
        if (t + d >= M) continue;
        double u = bget(up, r0 + t, d - 1);
#pragma unroll
        for (int n = 0; n < NC; ++n) w[n][t] -= u * w[n][t + d];
      }
      double iv = bget(inv, r0 + t, 0);
#pragma unroll
      for (int n = 0; n < NC; ++n) w[n][t] *= iv;
    }
#pragma unroll
    for (int n = 0; n < NC; ++n)
#pragma unroll
      for (int q = 0; q < P; ++q) c[n][q] = w[n][q];
  }
  double A[P][P];
#pragma unroll
  for (int b = 1; b <= P; ++b) {
    double g[M];
#pragma unroll
    for (int t = M - 1; t >= 0; --t) {
      double acc = 0.0;
#pragma unroll
      for (int d = 1; d <= P; ++d) {
        double pv = (t + d < M) ? g[t + d] : ((t + d - M + 1) == b ? 1.0 : 0.0);
        acc -= bget(up, r0 + t, d - 1) * pv;
      }
      g[t] = acc * bget(inv, r0 + t, 0);
    }
#pragma unroll
    for (int q = 0; q < P; ++q) A[q][b - 1] = g[q];
  }
  scan_affine_down<P, NC>(A, c, lane);
  double s[NC][P];
#pragma unroll
  for (int n = 0; n < NC; ++n)
#pragma unroll
    for (int q = 0; q < P; ++q) {
      double v = __shfl_down(c[n][q], 1);
      s[n][q] = lane == 63 ? 0.0 : v;
    }
#pragma unroll
  for (int t = M - 1; t >= 0; --t) {
#pragma unroll
    for (int d = 1; d <= P; ++d) {
      double u = bget(up, r0 + t, d - 1);
#pragma unroll
      for (int n = 0; n < NC; ++n) {
        double pv = (t + d < M) ? x[n][t + d] : s[n][t + d - M];
        x[n][t] -= u * pv;
      }
    }
    double iv = bget(inv, r0 + t, 0);
#pragma unroll
    for (int n = 0; n < NC; ++n) x[n][t] *= iv;
  }
}

// ---- banded matvec y_r = sum_d B(r,d) u_{r-P+d} ---------------------------
template <int M, int P, int NC>
__device__ __forceinline__ void bmatvec(const BandRef B, int r0, int lane,
                                        const double (&u)[NC][M], double (&y)[NC][M]) {
  double hl[NC][P], hr[NC][P];
#pragma unroll
  for (int n = 0; n < NC; ++n)
#pragma unroll
    for (int q = 0; q < P; ++q) {
      double a = __shfl_up(u[n][M - P + q], 1);
      double b = __shfl_down(u[n][q], 1);
      hl[n][q] = lane == 0 ? 0.0 : a;
      hr[n][q] = lane == 63 ? 0.0 : b;
    }
#pragma unroll
  for (int t = 0; t < M; ++t) {
    double acc[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) acc[n] = 0.0;
#pragma unroll
    for (int d = 0; d <= 2 * P; ++d) {
      const int col = t - P + d;
      double bv = bget(B, r0 + t, d);
#pragma unroll
      for (int n = 0; n < NC; ++n) {
        double v = col < 0 ? hl[n][col + P] : (col >= M ? hr[n][col - M] : u[n][col]);
        acc[n] += bv * v;
      }
    }
#pragma unroll
    for (int n = 0; n < NC; ++n) y[n][t] = acc[n];
  }
}

// broadcast the value of row `row` of a chunked column to all lanes
template <int M>
__device__ __forceinline__ double col_value(const double (&x)[M], int row) {
  const int l = row / M, t = row - l * M;
  double mine = 0.0;
#pragma unroll
  for (int q = 0; q < M; ++q)
    if (q == t) mine = x[q];
  return __shfl(mine, l);
}

// Device tables of one y grid, rows shifted by PADR, zero padded.
struct YTab {
  const double* b1;    // [rows][7]  RHS of d/deta (dt12)
  const double* a1;    // [rows][7]  LHS of d/deta (dt11, unfactored)
  const double* inv1;  // [rows]     LU of a1
  const double* up1;   // [rows][3]
  const double* lo1;   // [rows][3]
  const double* fmap;  // [rows]
  const double* b2;    // [rows][5]  RHS of d2/dy2 (dt22)
  const double* a2;    // [rows][5]  LHS of d2/dy2 (dt21, unfactored)
  const double* inv2;  // LU of a2
  const double* up2;   // [rows][2]
  const double* lo2;   // [rows][2]
  const double* invi;  // LU of b1 (integration)
  const double* upi;   // [rows][3]
  const double* loi;   // [rows][3]
  int my;
};

// first derivative du = fmap * A1^{-1} B1 u   (deryr2, cfdiff.f:518-601)
template <int M, int NC>
__device__ __forceinline__ void op_dy1(const YTab& T, int r0, int lane,
                                       const double (&u)[NC][M], double (&du)[NC][M],
                                       bool apply_fmap = true) {
  bmatvec<M, 3, NC>(BandRef{T.b1, 7, 0}, r0, lane, u, du);
  lsolve<M, 3, NC>(BandRef{T.lo1, 3, 0}, r0, lane, du);
  usolve<M, 3, NC>(BandRef{T.up1, 3, 0}, BandRef{T.inv1, 1, 0}, r0, lane, du);
  if (apply_fmap) {
#pragma unroll
    for (int t = 0; t < M; ++t) {
      double f = bget(BandRef{T.fmap, 1, 0}, r0 + t, 0);
#pragma unroll
      for (int n = 0; n < NC; ++n) du[n][t] *= f;
    }
  }
}

// second derivative d2u = A2^{-1} B2 u   (deryyr2, cfdiff.f:735-797)
template <int M, int NC>
__device__ __forceinline__ void op_dy2(const YTab& T, int r0, int lane,
                                       const double (&u)[NC][M], double (&d2u)[NC][M]) {
  bmatvec<M, 2, NC>(BandRef{T.b2, 5, 0}, r0, lane, u, d2u);
  lsolve<M, 2, NC>(BandRef{T.lo2, 2, 0}, r0, lane, d2u);
  usolve<M, 2, NC>(BandRef{T.up2, 2, 0}, BandRef{T.inv2, 1, 0}, r0, lane, d2u);
}

// Helmholtz/Poisson solve (dt22 - rK dt21) v = rhs with identity end rows,
// LU per mode class in `lu` ([rows][5] = inv, up1, up2, lo1, lo2), the RHS
// being dt21*f with zero end rows (calcrhslap, laps.f:17-46).
template <int M, int NC>
__device__ __forceinline__ void op_helm(const YTab& T, const double* __restrict__ lu,
                                        int r0, int lane, const double (&f)[NC][M],
                                        double (&v)[NC][M]) {
  bmatvec<M, 2, NC>(BandRef{T.a2, 5, 0}, r0, lane, f, v);
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const int r = r0 + t;
    if (r == 0 || r == T.my - 1) {
#pragma unroll
      for (int n = 0; n < NC; ++n) v[n][t] = 0.0;
    }
  }
  lsolve<M, 2, NC>(BandRef{lu, 5, 3}, r0, lane, v);
  usolve<M, 2, NC>(BandRef{lu, 5, 1}, BandRef{lu, 5, 0}, r0, lane, v);
}

}  // namespace loma
