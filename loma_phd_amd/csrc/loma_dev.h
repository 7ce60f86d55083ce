// loma_dev.h -- device building blocks for the y-direction operators.
//
// One 64-lane wavefront owns one complex y-column (one (kx,kz) mode).  Lane l
// holds the contiguous rows [l*M, l*M+M) in registers (rows >= my are zero
// padding whose coefficients are zero).  Banded triangular solves are split
// into an in-place local sweep per lane, a log-step scan of the affine carry
// maps across the wavefront (Hillis-Steele over 64 lanes), and a correction
// sweep driven by the exact incoming carry.  This gives the result of the
// reference's sequential banbks/banbks7 sweeps (cfdiff.f:954-1051) with all
// 64 lanes busy.
//
// Per-row coefficient tables are stored lane-major: entry `idx` of row
// (lane*M + t) sits at base[(t*ncol + idx)*64 + lane], so every table load of
// a wavefront is one contiguous 512-byte segment.  Lower-triangle
// multipliers are stored by destination row (lo(r, i-1) multiplies row r-i
// into row r), so every lane reads only its own rows.
#pragma once

#include <hip/hip_runtime.h>

namespace loma {

constexpr int WAVE = 64;

// Scheduling fence between the stages of an operator (and between
// operators): without it the scheduler hoists the table loads of later
// stages over the scans, holding every coefficient of the operator in
// registers at once.
__device__ __forceinline__ void op_fence() { __builtin_amdgcn_sched_barrier(0); }

// max that keeps a NaN (fmax drops it): a non-finite state reaches the CFL
// maxima and from there the step size, where it is flagged
__device__ __forceinline__ double max_nan(double a, double b) { return (b > a || b != b) ? b : a; }

// ---- buffer addressing ----------------------------------------------------
// Column and table accesses go through buffer resources held in scalar
// registers: the per-row and per-entry offsets are scalar (or instruction
// immediates), the only vector offset is the lane's, and the hardware range
// check replaces per-row bounds tests (out-of-range loads return 0,
// out-of-range stores are dropped).  The range check covers the vector
// offset and the immediate, not the scalar offset, so anything that may run
// past the end is folded into the resource base instead.
constexpr int RSRC_WORD3 = 0x00020000;  // raw 32-bit data format (CDNA)

__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0,
                                           bytes < 0 ? 0 : bytes, RSRC_WORD3);
}

// lane-major table view: resource over the table, entries per row, first
// entry used, this lane's byte offset and a scalar bias (see relaunder)
struct TabRef {
  __amdgpu_buffer_rsrc_t r;
  int ncol;
  int off;
  int lane_off;
  int sbias;
};
__device__ __forceinline__ TabRef tab(const double* base, int ncol, int off, int lane) {
  return TabRef{make_rsrc(base, 0x7fffffff), ncol, off, lane * 8, 0};
}
__device__ __forceinline__ double tget(const TabRef& b, int t, int idx) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(
      b.r, b.lane_off, b.sbias + (t * b.ncol + b.off + idx) * (WAVE * 8), 0);
  return __builtin_bit_cast(double, v);
}
// entry of a table in the arithmetic type V: the FP64 table, or its FP32
// copy (a view made by ytab_v<M, float>, whose lane offset counts 4 bytes)
template <typename V>
__device__ __forceinline__ V tget_v(const TabRef& b, int t, int idx) {
  if constexpr (sizeof(V) == 8) {
    return tget(b, t, idx);
  } else {
    const auto v = __builtin_amdgcn_raw_buffer_load_b32(
        b.r, b.lane_off, b.sbias + (t * b.ncol + b.off + idx) * (WAVE * 4), 0);
    return __builtin_bit_cast(float, v);
  }
}
// The same view, opaque to the optimiser: loads through it are issued again
// instead of keeping coefficients loaded earlier live across a scan (table
// reloads hit the cache; held coefficients cost registers and occupancy).
__device__ __forceinline__ TabRef relaunder(const TabRef& b) {
  int z = b.sbias;
  asm volatile("" : "+s"(z));
  return TabRef{b.r, b.ncol, b.off, b.lane_off, z};
}
// The same, additionally ordered after `dep` is computed: a level of a scan
// loads its table entries once the previous level's carry exists, all
// together, instead of the compiler hoisting every level's loads to the
// start (and spilling them) or serialising them one by one.
template <typename V>
__device__ __forceinline__ TabRef relaunder_after(const TabRef& b, V dep) {
  int z = b.sbias;
  asm volatile("" : "+s"(z) : "v"(dep));
  return TabRef{b.r, b.ncol, b.off, b.lane_off, z};
}
// Entries [0, K) of row t of a table view, issued together, and the wait
// for a row loaded ahead, taken once for the whole row: left to itself the
// scheduler issues each coefficient load next to its product, one memory
// round trip per coefficient.
template <int K, typename V>
__device__ __forceinline__ void row_ld(const TabRef& b, int t, V (&c)[K]) {
#pragma unroll
  for (int i = 0; i < K; ++i) c[i] = tget_v<V>(b, t, i);
}
template <int K, typename V>
__device__ __forceinline__ void row_take(const V (&n)[K], V (&c)[K]) {
#pragma unroll
  for (int i = 0; i < K; ++i) {
    c[i] = n[i];
    asm volatile("" : "+v"(c[i]));
  }
}
// index of entry `idx` of grid row `row` in a lane-major table
template <int M>
__host__ __device__ inline int lane_major_index(int row, int ncol, int idx) {
  const int l = row / M, t = row - l * M;
  return (t * ncol + idx) * WAVE + l;
}
// entry `idx` of grid row `row` (any lane's) of a lane-major table
template <int M>
__device__ __forceinline__ double trow(const double* base, int ncol, int idx, int row) {
  return base[lane_major_index<M>(row, ncol, idx)];
}

// ---- wavefront scans of affine maps s_out = A s_in + c -----------------
template <int P, int NC>
__device__ __forceinline__ void scan_affine_up(double (&A)[P][P], double (&c)[NC][P],
                                               int lane) {
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
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
  for (int d = 1; d < WAVE; d <<= 1) {
    double Ap[P][P], cp[NC][P];
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
      for (int b = 0; b < P; ++b) Ap[q][b] = __shfl_down(A[q][b], d);
#pragma unroll
    for (int n = 0; n < NC; ++n)
#pragma unroll
      for (int q = 0; q < P; ++q) cp[n][q] = __shfl_down(c[n][q], d);
    if (lane + d < WAVE) {
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

// ---- unit lower solve: w_r = x_r - sum_{i=1..P} L(r,i) w_{r-i} ------------
// lo(t, i-1) = L(r0+t, i): multiplier of row r0+t-i into row r0+t.
template <int M, int P, int NC>
__device__ __forceinline__ void lsolve(const TabRef lo, int lane, double (&x)[NC][M]) {
  static_assert(P <= M, "chunk shorter than bandwidth");
  // (1) in place, zero incoming state
#pragma unroll
  for (int t = 1; t < M; ++t)
#pragma unroll
    for (int i = 1; i <= P; ++i) {
      if (t - i < 0) continue;
      const double l = tget(lo, t, i - 1);
#pragma unroll
      for (int n = 0; n < NC; ++n) x[n][t] -= l * x[n][t - i];
    }
  op_fence();
  // carry: last P rows, c[q] = w_{r0+M-1-q}
  double c[NC][P];
#pragma unroll
  for (int n = 0; n < NC; ++n)
#pragma unroll
    for (int q = 0; q < P; ++q) c[n][q] = x[n][M - 1 - q];
  // (2) response of the carry to a unit incoming state s_b = w_{r0-1-b}
  double A[P][P];
#pragma unroll
  for (int b = 0; b < P; ++b) {
    double g[M];
#pragma unroll
    for (int t = 0; t < M; ++t) {
      double acc = 0.0;
#pragma unroll
      for (int i = 1; i <= P; ++i) {
        if (t - i >= 0) acc -= tget(lo, t, i - 1) * g[t - i];
        else if (i - t - 1 == b) acc -= tget(lo, t, i - 1);
      }
      g[t] = acc;
    }
#pragma unroll
    for (int q = 0; q < P; ++q) A[q][b] = g[M - 1 - q];
  }
  op_fence();
  scan_affine_up<P, NC>(A, c, lane);
  op_fence();
  // (3) exact incoming state and the correction it causes
  double s[NC][P];
#pragma unroll
  for (int n = 0; n < NC; ++n)
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const double v = __shfl_up(c[n][q], 1);
      s[n][q] = lane == 0 ? 0.0 : v;
    }
  const TabRef lo2 = relaunder(lo);
  double dl[NC][M];
#pragma unroll
  for (int t = 0; t < M; ++t) {
#pragma unroll
    for (int n = 0; n < NC; ++n) dl[n][t] = 0.0;
#pragma unroll
    for (int i = 1; i <= P; ++i) {
      const double l = tget(lo2, t, i - 1);
#pragma unroll
      for (int n = 0; n < NC; ++n) {
        const double pv = (t - i >= 0) ? dl[n][t - i] : s[n][i - t - 1];
        dl[n][t] -= l * pv;
      }
    }
#pragma unroll
    for (int n = 0; n < NC; ++n) x[n][t] += dl[n][t];
  }
}

// ---- upper solve: x_r = (w_r - sum_{d=1..P} U(r,d) x_{r+d}) * inv_r ------
template <int M, int P, int NC>
__device__ __forceinline__ void usolve(const TabRef up, const TabRef inv, int lane,
                                       double (&x)[NC][M]) {
#pragma unroll
  for (int t = M - 1; t >= 0; --t) {
#pragma unroll
    for (int d = 1; d <= P; ++d) {
      if (t + d >= M) continue;
      const double u = tget(up, t, d - 1);
#pragma unroll
      for (int n = 0; n < NC; ++n) x[n][t] -= u * x[n][t + d];
    }
    const double iv = tget(inv, t, 0);
#pragma unroll
    for (int n = 0; n < NC; ++n) x[n][t] *= iv;
  }
  op_fence();
  // carry: first P rows, c[q] = x_{r0+q}
  double c[NC][P];
#pragma unroll
  for (int n = 0; n < NC; ++n)
#pragma unroll
    for (int q = 0; q < P; ++q) c[n][q] = x[n][q];
  double A[P][P];
#pragma unroll
  for (int b = 0; b < P; ++b) {
    double g[M];
#pragma unroll
    for (int t = M - 1; t >= 0; --t) {
      double acc = 0.0;
#pragma unroll
      for (int d = 1; d <= P; ++d) {
        if (t + d < M) acc -= tget(up, t, d - 1) * g[t + d];
        else if (t + d - M == b) acc -= tget(up, t, d - 1);
      }
      g[t] = acc * tget(inv, t, 0);
    }
#pragma unroll
    for (int q = 0; q < P; ++q) A[q][b] = g[q];
  }
  op_fence();
  scan_affine_down<P, NC>(A, c, lane);
  op_fence();
  double s[NC][P];
#pragma unroll
  for (int n = 0; n < NC; ++n)
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const double v = __shfl_down(c[n][q], 1);
      s[n][q] = lane == WAVE - 1 ? 0.0 : v;
    }
  const TabRef up2 = relaunder(up), inv2 = relaunder(inv);
  double dl[NC][M];
#pragma unroll
  for (int t = M - 1; t >= 0; --t) {
#pragma unroll
    for (int n = 0; n < NC; ++n) dl[n][t] = 0.0;
#pragma unroll
    for (int d = 1; d <= P; ++d) {
      const double u = tget(up2, t, d - 1);
#pragma unroll
      for (int n = 0; n < NC; ++n) {
        const double pv = (t + d < M) ? dl[n][t + d] : s[n][t + d - M];
        dl[n][t] -= u * pv;
      }
    }
    const double iv = tget(inv2, t, 0);
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      dl[n][t] *= iv;
      x[n][t] += dl[n][t];
    }
  }
}

// ---- triangular solves with precomputed scan matrices -------------------
// For a fixed factor the carry maps of the lanes do not depend on the data,
// so the Hillis-Steele compositions are tabulated once per grid:
// H[k][q][b][lane] (lane-major) is the map from the carry of lane l-2^k
// (l+2^k for the upper solve) to that of lane l after the first k levels.
// Only the first `levels` levels are applied: the compositions beyond them
// are below 1e-22 in max norm (the factors' inverses decay geometrically),
// checked when the tables are built.
template <int M, int P, int NC, typename V>
__device__ __forceinline__ void lsolve_tab(const TabRef lo, const TabRef H, int levels,
                                           int lane, V (&x)[NC][M]) {
#pragma unroll
  for (int t = 1; t < M; ++t)
#pragma unroll
    for (int i = 1; i <= P; ++i) {
      if (t - i < 0) continue;
      const V l = tget_v<V>(lo, t, i - 1);
#pragma unroll
      for (int n = 0; n < NC; ++n) x[n][t] -= l * x[n][t - i];
    }
  op_fence();
  V c[NC][P];
#pragma unroll
  for (int n = 0; n < NC; ++n)
#pragma unroll
    for (int q = 0; q < P; ++q) c[n][q] = x[n][M - 1 - q];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (k < levels) {
      const int d = 1 << k;
      // the level's composed map is loaded before the shuffles and outside
      // the lane test, so its loads are in flight together
      const TabRef Hk = relaunder_after(H, c[0][0]);
      V h[P][P];
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int b = 0; b < P; ++b) h[q][b] = tget_v<V>(Hk, k * P + q, b);
      V cp[NC][P];
#pragma unroll
      for (int n = 0; n < NC; ++n)
#pragma unroll
        for (int q = 0; q < P; ++q) cp[n][q] = __shfl_up(c[n][q], d);
      if (lane >= d) {
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
          for (int b = 0; b < P; ++b)
#pragma unroll
            for (int n = 0; n < NC; ++n) c[n][q] += h[q][b] * cp[n][b];
      }
    }
  }
  op_fence();
  V s[NC][P];
#pragma unroll
  for (int n = 0; n < NC; ++n)
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const V v = __shfl_up(c[n][q], 1);
      s[n][q] = lane == 0 ? V(0) : v;
    }
  const TabRef lo2 = relaunder(lo);
  V dl[NC][M];
#pragma unroll
  for (int t = 0; t < M; ++t) {
#pragma unroll
    for (int n = 0; n < NC; ++n) dl[n][t] = V(0);
#pragma unroll
    for (int i = 1; i <= P; ++i) {
      const V l = tget_v<V>(lo2, t, i - 1);
#pragma unroll
      for (int n = 0; n < NC; ++n) {
        const V pv = (t - i >= 0) ? dl[n][t - i] : s[n][i - t - 1];
        dl[n][t] -= l * pv;
      }
    }
#pragma unroll
    for (int n = 0; n < NC; ++n) x[n][t] += dl[n][t];
  }
}

template <int M, int P, int NC, typename V>
__device__ __forceinline__ void usolve_tab(const TabRef up, const TabRef inv, const TabRef H,
                                           int levels, int lane, V (&x)[NC][M]) {
#pragma unroll
  for (int t = M - 1; t >= 0; --t) {
#pragma unroll
    for (int d = 1; d <= P; ++d) {
      if (t + d >= M) continue;
      const V u = tget_v<V>(up, t, d - 1);
#pragma unroll
      for (int n = 0; n < NC; ++n) x[n][t] -= u * x[n][t + d];
    }
    const V iv = tget_v<V>(inv, t, 0);
#pragma unroll
    for (int n = 0; n < NC; ++n) x[n][t] *= iv;
  }
  op_fence();
  V c[NC][P];
#pragma unroll
  for (int n = 0; n < NC; ++n)
#pragma unroll
    for (int q = 0; q < P; ++q) c[n][q] = x[n][q];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (k < levels) {
      const int d = 1 << k;
      const TabRef Hk = relaunder_after(H, c[0][0]);
      V h[P][P];
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int b = 0; b < P; ++b) h[q][b] = tget_v<V>(Hk, k * P + q, b);
      V cp[NC][P];
#pragma unroll
      for (int n = 0; n < NC; ++n)
#pragma unroll
        for (int q = 0; q < P; ++q) cp[n][q] = __shfl_down(c[n][q], d);
      if (lane + d < WAVE) {
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
          for (int b = 0; b < P; ++b)
#pragma unroll
            for (int n = 0; n < NC; ++n) c[n][q] += h[q][b] * cp[n][b];
      }
    }
  }
  op_fence();
  V s[NC][P];
#pragma unroll
  for (int n = 0; n < NC; ++n)
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const V v = __shfl_down(c[n][q], 1);
      s[n][q] = lane == WAVE - 1 ? V(0) : v;
    }
  const TabRef up2 = relaunder(up), inv2 = relaunder(inv);
  V dl[NC][M];
#pragma unroll
  for (int t = M - 1; t >= 0; --t) {
#pragma unroll
    for (int n = 0; n < NC; ++n) dl[n][t] = V(0);
#pragma unroll
    for (int d = 1; d <= P; ++d) {
      const V u = tget_v<V>(up2, t, d - 1);
#pragma unroll
      for (int n = 0; n < NC; ++n) {
        const V pv = (t + d < M) ? dl[n][t + d] : s[n][t + d - M];
        dl[n][t] -= u * pv;
      }
    }
    const V iv = tget_v<V>(inv2, t, 0);
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      dl[n][t] *= iv;
      x[n][t] += dl[n][t];
    }
  }
}

// ---- banded matvec y_r = sum_d B(r,d) u_{r-P+d} ---------------------------
// ROWS: each row's coefficients are loaded together one row ahead
// (row_ld / row_take); otherwise the scheduler places the loads.
template <int M, int P, int NC, typename V, bool ROWS = true>
__device__ __forceinline__ void bmatvec(const TabRef B, int lane, const V (&u)[NC][M],
                                        V (&y)[NC][M]) {
  V hl[NC][P], hr[NC][P];
#pragma unroll
  for (int n = 0; n < NC; ++n)
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const V a = __shfl_up(u[n][M - P + q], 1);
      const V b = __shfl_down(u[n][q], 1);
      hl[n][q] = lane == 0 ? V(0) : a;
      hr[n][q] = lane == WAVE - 1 ? V(0) : b;
    }
  V bn[2 * P + 1];
  if constexpr (ROWS) row_ld<2 * P + 1>(B, 0, bn);
#pragma unroll
  for (int t = 0; t < M; ++t) {
    V bc[2 * P + 1];
    if constexpr (ROWS) {
      row_take<2 * P + 1>(bn, bc);
      if (t + 1 < M) row_ld<2 * P + 1>(B, t + 1, bn);
      op_fence();
    } else {
      row_ld<2 * P + 1>(B, t, bc);
    }
    V acc[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) acc[n] = V(0);
#pragma unroll
    for (int d = 0; d <= 2 * P; ++d) {
      const int col = t - P + d;
      const V bv = bc[d];
#pragma unroll
      for (int n = 0; n < NC; ++n) {
        const V v = col < 0 ? hl[n][col + P] : (col >= M ? hr[n][col - M] : u[n][col]);
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

// Tables of one y grid, in order, in one allocation.  Per-row tables hold
// nb entries for each of 64*M rows (lane-major); the scan compositions of the
// fixed factors hold 6 levels of P x P maps for 64 lanes.
enum YTabId {
  YT_B1, YT_A1, YT_INV1, YT_UP1, YT_LO1, YT_FMAP,  // d/deta: B1, A1, LU of A1, fmap
  YT_B2, YT_A2, YT_INV2, YT_UP2, YT_LO2,           // d2/dy2: B2, A2, LU of A2
  YT_INVI, YT_UPI, YT_LOI,                         // integration: LU of B1
  YT_WBD,                                          // boundary d/dy functionals
  YT_NROW,                                         // (count of per-row tables)
  YT_HL1 = YT_NROW, YT_HU1, YT_HL2, YT_HU2, YT_HLI, YT_HUI,
  YT_COUNT
};
__host__ __device__ constexpr int ytab_width(int id) {
  // entries per row of the per-row tables; P of the scan tables
  return id == YT_B1 || id == YT_A1 ? 7
       : id == YT_B2 || id == YT_A2 ? 5
       : id == YT_UP1 || id == YT_LO1 || id == YT_UPI || id == YT_LOI ? 3
       : id == YT_UP2 || id == YT_LO2 || id == YT_WBD ? 2
       : id == YT_HL1 || id == YT_HU1 || id == YT_HLI || id == YT_HUI ? 3
       : id == YT_HL2 || id == YT_HU2 ? 2 : 1;
}
// element offset of table `id`; ytab_offset(YT_COUNT, M) is the total size
__host__ __device__ constexpr int ytab_offset(int id, int M) {
  int off = 0;
  for (int q = 0; q < id; ++q) {
    const int w = ytab_width(q);
    off += q < YT_NROW ? w * 64 * M : 6 * w * w * 64;
  }
  return off;
}

// Device tables of one y grid: one allocation, lane-major tables of 64*M
// rows each at the offsets ytab_offset(id, M) (loma_kernels.h), so that every
// table access goes through one buffer resource with a constant offset.
struct YTab {
  const double* base;
  const float* base32;  // the same tables rounded to FP32, same layout
  // scan level counts of the fixed factors (see lsolve_tab)
  int kL1, kU1, kL2, kU2, kLi, kUi;
  int my;
};

// view of table `id` (ncol entries per row) for this lane
template <int M>
__device__ __forceinline__ TabRef ytab(const YTab& T, int id, int ncol, int lane) {
  return TabRef{make_rsrc(T.base, 0x7fffffff), ncol, 0, lane * 8, ytab_offset(id, M) * 8};
}
// ... in the arithmetic type V (read with tget_v<V>)
template <int M, typename V>
__device__ __forceinline__ TabRef ytab_v(const YTab& T, int id, int ncol, int lane) {
  if constexpr (sizeof(V) == 8)
    return ytab<M>(T, id, ncol, lane);
  else
    return TabRef{make_rsrc(T.base32, 0x7fffffff), ncol, 0, lane * 4, ytab_offset(id, M) * 4};
}

// first derivative du = fmap * A1^{-1} B1 u   (deryr2, cfdiff.f:518-601)
template <int M, int NC, typename V, bool ROWS = true>
__device__ __forceinline__ void op_dy1(const YTab& T, int lane, const V (&u)[NC][M],
                                       V (&du)[NC][M], bool apply_fmap = true) {
  bmatvec<M, 3, NC, V, ROWS>(ytab_v<M, V>(T, YT_B1, 7, lane), lane, u, du);
  op_fence();
  lsolve_tab<M, 3, NC>(ytab_v<M, V>(T, YT_LO1, 3, lane), ytab_v<M, V>(T, YT_HL1, 3, lane), T.kL1, lane, du);
  op_fence();
  usolve_tab<M, 3, NC>(ytab_v<M, V>(T, YT_UP1, 3, lane), ytab_v<M, V>(T, YT_INV1, 1, lane),
                       ytab_v<M, V>(T, YT_HU1, 3, lane), T.kU1, lane, du);
  op_fence();
  if (apply_fmap) {
    const TabRef fm = ytab_v<M, V>(T, YT_FMAP, 1, lane);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      const V f = tget_v<V>(fm, t, 0);
#pragma unroll
      for (int n = 0; n < NC; ++n) du[n][t] *= f;
    }
  }
  op_fence();
}

// second derivative d2u = A2^{-1} B2 u   (deryyr2, cfdiff.f:735-797)
template <int M, int NC, typename V, bool ROWS = true>
__device__ __forceinline__ void op_dy2(const YTab& T, int lane, const V (&u)[NC][M],
                                       V (&d2u)[NC][M]) {
  bmatvec<M, 2, NC, V, ROWS>(ytab_v<M, V>(T, YT_B2, 5, lane), lane, u, d2u);
  op_fence();
  lsolve_tab<M, 2, NC>(ytab_v<M, V>(T, YT_LO2, 2, lane), ytab_v<M, V>(T, YT_HL2, 2, lane), T.kL2, lane, d2u);
  op_fence();
  usolve_tab<M, 2, NC>(ytab_v<M, V>(T, YT_UP2, 2, lane), ytab_v<M, V>(T, YT_INV2, 1, lane),
                       ytab_v<M, V>(T, YT_HU2, 2, lane), T.kU2, lane, d2u);
  op_fence();
}

// Helmholtz/Poisson solve (dt22 - rK dt21) v = rhs with identity end rows,
// LU per mode class in `lu` (lane-major, 5 entries per row: inv, up1, up2,
// lo1, lo2), the RHS being dt21*f with zero end rows (calcrhslap,
// laps.f:17-46).
template <int M, int NC>
__device__ __forceinline__ void op_helm(const YTab& T, const double* __restrict__ lu,
                                        int lane, const double (&f)[NC][M],
                                        double (&v)[NC][M]) {
  bmatvec<M, 2, NC, double, false>(ytab<M>(T, YT_A2, 5, lane), lane, f, v);
  const int r0 = lane * M;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const int r = r0 + t;
    if (r == 0 || r == T.my - 1) {
#pragma unroll
      for (int n = 0; n < NC; ++n) v[n][t] = 0.0;
    }
  }
  op_fence();
  lsolve<M, 2, NC>(tab(lu, 5, 3, lane), lane, v);
  op_fence();
  usolve<M, 2, NC>(tab(lu, 5, 1, lane), tab(lu, 5, 0, lane), lane, v);
  op_fence();
}

// ---- column staging through LDS ------------------------------------------
// A wavefront moves a complex column between HBM and its registers through
// a private LDS buffer of 64*M double2: the HBM side is contiguous across
// lanes (row p = t*64 + lane), the register side is the lane's own chunk.
// Only the owning wavefront touches its buffer, so a wavefront-scope fence
// orders the cross-lane accesses.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// one element of a complex column: b64 (float2) or b128 (double2) buffer op
// (the loaded vector is bit-cast as a whole: bit-casting its elements one by
// one miscompiles to two copies of element 0 with this toolchain)
__device__ __forceinline__ double2 buf_ld(__amdgpu_buffer_rsrc_t r, int voff, float2*) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0);
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  return make_double2((double)__builtin_bit_cast(float, (unsigned)u),
                      (double)__builtin_bit_cast(float, (unsigned)(u >> 32)));
}
__device__ __forceinline__ double2 buf_ld(__amdgpu_buffer_rsrc_t r, int voff, double2*) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0);
  const unsigned long long a = (unsigned long long)v[0] | ((unsigned long long)v[1] << 32);
  const unsigned long long b = (unsigned long long)v[2] | ((unsigned long long)v[3] << 32);
  return make_double2(__builtin_bit_cast(double, a), __builtin_bit_cast(double, b));
}
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, int voff, double2 x, float2*) {
  __attribute__((ext_vector_type(2))) unsigned v;
  v[0] = __builtin_bit_cast(unsigned, (float)x.x);
  v[1] = __builtin_bit_cast(unsigned, (float)x.y);
  __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, 0, 0);
}
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, int voff, double2 x, double2*) {
  const unsigned long long a = __builtin_bit_cast(unsigned long long, x.x);
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x.y);
  __attribute__((ext_vector_type(4))) unsigned v;
  v[0] = (unsigned)a; v[1] = (unsigned)(a >> 32);
  v[2] = (unsigned)b; v[3] = (unsigned)(b >> 32);
  __builtin_amdgcn_raw_buffer_store_b128(v, r, voff, 0, 0);
}

// Offset of row t*64 + lane of a column, opaque to the optimiser: the
// constant part must stay in the vector offset, which the range check
// covers (split off into the scalar offset it would escape the check).
template <typename T2>
__device__ __forceinline__ int col_voff(int lane, int t) {
  int v = (t * WAVE + lane) * (int)sizeof(T2);
  asm volatile("" : "+v"(v));
  return v;
}

// HBM column (T2 = float2 or double2; `col` wavefront-uniform) -> LDS stage.
// Row p = t*64 + lane; one resource spans rows 0..my-1, so rows >= my read
// as 0 without a test.  All M loads are issued before the first is used.
template <int M, typename T2>
__device__ __forceinline__ void stage_in(const T2* __restrict__ col, int my, int lane,
                                         double2* st) {
  const auto r = make_rsrc(col, my * (int)sizeof(T2));
  double2 q[M];
#pragma unroll
  for (int t = 0; t < M; ++t) q[t] = buf_ld(r, col_voff<T2>(lane, t), (T2*)nullptr);
  wave_sync();
#pragma unroll
  for (int t = 0; t < M; ++t) st[t * WAVE + lane] = q[t];
  wave_sync();
}
// LDS stage -> HBM column (rows >= my are dropped by the range check)
template <int M, typename T2>
__device__ __forceinline__ void stage_out(const double2* st, int my, int lane,
                                          T2* __restrict__ col) {
  wave_sync();
  double2 v[M];
#pragma unroll
  for (int t = 0; t < M; ++t) v[t] = st[t * WAVE + lane];
  wave_sync();
  const auto r = make_rsrc(col, my * (int)sizeof(T2));
#pragma unroll
  for (int t = 0; t < M; ++t) buf_st(r, col_voff<T2>(lane, t), v[t], (T2*)nullptr);
}
// stage -> registers (own chunk)
template <int M>
__device__ __forceinline__ void stage_get(const double2* st, int lane, double (&x)[2][M]) {
  const int r0 = lane * M;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const double2 v = st[r0 + t];
    x[0][t] = v.x;
    x[1][t] = v.y;
  }
}
// registers -> stage (own chunk)
template <int M>
__device__ __forceinline__ void stage_put(double2* st, int lane, const double (&x)[2][M]) {
  const int r0 = lane * M;
#pragma unroll
  for (int t = 0; t < M; ++t) st[r0 + t] = make_double2(x[0][t], x[1][t]);
}

template <int M, typename T2>
__device__ __forceinline__ void load_col(const T2* __restrict__ col, int my, int lane,
                                         double2* st, double (&x)[2][M]) {
  stage_in<M>(col, my, lane, st);
  stage_get<M>(st, lane, x);
}
template <int M, typename T2>
__device__ __forceinline__ void store_col(T2* __restrict__ col, int my, int lane,
                                          double2* st, const double (&x)[2][M]) {
  wave_sync();
  stage_put<M>(st, lane, x);
  stage_out<M>(st, my, lane, col);
}

// ---- FP32 staging (the mixed form's phase C) ------------------------------
// The same movements with a float2 stage (half the LDS of a double2 stage)
// for FP32 spectral columns, and FP64 state updates done in the coalesced
// order straight from HBM: the state never passes through registers in
// chunk order, so it needs no FP64 stage.
__device__ __forceinline__ float2 buf_ld32(__amdgpu_buffer_rsrc_t r, int voff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0);
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  return make_float2(__builtin_bit_cast(float, (unsigned)u),
                     __builtin_bit_cast(float, (unsigned)(u >> 32)));
}
template <int M>
__device__ __forceinline__ void stage_in32(const float2* __restrict__ col, int my, int lane,
                                           float2* st) {
  const auto r = make_rsrc(col, my * 8);
  float2 q[M];
#pragma unroll
  for (int t = 0; t < M; ++t) q[t] = buf_ld32(r, col_voff<float2>(lane, t));
  wave_sync();
#pragma unroll
  for (int t = 0; t < M; ++t) st[t * WAVE + lane] = q[t];
  wave_sync();
}
template <int M>
__device__ __forceinline__ void stage_get32(const float2* st, int lane, float (&x)[2][M]) {
  const int r0 = lane * M;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const float2 v = st[r0 + t];
    x[0][t] = v.x;
    x[1][t] = v.y;
  }
}
template <int M>
__device__ __forceinline__ void stage_put32(float2* st, int lane, const float (&x)[2][M]) {
  const int r0 = lane * M;
#pragma unroll
  for (int t = 0; t < M; ++t) st[r0 + t] = make_float2(x[0][t], x[1][t]);
}
template <int M>
__device__ __forceinline__ void load_col32(const float2* __restrict__ col, int my, int lane,
                                           float2* st, float (&x)[2][M]) {
  stage_in32<M>(col, my, lane, st);
  stage_get32<M>(st, lane, x);
}
// FP64 state column += s * x and the FP64 history column = x, for an FP32
// column x (own chunk): staged to the coalesced order, then each lane
// updates its rows p = t*64 + lane in FP64 (rows >= my dropped by the range
// checks)
template <int M>
__device__ __forceinline__ void update_col32(double2* __restrict__ state,
                                             double2* __restrict__ hist, int my, int lane,
                                             float2* st, const float (&x)[2][M], double s) {
  wave_sync();
  stage_put32<M>(st, lane, x);
  wave_sync();
  float2 v[M];
#pragma unroll
  for (int t = 0; t < M; ++t) v[t] = st[t * WAVE + lane];
  wave_sync();
  const auto rs = make_rsrc(state, my * 16), rh = make_rsrc(hist, my * 16);
  double2 h[M];
#pragma unroll
  for (int t = 0; t < M; ++t) h[t] = buf_ld(rs, col_voff<double2>(lane, t), (double2*)nullptr);
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const double2 a = make_double2((double)v[t].x, (double)v[t].y);
    buf_st(rs, col_voff<double2>(lane, t), make_double2(h[t].x + s * a.x, h[t].y + s * a.y),
           (double2*)nullptr);
    buf_st(rh, col_voff<double2>(lane, t), a, (double2*)nullptr);
  }
}
// FP64 state column += a spectral increment (T2 = float2 or double2; plus,
// for the 00 mode, its FP64 mean profile `mean`; null otherwise), in the
// coalesced order
template <int M, typename T2>
__device__ __forceinline__ void add_col(double2* __restrict__ state, const T2* __restrict__ inc,
                                        int my, int lane, const double* __restrict__ mean) {
  const auto rs = make_rsrc(state, my * 16), ri = make_rsrc(inc, my * (int)sizeof(T2));
  double2 h[M], d[M];
#pragma unroll
  for (int t = 0; t < M; ++t) {
    h[t] = buf_ld(rs, col_voff<double2>(lane, t), (double2*)nullptr);
    d[t] = buf_ld(ri, col_voff<T2>(lane, t), (T2*)nullptr);
  }
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const int p = t * WAVE + lane;
    double dr = d[t].x;
    if (mean != nullptr && p < my) dr += mean[p];
    buf_st(rs, col_voff<double2>(lane, t), make_double2(h[t].x + dr, h[t].y + d[t].y),
           (double2*)nullptr);
  }
}
template <int M>
__device__ __forceinline__ void add_col32(double2* __restrict__ state,
                                          const float2* __restrict__ inc, int my, int lane,
                                          const double* __restrict__ mean) {
  add_col<M, float2>(state, inc, my, lane, mean);
}
// FP64 state column += s * x and the history column = x, for an FP64 column
// x (own chunk), through the stage in the coalesced order (update_col32's
// FP64 form)
template <int M>
__device__ __forceinline__ void update_col(double2* __restrict__ state,
                                           double2* __restrict__ hist, int my, int lane,
                                           double2* st, const double (&x)[2][M], double s) {
  wave_sync();
  stage_put<M>(st, lane, x);
  wave_sync();
  double2 v[M];
#pragma unroll
  for (int t = 0; t < M; ++t) v[t] = st[t * WAVE + lane];
  wave_sync();
  const auto rs = make_rsrc(state, my * 16), rh = make_rsrc(hist, my * 16);
  double2 h[M];
#pragma unroll
  for (int t = 0; t < M; ++t) h[t] = buf_ld(rs, col_voff<double2>(lane, t), (double2*)nullptr);
#pragma unroll
  for (int t = 0; t < M; ++t) {
    buf_st(rs, col_voff<double2>(lane, t), make_double2(h[t].x + s * v[t].x, h[t].y + s * v[t].y),
           (double2*)nullptr);
    buf_st(rh, col_voff<double2>(lane, t), v[t], (double2*)nullptr);
  }
}

}  // namespace loma
