// loma_yops_impl.h -- y-direction kernels: per-operator entry points and the
// fused per-mode phase kernels of an RK3 substep.
//
// Phase A  = trunk/lomacte3/src/cross.f:315-554 (BCs, Lapvdvhom, m_x/m_z,
//   xi-update, rho*u_i, T derivatives) producing the transform inputs.
// Phase C  = cross.f:619-1023 (T update, N_i, vorticities, Laplacians,
//   RHS, gamma-update, 00 modes of u and w) consuming the transform outputs.
// Phase C2 = cross.f:1027-1104 (00 mode of rho*v, solvepsi).
// One 64-lane wavefront per (kx,kz) mode; see loma_dev.h for the solver.
// Columns move between HBM and registers through a per-wavefront LDS stage,
// and every phase is ordered so that at most four columns are live at once.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "loma_kernels.h"

namespace loma {

static __constant__ double c_gama[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static __constant__ double c_xi[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};

// per-wavefront LDS stage of a 256-thread block
#define LOMA_STAGE(M)                                 \
  __shared__ double2 stage_buf_[4][WAVE * (M)];       \
  double2* st = stage_buf_[threadIdx.x >> 6]
// stage plus a second per-wavefront LDS column where a phase parks a value
#define LOMA_STAGE2(M)                                \
  __shared__ double2 stage_buf_[4][2][WAVE * (M)];    \
  double2* st = stage_buf_[threadIdx.x >> 6][0];      \
  double2* pk = stage_buf_[threadIdx.x >> 6][1]
// the same with float2 columns (FP32 phase C)
#define LOMA_STAGE2_F32(M)                            \
  __shared__ float2 stage_buf_[4][2][WAVE * (M)];     \
  float2* st = stage_buf_[threadIdx.x >> 6][0];       \
  float2* pk = stage_buf_[threadIdx.x >> 6][1]

// sum over the wavefront
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// ------------------------------------------------------------------------
// Neumann (solvepsi) pieces
// ------------------------------------------------------------------------
// dot products of a complex column with the boundary-derivative functionals
// w0 (d/dy at row 0) and w1 (d/dy at row my-1): out = {w0.re, w0.im, w1.re,
// w1.im}.
template <int M>
__device__ __forceinline__ void bd_dots(const YTab& T, int lane, const double (&x)[2][M],
                                        double (&out)[4]) {
  const TabRef w = ytab<M>(T, YT_WBD, 2, lane);
  double s0r = 0.0, s0i = 0.0, s1r = 0.0, s1i = 0.0;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const double a = tget(w, t, 0), b = tget(w, t, 1);
    s0r += a * x[0][t];
    s0i += a * x[1][t];
    s1r += b * x[0][t];
    s1i += b * x[1][t];
  }
  out[0] = wave_sum(s0r);
  out[1] = wave_sum(s0i);
  out[2] = wave_sum(s1r);
  out[3] = wave_sum(s1i);
}

// v solving (dt22 - rK dt21) v = dt21 f in the interior with dv/dy = 0 at
// both ends (solvepsi, laps.f:580-726).  The reference adds two homogeneous
// Dirichlet solutions h1, h2 to the zero-Dirichlet particular solution; here
// the boundary derivatives of h1, h2 are per-class constants
// bc = {w0.h1, w0.h2, w1.h1, w1.h2}, the particular solution's are two dot
// products, and the combination is a second Dirichlet solve whose end values
// are the 2x2 solution (same linear algebra, half the live columns).
template <int M>
__device__ __forceinline__ void solve_neumann(const YTab& T, const double* lu,
                                              const double* bc, int lane,
                                              const double (&f)[2][M], double (&v)[2][M]) {
  op_helm<M, 2>(T, lu, lane, f, v);
  double dp[4];
  bd_dots<M>(T, lane, v, dp);
  const double a11 = bc[0], a12 = bc[1], a21 = bc[2], a22 = bc[3];
  const double det = a11 * a22 - a12 * a21;
  const double cr1 = (-dp[0] * a22 + a12 * dp[2]) / det;
  const double cr2 = (-a11 * dp[2] + a21 * dp[0]) / det;
  const double ci1 = (-dp[1] * a22 + a12 * dp[3]) / det;
  const double ci2 = (-a11 * dp[3] + a21 * dp[1]) / det;
  // second solve with Dirichlet end values (c1, c2)
  bmatvec<M, 2, 2>(ytab<M>(T, YT_A2, 5, lane), lane, f, v);
  const int r0 = lane * M;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const int r = r0 + t;
    if (r == 0) { v[0][t] = cr1; v[1][t] = ci1; }
    if (r == T.my - 1) { v[0][t] = cr2; v[1][t] = ci2; }
  }
  lsolve<M, 2, 2>(tab(lu, 5, 3, lane), lane, v);
  usolve<M, 2, 2>(tab(lu, 5, 1, lane), tab(lu, 5, 0, lane), lane, v);
}

// per-class Neumann constants: h1 = LU^{-1} e_0, h2 = LU^{-1} e_{my-1};
// bc = {w0.h1, w0.h2, w1.h1, w1.h2}
template <int M>
__global__ __launch_bounds__(256) void k_helm_bc(YTab T, const double* lu, int lu_stride,
                                                 double* bc, double* hout, int nclass) {
  const int wave = wave_uniform(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (wave >= nclass) return;
  const double* L = lu + (size_t)wave * lu_stride;
  const int r0 = lane * M;
  double h[2][M];
#pragma unroll
  for (int t = 0; t < M; ++t) {
    h[0][t] = (r0 + t == 0) ? 1.0 : 0.0;
    h[1][t] = (r0 + t == T.my - 1) ? 1.0 : 0.0;
  }
  lsolve<M, 2, 2>(tab(L, 5, 3, lane), lane, h);
  usolve<M, 2, 2>(tab(L, 5, 1, lane), tab(L, 5, 0, lane), lane, h);
  if (hout) {
    // h1, h2 lane-major (entry 0: h1, entry 1: h2 of row lane*M + t)
    double* H = hout + (size_t)wave * 2 * WAVE * M;
#pragma unroll
    for (int t = 0; t < M; ++t) {
      H[(t * 2 + 0) * WAVE + lane] = h[0][t];
      H[(t * 2 + 1) * WAVE + lane] = h[1][t];
    }
  }
  double d[4];
  bd_dots<M>(T, lane, h, d);  // {w0.h1, w0.h2, w1.h1, w1.h2}
  if (lane == 0) {
    bc[(size_t)wave * 4 + 0] = d[0];
    bc[(size_t)wave * 4 + 1] = d[1];
    bc[(size_t)wave * 4 + 2] = d[2];
    bc[(size_t)wave * 4 + 3] = d[3];
  }
}

// ------------------------------------------------------------------------
// per-operator kernels
// ------------------------------------------------------------------------
template <int M, int ORDER>
__global__ __launch_bounds__(256) void k_dy(YTab T, const double2* u, double2* du,
                                            int ncols) {
  LOMA_STAGE(M);
  const int wave = wave_uniform(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (wave >= ncols) return;
  const int my = T.my;
  double x[2][M], y[2][M];
  load_col<M>(u + (size_t)wave * my, my, lane, st, x);
  if (ORDER == 1)
    op_dy1<M, 2>(T, lane, x, y);
  else
    op_dy2<M, 2>(T, lane, x, y);
  store_col<M>(du + (size_t)wave * my, my, lane, st, y);
}

// mode 0: Lapvdvhom (laps.f:319); mode 1: solvepsi (laps.f:580);
// mode 2: Dirichlet Helmholtz (visc kernel), no derivative output
template <int M>
__global__ __launch_bounds__(256) void k_lapvdvhom(YTab T, const double* lu, int lu_stride,
                                                   const double* bc, const double2* phi,
                                                   double2* v, double2* dv,
                                                   const double* rK, int ncols, int mode) {
  LOMA_STAGE(M);
  const int wave = wave_uniform(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (wave >= ncols) return;
  const int my = T.my;
  double f[2][M], x[2][M];
  load_col<M>(phi + (size_t)wave * my, my, lane, st, f);
  const double* L = lu + (size_t)wave * lu_stride;
  if (mode != 2 && rK[wave] == 0.0) {
#pragma unroll
    for (int t = 0; t < M; ++t) x[0][t] = x[1][t] = 0.0;
  } else if (mode == 1) {
    solve_neumann<M>(T, L, bc + (size_t)wave * 4, lane, f, x);
  } else {
    op_helm<M, 2>(T, L, lane, f, x);
  }
  store_col<M>(v + (size_t)wave * my, my, lane, st, x);
  if (mode != 2) {
    op_dy1<M, 2>(T, lane, x, f);
    store_col<M>(dv + (size_t)wave * my, my, lane, st, f);
  }
}

// visc (trunk/mixing/src/laps.f:535-857) for a batch of mode columns, with
// the mixing-layer boundary functions (laps.f:893-1060).  lu1 / lu2: per
// column LU of the Dirichlet Helmholtz matrices dt22 - rk1 dt21 and
// dt22 - rk2 dt21; bc2: per column {w0.h1, w0.h2, w1.h1, w1.h2} of the rk2
// homogeneous solutions (w0, w1 the d/dy functionals at the ends).
//   omega = (D2 - rk1)^{-1} vorwk with omega = 0 at both ends (C_ome = 1,
//     C_dyo = 0 make the 2x2 boundary system the identity);
//   phi   = (D2 - rk1)^{-1} phiwk, zero at both ends (C_phi rows bot2/top2);
//   v     = (D2 - rk2)^{-1} phi with the Robin rows sqrt(rk2) v - dv/dy = 0
//     (bottom) and sqrt(rk2) v + dv/dy = 0 (top): of the reference's 4x4
//     system the phi rows fix the phi coefficients to zero and leave this
//     2x2 for the end values (c3, c4) of v, solved here in closed form;
//   vor = dalbe omega + dtri vorwk + dtxi v_in, phi likewise with dvdy_in
//     (dalbe != 0), else omega and phi; v = dvdy = 0 when rk2 = 0.
template <int M>
__global__ __launch_bounds__(256) void k_visc(YTab T, const double* lu1, const double* lu2,
                                              int lu_stride, const double* bc2,
                                              double2* vorwk, double2* vor, double2* phi,
                                              double2* v, double2* dvdy, double2* phiwk,
                                              const double* rk2, double dalbe, double dtri,
                                              double dtxi, int ncols) {
  LOMA_STAGE(M);
  const int col = wave_uniform(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (col >= ncols) return;
  const int my = T.my, r0 = lane * M;
  const size_t o = (size_t)col * my;
  const double* L1 = lu1 + (size_t)col * lu_stride;
  const double* L2 = lu2 + (size_t)col * lu_stride;
  double f[2][M], x[2][M];
  // omega and vor
  load_col<M>(vorwk + o, my, lane, st, f);
  op_helm<M, 2>(T, L1, lane, f, x);
  store_col<M>(vorwk + o, my, lane, st, x);
  if (dalbe != 0.0) {
    double g[2][M];
    load_col<M>(v + o, my, lane, st, g);
#pragma unroll
    for (int t = 0; t < M; ++t)
#pragma unroll
      for (int c = 0; c < 2; ++c) x[c][t] = dalbe * x[c][t] + dtri * f[c][t] + dtxi * g[c][t];
  }
  store_col<M>(vor + o, my, lane, st, x);
  // phi
  load_col<M>(phiwk + o, my, lane, st, f);
  op_helm<M, 2>(T, L1, lane, f, x);
  store_col<M>(phiwk + o, my, lane, st, x);
  {
    double g[2][M];
    if (dalbe != 0.0) {
      load_col<M>(dvdy + o, my, lane, st, g);
#pragma unroll
      for (int t = 0; t < M; ++t)
#pragma unroll
        for (int c = 0; c < 2; ++c) g[c][t] = dalbe * x[c][t] + dtri * f[c][t] + dtxi * g[c][t];
    } else {
#pragma unroll
      for (int t = 0; t < M; ++t) g[0][t] = x[0][t], g[1][t] = x[1][t];
    }
    store_col<M>(phi + o, my, lane, st, g);
  }
  // v, dv/dy
  const double k2 = rk2[col];
  if (k2 == 0.0) {
#pragma unroll
    for (int t = 0; t < M; ++t) f[0][t] = f[1][t] = 0.0;
    store_col<M>(v + o, my, lane, st, f);
    store_col<M>(dvdy + o, my, lane, st, f);
    return;
  }
  {
    double vp[2][M], dp[4];
    op_helm<M, 2>(T, L2, lane, x, vp);
    bd_dots<M>(T, lane, vp, dp);  // {w0.vp re, im, w1.vp re, im}
    const double* bc = bc2 + (size_t)col * 4;
    const double k = sqrt(k2);
    const double m11 = k - bc[0], m12 = -bc[1], m21 = bc[2], m22 = k + bc[3];
    const double det = m11 * m22 - m12 * m21;
    const double c3r = (dp[0] * m22 + m12 * dp[2]) / det;
    const double c4r = (-m11 * dp[2] - m21 * dp[0]) / det;
    const double c3i = (dp[1] * m22 + m12 * dp[3]) / det;
    const double c4i = (-m11 * dp[3] - m21 * dp[1]) / det;
    // second Dirichlet solve with end values (c3, c4)
    bmatvec<M, 2, 2>(ytab<M>(T, YT_A2, 5, lane), lane, x, vp);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      const int r = r0 + t;
      if (r == 0) { vp[0][t] = c3r; vp[1][t] = c3i; }
      if (r == my - 1) { vp[0][t] = c4r; vp[1][t] = c4i; }
    }
    lsolve<M, 2, 2>(tab(L2, 5, 3, lane), lane, vp);
    usolve<M, 2, 2>(tab(L2, 5, 1, lane), tab(L2, 5, 0, lane), lane, vp);
    store_col<M>(v + o, my, lane, st, vp);
    op_dy1<M, 2>(T, lane, vp, f);
    store_col<M>(dvdy + o, my, lane, st, f);
  }
}

// inty8 (laps.f:484-575): u = B1^{-1} A1 (du/fmap) - value at row 0, with
// the last row of B1 replaced by the gauge row u(my-1) = 0 (the reference's
// B1 is singular up to rounding; this is the well-posed form of the same
// integral, fixed afterwards by the row-0 shift).
template <int M>
__device__ __forceinline__ void op_inty(const YTab& T, int lane, double (&x)[1][M]) {
  double f[1][M];
  const int r0 = lane * M;
  const TabRef fm = ytab<M>(T, YT_FMAP, 1, lane);
#pragma unroll
  for (int t = 0; t < M; ++t) f[0][t] = (r0 + t < T.my) ? x[0][t] / tget(fm, t, 0) : 0.0;
  bmatvec<M, 3, 1>(ytab<M>(T, YT_A1, 7, lane), lane, f, x);
#pragma unroll
  for (int t = 0; t < M; ++t)
    if (r0 + t == T.my - 1) x[0][t] = 0.0;  // gauge row x(my) = 0
  lsolve_tab<M, 3, 1>(ytab<M>(T, YT_LOI, 3, lane), ytab<M>(T, YT_HLI, 3, lane), T.kLi, lane, x);
  usolve_tab<M, 3, 1>(ytab<M>(T, YT_UPI, 3, lane), ytab<M>(T, YT_INVI, 1, lane),
                      ytab<M>(T, YT_HUI, 3, lane),
                      T.kUi, lane, x);
  const double v0 = col_value<M>(x[0], 0);
#pragma unroll
  for (int t = 0; t < M; ++t) x[0][t] -= v0;
}

template <int M>
__global__ __launch_bounds__(64) void k_inty8(YTab T, const double* du, double* u) {
  const int lane = threadIdx.x & 63, r0 = lane * M;
  double x[1][M];
#pragma unroll
  for (int t = 0; t < M; ++t) x[0][t] = (r0 + t < T.my) ? du[r0 + t] : 0.0;
  op_inty<M>(T, lane, x);
#pragma unroll
  for (int t = 0; t < M; ++t)
    if (r0 + t < T.my) u[r0 + t] = x[0][t];
}

// ------------------------------------------------------------------------
// fused phase kernels
// ------------------------------------------------------------------------
__device__ __forceinline__ void mode_of(const RunConst& rc, int w, int& i, int& k,
                                        double& kx, double& kz) {
  i = w / rc.mz;
  k = w - i * rc.mz;
  kx = rc.alp * i;
  kz = (k <= rc.nz) ? rc.bet * k : -rc.bet * (rc.mz - k);
}
// Wavefront sequence index -> mode index i*mz + k.  Within a kx plane the
// order is k = 0, then the pairs (q, mz - q): +kz and -kz share a Helmholtz
// class, so the second wavefront of a pair finds the class's LU tables in
// the cache the first one filled.
__device__ __forceinline__ int mode_seq(const RunConst& rc, int seq) {
  const int i = seq / rc.mz, s = seq - i * rc.mz;
  const int k = s == 0 ? 0 : ((s & 1) ? (s + 1) >> 1 : rc.mz - (s >> 1));
  return i * rc.mz + k;
}
__device__ __forceinline__ int helm_class(const RunConst& rc, int i, int k) {
  const int ka = (k <= rc.nz) ? k : rc.mz - k;
  return i * (rc.nz + 1) + ka;
}

// y = x * (i c)
template <int M>
__device__ __forceinline__ void mul_i(const double (&x)[2][M], double c, double (&y)[2][M]) {
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const double re = -x[1][t] * c, im = x[0][t] * c;
    y[0][t] = re;
    y[1][t] = im;
  }
}

// stage <- x + s * stage   (own rows), the xi updates
template <int M>
__device__ __forceinline__ void stage_axpy(double2* st, int lane, const double (&x)[2][M],
                                           double s) {
  const int r0 = lane * M;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const double2 h = st[r0 + t];
    st[r0 + t] = make_double2(x[0][t] + h.x * s, x[1][t] + h.y * s);
  }
}
// stage <- stage + s * x   (own rows), the gamma updates
template <int M>
__device__ __forceinline__ void stage_add(double2* st, int lane, const double (&x)[2][M],
                                          double s) {
  const int r0 = lane * M;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const double2 h = st[r0 + t];
    st[r0 + t] = make_double2(h.x + s * x[0][t], h.y + s * x[1][t]);
  }
}

// Mean split of the transforms: the 00 mode of a transform input travels in
// FP64 in `mean` (its real part; the imaginary part of a mean is zero) and
// the transform buffer carries zero in that column, so FP32 transforms only
// ever see perturbation-sized values.  The x-line kernel adds the profiles
// back in physical space.
template <int M, typename T2>
__device__ __forceinline__ void store_spec(T2* __restrict__ col, int my, int lane, double2* st,
                                           const double (&x)[2][M], bool m00,
                                           double* __restrict__ mean) {
  if (!m00) {
    store_col<M>(col, my, lane, st, x);
    return;
  }
  const int r0 = lane * M;
  wave_sync();
#pragma unroll
  for (int t = 0; t < M; ++t) {
    if (r0 + t < my) mean[r0 + t] = x[0][t];
    st[r0 + t] = make_double2(0.0, 0.0);
  }
  stage_out<M>(st, my, lane, col);
}
// ... and the outputs' 00 mode is the transformed remainder plus the FP64
// estimate the x-line kernel subtracted (own rows of the stage)
template <int M>
__device__ __forceinline__ void stage_add_mean(double2* st, int lane, int my, bool m00,
                                               const double* __restrict__ mean) {
  if (!m00) return;
  const int r0 = lane * M;
#pragma unroll
  for (int t = 0; t < M; ++t)
    if (r0 + t < my) st[r0 + t].x += mean[r0 + t];
}
template <int M>
__device__ __forceinline__ void add_mean(double (&x)[2][M], int lane, int my, bool m00,
                                         const double* __restrict__ mean) {
  if (!m00) return;
  const int r0 = lane * M;
#pragma unroll
  for (int t = 0; t < M; ++t)
    if (r0 + t < my) x[0][t] += mean[r0 + t];
}

// scalar-field block of phase A: BCs, then f, i kx f, df/dy, i kz f, lap f
template <int M, typename T2>
__device__ __forceinline__ void scalar_derivs(const PhaseArgs& a, int lane, double2* st,
                                              size_t o, size_t os, bool m00, double kx, double kz,
                                              double rK, double2* fld, double bot, double top,
                                              T2* const* sp, int q_f, int q_dx, int q_dy,
                                              int q_dz, int q_lap) {
  const int my = a.rc.my, r0 = lane * M;
  double* mi = a.mean_in;
  double F[2][M], d[2][M];
  load_col<M>(fld + o, my, lane, st, F);
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const int r = r0 + t;
    if (r == 0 || r == my - 1) {
      F[0][t] = m00 ? (r == 0 ? bot : top) : 0.0;
      F[1][t] = 0.0;
      fld[o + r] = make_double2(F[0][t], 0.0);
    }
  }
  if (q_f >= 0) store_spec<M>(sp[q_f] + os, my, lane, st, F, m00, mi + (size_t)q_f * my);
  mul_i<M>(F, kx, d);
  store_spec<M>(sp[q_dx] + os, my, lane, st, d, m00, mi + (size_t)q_dx * my);
  mul_i<M>(F, kz, d);
  store_spec<M>(sp[q_dz] + os, my, lane, st, d, m00, mi + (size_t)q_dz * my);
  op_dy1<M, 2>(a.T, lane, F, d);
  store_spec<M>(sp[q_dy] + os, my, lane, st, d, m00, mi + (size_t)q_dy * my);
  op_dy2<M, 2>(a.T, lane, F, d);
#pragma unroll
  for (int t = 0; t < M; ++t) {
    d[0][t] = d[0][t] - rK * F[0][t];
    d[1][t] = d[1][t] - rK * F[1][t];
  }
  store_spec<M>(sp[q_lap] + os, my, lane, st, d, m00, mi + (size_t)q_lap * my);
}

template <int M, typename TS>
__global__ __launch_bounds__(256, 2) void k_phaseA(PhaseArgs a) {
  const int bx = blockIdx.x;
#include "loma_phaseA_body.inc"
}
template <int M, typename TS>
__global__ __launch_bounds__(256, 2) void k_phaseA_batch(const PhaseArgs* __restrict__ args,
    const int* __restrict__ first, int nrun) {
  const int r = wave_uniform(batch_run(first, nrun, blockIdx.x));
  const PhaseArgs& a = args[r];
  const int bx = blockIdx.x - first[r];
#include "loma_phaseA_body.inc"
}

template <int M, typename TS>
__global__ __launch_bounds__(256, 2) void k_phaseC(PhaseArgs a) {
  const int bx = blockIdx.x;
#include "loma_phaseC_body.inc"
}
template <int M, typename TS>
__global__ __launch_bounds__(256, 2) void k_phaseC_batch(const PhaseArgs* __restrict__ args,
    const int* __restrict__ first, int nrun) {
  const int r = wave_uniform(batch_run(first, nrun, blockIdx.x));
  const PhaseArgs& a = args[r];
  const int bx = blockIdx.x - first[r];
#include "loma_phaseC_body.inc"
}

// phase C of the mixed form: RHS in FP32, state updates in FP64
template <int M>
__global__ __launch_bounds__(256, 3) void k_phaseC32(PhaseArgs a) {
  const int bx = blockIdx.x;
#include "loma_phaseC32_body.inc"
}
template <int M>
__global__ __launch_bounds__(256, 3) void k_phaseC32_batch(const PhaseArgs* __restrict__ args,
    const int* __restrict__ first, int nrun) {
  const int r = wave_uniform(batch_run(first, nrun, blockIdx.x));
  const PhaseArgs& a = args[r];
  const int bx = blockIdx.x - first[r];
#include "loma_phaseC32_body.inc"
}

// (8) psi from the low-Mach constraint and the 00 mode of rho*v
// (cross.f:1027-1104).  A kernel of its own: it reads only drho and writes
// psi, so splitting it from the rest of phase C adds no HBM traffic.
// One wavefront per Helmholtz class (kx, |kz|): the +kz and -kz modes of a
// class share its LU and its homogeneous solutions h1, h2, so the wavefront
// solves both at once (four real columns) and reads the class tables once
// for two modes.  The Neumann combination is v + c1 h1 + c2 h2 with h1, h2
// from the class table (the second Dirichlet solve of solvepsi, laps.f:
// 700-726, by linearity).  The class with kx = kz = 0 is the 00 mode alone.
template <int M>
__device__ __forceinline__ void bd_dots4(const YTab& T, int lane, const double (&x)[4][M],
                                         double (&out)[8]) {
  const TabRef w = ytab<M>(T, YT_WBD, 2, lane);
  double s[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) s[q] = 0.0;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const double a = tget(w, t, 0), b = tget(w, t, 1);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      s[n] += a * x[n][t];
      s[4 + n] += b * x[n][t];
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) out[q] = wave_sum(s[q]);
}

template <int M, typename TS>
__global__ __launch_bounds__(256) void k_phaseC2(PhaseArgs a) {
  const int bx = blockIdx.x;
#include "loma_phaseC2_body.inc"
}
template <int M, typename TS>
__global__ __launch_bounds__(256) void k_phaseC2_batch(const PhaseArgs* __restrict__ args,
    const int* __restrict__ first, int nrun) {
  const int r = wave_uniform(batch_run(first, nrun, blockIdx.x));
  const PhaseArgs& a = args[r];
  const int bx = blockIdx.x - first[r];
#include "loma_phaseC2_body.inc"
}

// Seeded mixing-layer initial condition on the device, one wavefront per
// (kx, kz) mode (the recipe of trunk/initloma3/src/main.f:95-463 with the
// seeded draws of the host: coef[c][mode][n] = amp_n exp(2 pi i phs_n)).
//   u_c(y) = shape(kh) ywin(y) sum_n coef_n e1^(n+1),  e1 = exp(i 2 kh/nymax y)
//   (cu, cv, cw) = curl(u, v, w), zero on the kx = 0 plane and kz = 0 modes;
//   pass 1: 2 |c_.|^2 at the centre row per mode (the host sums them in
//   mode order into the centre-plane rms of each component);
//   pass 2: rho u_i = rho scale_c c_i; psi from lap psi = div(rho u)
//   (Lapvdvhom, main.f:452-463); Omega_y = ikz rho u - ikx rho w;
//   phi = (D2 - k^2)(rho v - d psi/dy); T and Z in the 00 mode only.
template <int M>
__global__ __launch_bounds__(256) void k_ic(IcArgs a) {
  LOMA_STAGE(M);
  const RunConst& rc = a.rc;
  const int nmodes = rc.nplanes * rc.mz;
  const int w = wave_uniform(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (w >= nmodes) return;
  int i, k;
  double kx, kz;
  mode_of(rc, w, i, k, kx, kz);
  const int my = rc.my, r0 = lane * M;
  const size_t o = (size_t)w * my;
  double yv[M], rho[M];
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const int r = r0 + t;
    yv[t] = r < my ? a.y[r] : 0.0;
    rho[t] = 1.0 + (a.s - 1.0) / (a.s + 1.0) * tanh(-2.0 * yv[t]);
  }
  if (i == 0 || k == 0) {
    // no perturbation on the kx = 0 plane and the kz = 0 modes
    if (a.pass == 1) {
      if (lane == 0) a.partial[(size_t)w * 3] = a.partial[(size_t)w * 3 + 1] =
          a.partial[(size_t)w * 3 + 2] = 0.0;
      return;
    }
    double x[2][M];
#pragma unroll
    for (int t = 0; t < M; ++t) x[0][t] = x[1][t] = 0.0;
    store_col<M>(a.vor + o, my, lane, st, x);
    store_col<M>(a.phi + o, my, lane, st, x);
    store_col<M>(a.psi + o, my, lane, st, x);
    const bool m00 = i == 0 && k == 0;
#pragma unroll
    for (int t = 0; t < M; ++t) x[0][t] = m00 ? 1.0 / rho[t] : 0.0;  // T = 1/rho00
    store_col<M>(a.scal + o, my, lane, st, x);
    if (a.with_z) {
#pragma unroll
      for (int t = 0; t < M; ++t) x[0][t] = m00 ? 0.5 * (1.0 + tanh(2.0 * yv[t])) : 0.0;
      store_col<M>(a.zf + o, my, lane, st, x);
    }
    return;
  }
  const double kh = sqrt(kx * kx + kz * kz);
  const double q = kh / a.k0;
  const double shape = q * q * exp(-(q * q));
  const double ky1 = 2.0 * kh / a.nymax;
  double er[M], ei[M], win[M];
#pragma unroll
  for (int t = 0; t < M; ++t) {
    sincos(ky1 * yv[t], &ei[t], &er[t]);
    const double yw = yv[t] / a.env_width;
    win[t] = shape * exp(-(yw * yw));
  }
  // component c of the velocity into f[n0], f[n0+1] (Horner in e1)
  auto build = [&](int c, double (&f)[4][M], int n0) {
    const double2* cf = a.coef + ((size_t)c * nmodes + w) * a.nymax;
    const double2 top = cf[a.nymax - 1];
#pragma unroll
    for (int t = 0; t < M; ++t) {
      f[n0][t] = top.x * er[t] - top.y * ei[t];
      f[n0 + 1][t] = top.x * ei[t] + top.y * er[t];
    }
    for (int n = a.nymax - 2; n >= 0; --n) {
      const double2 cn = cf[n];
#pragma unroll
      for (int t = 0; t < M; ++t) {
        const double pr = f[n0][t] + cn.x, pi = f[n0 + 1][t] + cn.y;
        f[n0][t] = pr * er[t] - pi * ei[t];
        f[n0 + 1][t] = pr * ei[t] + pi * er[t];
      }
    }
#pragma unroll
    for (int t = 0; t < M; ++t) {
      f[n0][t] *= win[t];
      f[n0 + 1][t] *= win[t];
    }
  };
  // u in rows 0,1 and w in rows 2,3; one derivative pass for both
  double uw[4][M], duw[4][M];
  build(0, uw, 0);
  build(2, uw, 2);
  op_dy1<M, 4>(a.T, lane, uw, duw);
  double cu[2][M], cv[2][M], cw[2][M];
  {
    double v[4][M];
    build(1, v, 0);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      // cu = dw - ikz v, cv = ikz u - ikx w, cw = ikx v - du   (main.f:240-257)
      cu[0][t] = duw[2][t] + kz * v[1][t];
      cu[1][t] = duw[3][t] - kz * v[0][t];
      cv[0][t] = -kz * uw[1][t] + kx * uw[3][t];
      cv[1][t] = kz * uw[0][t] - kx * uw[2][t];
      cw[0][t] = -kx * v[1][t] - duw[0][t];
      cw[1][t] = kx * v[0][t] - duw[1][t];
    }
  }
  const int jc = my / 2;
  if (a.pass == 1) {
    // kx > 0 planes count twice (the kx < 0 half is their conjugate)
    auto energy = [&](const double (&x)[2][M]) {
      const double re = col_value<M>(x[0], jc), im = col_value<M>(x[1], jc);
      return 2.0 * (re * re + im * im);
    };
    const double e[3] = {energy(cu), energy(cv), energy(cw)};
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) a.partial[(size_t)w * 3 + c] = e[c];
    }
    return;
  }
  // rho u_i with the normalisation of each component
#pragma unroll
  for (int t = 0; t < M; ++t)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      cu[h][t] = rho[t] * (cu[h][t] * a.scale[0]);
      cv[h][t] = rho[t] * (cv[h][t] * a.scale[1]);
      cw[h][t] = rho[t] * (cw[h][t] * a.scale[2]);
    }
  {
    double vo[2][M];
#pragma unroll
    for (int t = 0; t < M; ++t) {  // Omega_y = ikz rho u - ikx rho w
      vo[0][t] = -kz * cu[1][t] - (-kx * cw[1][t]);
      vo[1][t] = kz * cu[0][t] - kx * cw[0][t];
    }
    store_col<M>(a.vor + o, my, lane, st, vo);
  }
  double x[2][M];
  {
    double dv[2][M], dvv[2][M];
    op_dy1<M, 2>(a.T, lane, cv, dv);
#pragma unroll
    for (int t = 0; t < M; ++t) {  // div = ikx rho u + d(rho v)/dy + ikz rho w
      dvv[0][t] = (-kx * cu[1][t] + dv[0][t]) + (-kz * cw[1][t]);
      dvv[1][t] = (kx * cu[0][t] + dv[1][t]) + kz * cw[0][t];
    }
    const double* lu = a.helm_lu + (size_t)helm_class(rc, i, k) * a.lu_stride;
    op_helm<M, 2>(a.T, lu, lane, dvv, x);  // psi
  }
  store_col<M>(a.psi + o, my, lane, st, x);
  {
    double d[2][M];
    op_dy1<M, 2>(a.T, lane, x, d);
#pragma unroll
    for (int t = 0; t < M; ++t) {  // m_y = rho v - d psi/dy
      x[0][t] = cv[0][t] - d[0][t];
      x[1][t] = cv[1][t] - d[1][t];
    }
    op_dy2<M, 2>(a.T, lane, x, d);
    const double rK = kx * kx + kz * kz;
#pragma unroll
    for (int t = 0; t < M; ++t) {
      d[0][t] = d[0][t] - rK * x[0][t];
      d[1][t] = d[1][t] - rK * x[1][t];
    }
    store_col<M>(a.phi + o, my, lane, st, d);
  }
#pragma unroll
  for (int t = 0; t < M; ++t) x[0][t] = x[1][t] = 0.0;
  store_col<M>(a.scal + o, my, lane, st, x);
  if (a.with_z) store_col<M>(a.zf + o, my, lane, st, x);
}

// Momentum and vorticity thickness from the x-z mean velocity u00(y): the
// 00 column of the transformed u (spec slot 0) plus its FP64 mean-split
// profile, as the reference takes them from the 00 mode of foup2's output.
// which = 1 (rkstep 1): delta_m = Ly/4 - sum_j trp_j (u_{j+1}^2 + u_{j-1}^2)
// / (u_top - u_bot)^2 with trp = hy/2 (cross.f:1414-1429, main.f:689-696);
// which = 3 (rkstep 3): delta_w = 1 / max |d u00/dy| (cross.f:692-699).
// One wavefront; the result goes to the run's device scalars.
template <int M, typename TS>
__global__ __launch_bounds__(64) void k_thickness(PhaseArgs a, const double* hy, double ly,
                                                  int which) {
  using T2 = typename std::conditional<sizeof(TS) == 4, float2, double2>::type;
  LOMA_STAGE(M);
  const int lane = threadIdx.x & 63, my = a.rc.my, r0 = lane * M;
  double u[2][M];
  load_col<M>((const T2*)a.spec[0], my, lane, st, u);
  add_mean<M>(u, lane, my, true, a.mean_out);
  if (which == 1) {
    wave_sync();
    stage_put<M>(st, lane, u);
    wave_sync();
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < M; ++t) {
      const int r = r0 + t;
      if (r >= my) continue;
      const int lo = r == 0 ? 0 : r - 1, hi = r == my - 1 ? my - 1 : r + 1;
      const double ul = st[lo].x, uh = st[hi].x;
      s += 0.5 * hy[r] * (uh * uh + ul * ul);
    }
    s = wave_sum(s);
    const double du = st[my - 1].x - st[0].x;
    if (lane == 0) a.sc->dm = 0.25 * ly - s / (du * du);
  } else {
    double d[2][M];
    op_dy1<M, 2>(a.T, lane, u, d);
    double m = 0.0;
#pragma unroll
    for (int t = 0; t < M; ++t)
      if (r0 + t < my) m = fmax(m, fabs(d[0][t]));
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) m = fmax(m, __shfl_xor(m, dd));
    if (lane == 0) a.sc->dw = 1.0 / m;
  }
}

// per-M launchers (instantiated in loma_yops_m*.hip, one chunk size each)
template <int M>
int launch_dy_m(int order, const YTab& T, const double* u, double* du, int ncols,
                hipStream_t st) {
  dim3 g((ncols + 3) / 4), b(256);
  if (order == 1)
    hipLaunchKernelGGL((k_dy<M, 1>), g, b, 0, st, T, (const double2*)u, (double2*)du, ncols);
  else
    hipLaunchKernelGGL((k_dy<M, 2>), g, b, 0, st, T, (const double2*)u, (double2*)du, ncols);
  return (int)hipGetLastError();
}
template <int M>
int launch_lapvdvhom_m(const YTab& T, const double* lu, int lu_stride, const double* bc,
                       const double* phi, double* v, double* dv, const double* rK, int ncols,
                       int mode, hipStream_t st) {
  dim3 g((ncols + 3) / 4), b(256);
  hipLaunchKernelGGL((k_lapvdvhom<M>), g, b, 0, st, T, lu, lu_stride, bc,
                     (const double2*)phi, (double2*)v, (double2*)dv, rK, ncols, mode);
  return (int)hipGetLastError();
}
template <int M>
int launch_visc_m(const YTab& T, const double* lu1, const double* lu2, int lu_stride,
                  const double* bc2, double* vorwk, double* vor, double* phi, double* v,
                  double* dvdy, double* phiwk, const double* rk2, double dalbe, double dtri,
                  double dtxi, int ncols, hipStream_t st) {
  dim3 g((ncols + 3) / 4), b(256);
  hipLaunchKernelGGL((k_visc<M>), g, b, 0, st, T, lu1, lu2, lu_stride, bc2, (double2*)vorwk,
                     (double2*)vor, (double2*)phi, (double2*)v, (double2*)dvdy,
                     (double2*)phiwk, rk2, dalbe, dtri, dtxi, ncols);
  return (int)hipGetLastError();
}
template <int M>
int launch_helm_bc_m(const YTab& T, const double* lu, int lu_stride, double* bc, double* h,
                     int nclass, hipStream_t st) {
  dim3 g((nclass + 3) / 4), b(256);
  hipLaunchKernelGGL((k_helm_bc<M>), g, b, 0, st, T, lu, lu_stride, bc, h, nclass);
  return (int)hipGetLastError();
}
template <int M>
int launch_inty8_m(const YTab& T, const double* du, double* u, hipStream_t st) {
  hipLaunchKernelGGL((k_inty8<M>), dim3(1), dim3(64), 0, st, T, du, u);
  return (int)hipGetLastError();
}
template <int M>
int launch_phase_m(int which, bool fp32, const PhaseArgs& a, hipStream_t st) {
  dim3 g((a.nmodes + 3) / 4), b(256);
  if (which == 0) {
    if (fp32) hipLaunchKernelGGL((k_phaseA<M, float>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_phaseA<M, double>), g, b, 0, st, a);
  } else if (which == 1) {
    if (fp32) hipLaunchKernelGGL((k_phaseC32<M>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_phaseC<M, double>), g, b, 0, st, a);
  } else {
    // one wavefront per Helmholtz class (the +-kz pair)
    dim3 gc((a.rc.nplanes * (a.rc.nz + 1) + 3) / 4);
    if (fp32) hipLaunchKernelGGL((k_phaseC2<M, float>), gc, b, 0, st, a);
    else hipLaunchKernelGGL((k_phaseC2<M, double>), gc, b, 0, st, a);
  }
  return (int)hipGetLastError();
}

template <int M>
int launch_phase_batch_m(int which, bool fp32, const PhaseArgs* args, const int* first,
                         int nrun, int nblocks, hipStream_t st) {
  dim3 g(nblocks), b(256);
  if (which == 0) {
    if (fp32) hipLaunchKernelGGL((k_phaseA_batch<M, float>), g, b, 0, st, args, first, nrun);
    else hipLaunchKernelGGL((k_phaseA_batch<M, double>), g, b, 0, st, args, first, nrun);
  } else if (which == 1) {
    if (fp32) hipLaunchKernelGGL((k_phaseC32_batch<M>), g, b, 0, st, args, first, nrun);
    else hipLaunchKernelGGL((k_phaseC_batch<M, double>), g, b, 0, st, args, first, nrun);
  } else {
    if (fp32) hipLaunchKernelGGL((k_phaseC2_batch<M, float>), g, b, 0, st, args, first, nrun);
    else hipLaunchKernelGGL((k_phaseC2_batch<M, double>), g, b, 0, st, args, first, nrun);
  }
  return (int)hipGetLastError();
}

template <int M>
int launch_thickness_m(bool fp32, const PhaseArgs& a, const double* hy, double ly, int which,
                       hipStream_t st) {
  if (fp32) hipLaunchKernelGGL((k_thickness<M, float>), dim3(1), dim3(64), 0, st, a, hy, ly, which);
  else hipLaunchKernelGGL((k_thickness<M, double>), dim3(1), dim3(64), 0, st, a, hy, ly, which);
  return (int)hipGetLastError();
}
template <int M>
int launch_ic_m(const IcArgs& a, hipStream_t st) {
  const int nmodes = a.rc.nplanes * a.rc.mz;
  hipLaunchKernelGGL((k_ic<M>), dim3((nmodes + 3) / 4), dim3(256), 0, st, a);
  return (int)hipGetLastError();
}

}  // namespace loma
#include "loma_stats_impl.h"
namespace loma {

#define LOMA_INSTANTIATE_M(MV)                                                        \
  template int launch_dy_m<MV>(int, const YTab&, const double*, double*, int, hipStream_t); \
  template int launch_lapvdvhom_m<MV>(const YTab&, const double*, int, const double*,     \
                                      const double*, double*, double*, const double*, int, \
                                      int, hipStream_t);                                  \
  template int launch_helm_bc_m<MV>(const YTab&, const double*, int, double*, double*,    \
                                    int, hipStream_t);                                    \
  template int launch_inty8_m<MV>(const YTab&, const double*, double*, hipStream_t);      \
  template int launch_visc_m<MV>(const YTab&, const double*, const double*, int,          \
                                 const double*, double*, double*, double*, double*,       \
                                 double*, double*, const double*, double, double, double, \
                                 int, hipStream_t);                                       \
  template int launch_phase_m<MV>(int, bool, const PhaseArgs&, hipStream_t);              \
  template int launch_phase_batch_m<MV>(int, bool, const PhaseArgs*, const int*, int, int,  \
                                        hipStream_t);                                     \
  template int launch_ic_m<MV>(const IcArgs&, hipStream_t);                              \
  template int launch_thickness_m<MV>(bool, const PhaseArgs&, const double*, double, int,  \
                                      hipStream_t);                                       \
  template int launch_stats_m<MV>(bool, int, const StatsArgs&, hipStream_t);

}  // namespace loma
