// loma_yops.hip -- y-direction kernels: per-operator entry points and the two
// fused per-mode phase kernels of an RK3 substep.
//
// Phase A = trunk/lomacte3/src/cross.f:315-554 (BCs, Lapvdvhom, m_x/m_z,
//   xi-update, rho*u_i, T derivatives) producing the 8 transform inputs.
// Phase C = cross.f:619-1104 (T update, N_i, vorticities, Laplacians,
//   RHS, gamma-update, 00 modes, solvepsi) consuming the 11 transform outputs.
// One 64-lane wavefront per (kx,kz) mode; see loma_dev.h for the solver.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "loma_kernels.h"

namespace loma {

static __constant__ double c_gama[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static __constant__ double c_xi[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};

// ------------------------------------------------------------------------
// column I/O: complex column of length my into re/im register chunks
// ------------------------------------------------------------------------
template <int M, typename T2>
__device__ __forceinline__ void load_col(const T2* __restrict__ col, int my, int r0,
                                         double (&x)[2][M]) {
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const int r = r0 + t;
    if (r < my) {
      T2 v = col[r];
      x[0][t] = (double)v.x;
      x[1][t] = (double)v.y;
    } else {
      x[0][t] = 0.0;
      x[1][t] = 0.0;
    }
  }
}
template <int M, typename T2>
__device__ __forceinline__ void store_col(T2* __restrict__ col, int my, int r0,
                                          const double (&x)[2][M]) {
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const int r = r0 + t;
    if (r < my) {
      T2 v;
      v.x = x[0][t];
      v.y = x[1][t];
      col[r] = v;
    }
  }
}

// ------------------------------------------------------------------------
// per-operator kernels
// ------------------------------------------------------------------------
template <int M, int ORDER>
__global__ __launch_bounds__(256) void k_dy(YTab T, const double2* u, double2* du,
                                            int ncols) {
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (wave >= ncols) return;
  const int my = T.my, r0 = lane * M;
  double x[2][M], y[2][M];
  load_col<M>(u + (size_t)wave * my, my, r0, x);
  if (ORDER == 1)
    op_dy1<M, 2>(T, r0, lane, x, y);
  else
    op_dy2<M, 2>(T, r0, lane, x, y);
  store_col<M>(du + (size_t)wave * my, my, r0, y);
}

// mode 0: Lapvdvhom (laps.f:319); mode 1: solvepsi (laps.f:580);
// mode 2: Dirichlet Helmholtz (visc kernel), no derivative output
template <int M>
__device__ __forceinline__ void solve_psi_col(const YTab& T, const double* lu, int r0,
                                              int lane, const double (&f)[2][M],
                                              double (&v)[2][M], double (&dv)[2][M]) {
  // particular (re, im) and the two homogeneous solutions (laps.f:658-689)
  double x[4][M];
  {
    double rhs[2][M];
    bmatvec<M, 2, 2>(BandRef{T.a2, 5, 0}, r0, lane, f, rhs);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      const int r = r0 + t;
      const bool end = (r == 0 || r == T.my - 1);
      x[0][t] = end ? 0.0 : rhs[0][t];
      x[1][t] = end ? 0.0 : rhs[1][t];
      x[2][t] = (r == 0) ? 1.0 : 0.0;
      x[3][t] = (r == T.my - 1) ? 1.0 : 0.0;
    }
  }
  lsolve<M, 2, 4>(BandRef{lu, 5, 3}, r0, lane, x);
  usolve<M, 2, 4>(BandRef{lu, 5, 1}, BandRef{lu, 5, 0}, r0, lane, x);
  // d/deta of all four (calcrhsd1 + banbks7, laps.f:696-705)
  double d[4][M];
  bmatvec<M, 3, 4>(BandRef{T.b1, 7, 0}, r0, lane, x, d);
  lsolve<M, 3, 4>(BandRef{T.lo1, 3, 0}, r0, lane, d);
  usolve<M, 3, 4>(BandRef{T.up1, 3, 0}, BandRef{T.inv1, 1, 0}, r0, lane, d);
  // d psi/dy = 0 at both ends: 2x2 system (laps.f:707-721)
  const int my = T.my;
  const double f0 = T.fmap[PADR], f1 = T.fmap[PADR + my - 1];
  const double d1b = col_value<M>(d[2], 0), d2b = col_value<M>(d[3], 0);
  const double d1t = col_value<M>(d[2], my - 1), d2t = col_value<M>(d[3], my - 1);
  const double dpr0 = col_value<M>(d[0], 0), dpr1 = col_value<M>(d[0], my - 1);
  const double dpi0 = col_value<M>(d[1], 0), dpi1 = col_value<M>(d[1], my - 1);
  const double a11 = d1b * f0, a12 = d2b * f0, a21 = d1t * f1, a22 = d2t * f1;
  const double det = a11 * a22 - a12 * a21;
  const double rr0 = -dpr0 * f0, rr1 = -dpr1 * f1;
  const double ri0 = -dpi0 * f0, ri1 = -dpi1 * f1;
  const double cr1 = (rr0 * a22 - a12 * rr1) / det, cr2 = (a11 * rr1 - a21 * rr0) / det;
  const double ci1 = (ri0 * a22 - a12 * ri1) / det, ci2 = (a11 * ri1 - a21 * ri0) / det;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const double fm = bget(BandRef{T.fmap, 1, 0}, r0 + t, 0);
    v[0][t] = x[0][t] + cr1 * x[2][t] + cr2 * x[3][t];
    v[1][t] = x[1][t] + ci1 * x[2][t] + ci2 * x[3][t];
    dv[0][t] = fm * (d[0][t] + cr1 * d[2][t] + cr2 * d[3][t]);
    dv[1][t] = fm * (d[1][t] + ci1 * d[2][t] + ci2 * d[3][t]);
  }
}

template <int M>
__global__ __launch_bounds__(256) void k_lapvdvhom(YTab T, const double* lu, int lu_rows,
                                                   const double2* phi, double2* v,
                                                   double2* dv, const double* rK,
                                                   int ncols, int mode) {
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (wave >= ncols) return;
  const int my = T.my, r0 = lane * M;
  double f[2][M], x[2][M], d[2][M];
  load_col<M>(phi + (size_t)wave * my, my, r0, f);
  const double* L = lu + (size_t)wave * lu_rows * 5;
  if (mode != 2 && rK[wave] == 0.0) {
#pragma unroll
    for (int t = 0; t < M; ++t) x[0][t] = x[1][t] = d[0][t] = d[1][t] = 0.0;
  } else if (mode == 1) {
    solve_psi_col<M>(T, L, r0, lane, f, x, d);
  } else {
    op_helm<M, 2>(T, L, r0, lane, f, x);
    if (mode == 0) op_dy1<M, 2>(T, r0, lane, x, d);
  }
  store_col<M>(v + (size_t)wave * my, my, r0, x);
  if (mode != 2) store_col<M>(dv + (size_t)wave * my, my, r0, d);
}

// inty8 (laps.f:484-575): u = B1^{-1} A1 (du/fmap) - value at row 0
template <int M>
__device__ __forceinline__ void op_inty(const YTab& T, int r0, int lane, double (&x)[1][M]) {
  double f[1][M];
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const int r = r0 + t;
    f[0][t] = (r < T.my) ? x[0][t] / bget(BandRef{T.fmap, 1, 0}, r, 0) : 0.0;
  }
  bmatvec<M, 3, 1>(BandRef{T.a1, 7, 0}, r0, lane, f, x);
#pragma unroll
  for (int t = 0; t < M; ++t)
    if (r0 + t == T.my - 1) x[0][t] = 0.0;  // gauge row x(my) = 0
  lsolve<M, 3, 1>(BandRef{T.loi, 3, 0}, r0, lane, x);
  usolve<M, 3, 1>(BandRef{T.upi, 3, 0}, BandRef{T.invi, 1, 0}, r0, lane, x);
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
  op_inty<M>(T, r0, lane, x);
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
__device__ __forceinline__ int helm_class(const RunConst& rc, int i, int k) {
  const int ka = (k <= rc.nz) ? k : rc.mz - k;
  return i * (rc.nz + 1) + ka;
}

// z = (a + i b) * (i c)  ->  (-b c, a c)
template <int M>
__device__ __forceinline__ void mul_i(const double (&x)[2][M], double c, double (&y)[2][M]) {
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const double re = -x[1][t] * c, im = x[0][t] * c;
    y[0][t] = re;
    y[1][t] = im;
  }
}

template <int M, typename TS>
__global__ __launch_bounds__(256) void k_phaseA(PhaseArgs a) {
  using T2 = typename std::conditional<sizeof(TS) == 4, float2, double2>::type;
  const RunConst& rc = a.rc;
  const int w = a.mode_base + blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= a.mode_base + a.nmodes) return;
  int i, k;
  double kx, kz;
  mode_of(rc, w, i, k, kx, kz);
  const int my = rc.my, r0 = lane * M;
  const double rK = kx * kx + kz * kz;
  const double dt = a.sc->dt;
  const double dtxi = dt * c_xi[a.rkstep - 1];
  const size_t o = (size_t)w * my;
  const double* lu = a.helm_lu + (size_t)helm_class(rc, i, k) * a.lu_rows * 5;
  const bool m00 = (i == 0 && k == 0);

  // (1) phi BCs + Lapvdvhom -> m_y (v), dm_y/dy (dv)    cross.f:328-372
  double f[2][M], v[2][M], dv[2][M];
  load_col<M>(a.phi + o, my, r0, f);
#pragma unroll
  for (int t = 0; t < M; ++t)
    if (r0 + t == 0 || r0 + t == my - 1) f[0][t] = f[1][t] = 0.0;
  if (rK == 0.0) {
#pragma unroll
    for (int t = 0; t < M; ++t) v[0][t] = v[1][t] = dv[0][t] = dv[1][t] = 0.0;
  } else {
    op_helm<M, 2>(a.T, lu, r0, lane, f, v);
    op_dy1<M, 2>(a.T, r0, lane, v, dv);
  }
  {  // phi** = phi + dtxi * RHS(phi)^(n-1)     cross.f:420-430
    double h[2][M];
    load_col<M>(a.ten13 + o, my, r0, h);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      f[0][t] = f[0][t] + h[0][t] * dtxi;
      f[1][t] = f[1][t] + h[1][t] * dtxi;
    }
    store_col<M>(a.phi + o, my, r0, f);
  }
  // (2) vor BCs, m_x, m_z                               cross.f:315-320, 393-407
  double om[2][M];
  load_col<M>(a.vor + o, my, r0, om);
  {
    const double b0r = col_value<M>(om[0], 1), b0i = col_value<M>(om[1], 1);
    const double b1r = col_value<M>(om[0], my - 2), b1i = col_value<M>(om[1], my - 2);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      if (r0 + t == 0) { om[0][t] = b0r; om[1][t] = b0i; }
      if (r0 + t == my - 1) { om[0][t] = b1r; om[1][t] = b1i; }
    }
  }
  double mx[2][M], mz[2][M];
  {
    const double ta = (rK != 0.0) ? kx / rK : 0.0, tb = (rK != 0.0) ? kz / rK : 0.0;
    // m_x = dv*(i ta) - om*(i tb); m_z = dv*(i tb) + om*(i ta)
#pragma unroll
    for (int t = 0; t < M; ++t) {
      mx[0][t] = (-dv[1][t] * ta) - (-om[1][t] * tb);
      mx[1][t] = (dv[0][t] * ta) - (om[0][t] * tb);
      mz[0][t] = (-dv[1][t] * tb) + (-om[1][t] * ta);
      mz[1][t] = (dv[0][t] * tb) + (om[0][t] * ta);
    }
  }
  {  // vor** = vor + dtxi * RHS(vor)^(n-1)
    double h[2][M];
    load_col<M>(a.ten12 + o, my, r0, h);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      om[0][t] = om[0][t] + h[0][t] * dtxi;
      om[1][t] = om[1][t] + h[1][t] * dtxi;
    }
    store_col<M>(a.vor + o, my, r0, om);
  }
  // (3) rho u_i = m_i + grad psi                        cross.f:434-482
  {
    double ps[2][M], dps[2][M];
    load_col<M>(a.psi + o, my, r0, ps);
    op_dy1<M, 2>(a.T, r0, lane, ps, dps);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      mx[0][t] = mx[0][t] + (-ps[1][t] * kx);
      mx[1][t] = mx[1][t] + (ps[0][t] * kx);
      mz[0][t] = mz[0][t] + (-ps[1][t] * kz);
      mz[1][t] = mz[1][t] + (ps[0][t] * kz);
      v[0][t] = v[0][t] + dps[0][t];
      v[1][t] = v[1][t] + dps[1][t];
    }
  }
  if (m00) {
    // 00 modes of rho u, rho v, rho w and the xi part of their update
    // (cross.f:486-499)
#pragma unroll
    for (int t = 0; t < M; ++t) {
      const int r = r0 + t;
      if (r < my) {
        mx[0][t] = a.u00wk[r]; mx[1][t] = 0.0;
        v[0][t] = a.v00wk[r];  v[1][t] = 0.0;
        mz[0][t] = a.w00wk[r]; mz[1][t] = 0.0;
        if (r >= 1 && r <= my - 2) {
          a.u00wk[r] = a.u00wk[r] + dtxi * a.rf0u[r];
          a.w00wk[r] = a.w00wk[r] + dtxi * a.rf0w[r];
        }
      }
    }
  }
  store_col<M>((T2*)a.spec[0] + o, my, r0, mx);
  store_col<M>((T2*)a.spec[1] + o, my, r0, v);
  store_col<M>((T2*)a.spec[2] + o, my, r0, mz);
  // (4) T BCs and derivatives                           cross.f:338-349, 521-554
  {
    double T[2][M], d1[2][M], d2[2][M];
    load_col<M>(a.scal + o, my, r0, T);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      const int r = r0 + t;
      if (r == 0 || r == my - 1) {
        T[0][t] = m00 ? (r == 0 ? 1.0 / rc.rhobot : 1.0 / rc.rhotop) : 0.0;
        T[1][t] = 0.0;
        double2 bv;
        bv.x = T[0][t];
        bv.y = 0.0;
        a.scal[o + r] = bv;
      }
    }
    op_dy1<M, 2>(a.T, r0, lane, T, d1);
    op_dy2<M, 2>(a.T, r0, lane, T, d2);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      d2[0][t] = d2[0][t] - rK * T[0][t];
      d2[1][t] = d2[1][t] - rK * T[1][t];
    }
    store_col<M>((T2*)a.spec[3] + o, my, r0, d2);  // lap T
    store_col<M>((T2*)a.spec[5] + o, my, r0, d1);  // dT/dy
    mul_i<M>(T, kx, d1);
    store_col<M>((T2*)a.spec[4] + o, my, r0, d1);  // dT/dx
    mul_i<M>(T, kz, d2);
    store_col<M>((T2*)a.spec[6] + o, my, r0, d2);  // dT/dz
    store_col<M>((T2*)a.spec[7] + o, my, r0, T);   // T
  }
  if (a.with_z) {
    // passive scalar: BCs and derivatives (lomaHZB cross.f:343-355, 626-666)
    double Z[2][M], d1[2][M], d2[2][M];
    load_col<M>(a.zf + o, my, r0, Z);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      const int r = r0 + t;
      if (r == 0 || r == my - 1) {
        Z[0][t] = m00 ? (r == 0 ? rc.zbot : rc.ztop) : 0.0;
        Z[1][t] = 0.0;
        double2 bv;
        bv.x = Z[0][t];
        bv.y = 0.0;
        a.zf[o + r] = bv;
      }
    }
    op_dy1<M, 2>(a.T, r0, lane, Z, d1);
    op_dy2<M, 2>(a.T, r0, lane, Z, d2);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      d2[0][t] = d2[0][t] - rK * Z[0][t];
      d2[1][t] = d2[1][t] - rK * Z[1][t];
    }
    store_col<M>((T2*)a.spec[9] + o, my, r0, d1);   // dZ/dy
    store_col<M>((T2*)a.spec[11] + o, my, r0, d2);  // lap Z
    mul_i<M>(Z, kx, d1);
    store_col<M>((T2*)a.spec[8] + o, my, r0, d1);   // dZ/dx
    mul_i<M>(Z, kz, d2);
    store_col<M>((T2*)a.spec[10] + o, my, r0, d2);  // dZ/dz
  }
}

template <int M, typename TS>
__global__ __launch_bounds__(256) void k_phaseC(PhaseArgs a) {
  using T2 = typename std::conditional<sizeof(TS) == 4, float2, double2>::type;
  const RunConst& rc = a.rc;
  const int w = a.mode_base + blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= a.mode_base + a.nmodes) return;
  int i, k;
  double kx, kz;
  mode_of(rc, w, i, k, kx, kz);
  const int my = rc.my, r0 = lane * M;
  const double rK = kx * kx + kz * kz;
  const double dt = a.sc->dt;
  const double dtgamma = dt * c_gama[a.rkstep - 1];
  const double dtxi = dt * c_xi[a.rkstep - 1];
  const double ire = rc.ire;
  const size_t o = (size_t)w * my;
  const bool m00 = (i == 0 && k == 0);
  // spec slots: 0 u, 1 v, 2 w, 3 ruu, 4 rvv, 5 rww, 6 ruv, 7 ruw, 8 rvw,
  //             9 drho, 10 DT, 11 DZ
  const T2* sp[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) sp[q] = (const T2*)a.spec[q] + o;

  // (6a) T += DT                                        cross.f:619-627
  {
    double T[2][M], D[2][M];
    load_col<M>(a.scal + o, my, r0, T);
    load_col<M>(sp[10], my, r0, D);
#pragma unroll
    for (int t = 0; t < M; ++t) { T[0][t] += D[0][t]; T[1][t] += D[1][t]; }
    store_col<M>(a.scal + o, my, r0, T);
    if (a.with_z) {
      load_col<M>(a.zf + o, my, r0, T);
      load_col<M>(sp[11], my, r0, D);
#pragma unroll
      for (int t = 0; t < M; ++t) { T[0][t] += D[0][t]; T[1][t] += D[1][t]; }
      store_col<M>(a.zf + o, my, r0, T);
    }
  }
  // (6b) N_i                                            cross.f:651-682
  double P1[2][M], Q[2][M];
  {
    double ruv[2][M], rvw[2][M], Ny[2][M];
    {
      double rvv[2][M];
      load_col<M>(sp[4], my, r0, rvv);
      op_dy1<M, 2>(a.T, r0, lane, rvv, Ny);
    }
    load_col<M>(sp[6], my, r0, ruv);
    load_col<M>(sp[8], my, r0, rvw);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      // Ny = -ikx ruv - d(rvv)/dy - ikz rvw
      const double re = (-(-ruv[1][t] * kx) - Ny[0][t]) - (-rvw[1][t] * kz);
      const double im = (-(ruv[0][t] * kx) - Ny[1][t]) - (rvw[0][t] * kz);
      Ny[0][t] = re;
      Ny[1][t] = im;
    }
    double druv[2][M], drvw[2][M];
    op_dy1<M, 2>(a.T, r0, lane, ruv, druv);
    op_dy1<M, 2>(a.T, r0, lane, rvw, drvw);
    double Nx[2][M], Nz[2][M];
    {
      double ruu[2][M], ruw[2][M], rww[2][M];
      load_col<M>(sp[3], my, r0, ruu);
      load_col<M>(sp[7], my, r0, ruw);
      load_col<M>(sp[5], my, r0, rww);
#pragma unroll
      for (int t = 0; t < M; ++t) {
        Nx[0][t] = (-(-ruu[1][t] * kx) - druv[0][t]) - (-ruw[1][t] * kz);
        Nx[1][t] = (-(ruu[0][t] * kx) - druv[1][t]) - (ruw[0][t] * kz);
        Nz[0][t] = (-(-ruw[1][t] * kx) - drvw[0][t]) - (-rww[1][t] * kz);
        Nz[1][t] = (-(ruw[0][t] * kx) - drvw[1][t]) - (rww[0][t] * kz);
      }
    }
    if (m00) {
      // 00-mode RHS (cross.f:686-710): -d(rhouv)/dy + u00''/Re
      double u[2][M], wv[2][M], d2u[2][M], d2w[2][M];
      load_col<M>(sp[0], my, r0, u);
      load_col<M>(sp[2], my, r0, wv);
#pragma unroll
      for (int t = 0; t < M; ++t) { u[1][t] = 0.0; wv[1][t] = 0.0; }
      op_dy2<M, 2>(a.T, r0, lane, u, d2u);
      op_dy2<M, 2>(a.T, r0, lane, wv, d2w);
#pragma unroll
      for (int t = 0; t < M; ++t) {
        const int r = r0 + t;
        if (r < my) {
          const double fu = -druv[0][t] + ire * d2u[0][t];
          const double fw = -drvw[0][t] + ire * d2w[0][t];
          a.rf0u[r] = fu;
          a.rf0w[r] = fw;
          // 00 gamma update (cross.f:1006-1011)
          if (r >= 1 && r <= my - 2) {
            a.u00wk[r] = a.u00wk[r] + dtgamma * fu;
            a.w00wk[r] = a.w00wk[r] + dtgamma * fw;
          }
        }
      }
    }
    // P1 = ikz Nx - ikx Nz ; S = ikx Nx + ikz Nz        cross.f:885-906
    double S[2][M];
#pragma unroll
    for (int t = 0; t < M; ++t) {
      P1[0][t] = (-Nx[1][t] * kz) - (-Nz[1][t] * kx);
      P1[1][t] = (Nx[0][t] * kz) - (Nz[0][t] * kx);
      S[0][t] = (-Nx[1][t] * kx) + (-Nz[1][t] * kz);
      S[1][t] = (Nx[0][t] * kx) + (Nz[0][t] * kz);
    }
    double dS[2][M];
    op_dy1<M, 2>(a.T, r0, lane, S, dS);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      Q[0][t] = -rK * Ny[0][t] - dS[0][t];
      Q[1][t] = -rK * Ny[1][t] - dS[1][t];
    }
  }
  // (6c,6d) vorticities and their Laplacians             cross.f:748-857
  {
    double u[2][M], vv[2][M], wv[2][M], du[2][M], dw[2][M];
    load_col<M>(sp[0], my, r0, u);
    load_col<M>(sp[2], my, r0, wv);
    op_dy1<M, 2>(a.T, r0, lane, u, du);
    op_dy1<M, 2>(a.T, r0, lane, wv, dw);
    load_col<M>(sp[1], my, r0, vv);
    double wx[2][M], wy[2][M], wzv[2][M];
#pragma unroll
    for (int t = 0; t < M; ++t) {
      wx[0][t] = dw[0][t] - (-vv[1][t] * kz);
      wx[1][t] = dw[1][t] - (vv[0][t] * kz);
      wy[0][t] = (-u[1][t] * kz) - (-wv[1][t] * kx);
      wy[1][t] = (u[0][t] * kz) - (wv[0][t] * kx);
      wzv[0][t] = (-vv[1][t] * kx) - du[0][t];
      wzv[1][t] = (vv[0][t] * kx) - du[1][t];
    }
    double lap[2][M];
    op_dy2<M, 2>(a.T, r0, lane, wy, lap);
    double rhs[2][M];
#pragma unroll
    for (int t = 0; t < M; ++t) {
      const double l0 = lap[0][t] - rK * wy[0][t], l1 = lap[1][t] - rK * wy[1][t];
      rhs[0][t] = P1[0][t] + ire * l0;
      rhs[1][t] = P1[1][t] + ire * l1;
    }
    {  // RHS(Omega_y) and gamma update of vor           cross.f:966-972
      double vr[2][M];
      load_col<M>(a.vor + o, my, r0, vr);
#pragma unroll
      for (int t = 0; t < M; ++t) {
        vr[0][t] = vr[0][t] + dtgamma * rhs[0][t];
        vr[1][t] = vr[1][t] + dtgamma * rhs[1][t];
      }
      store_col<M>(a.vor + o, my, r0, vr);
      store_col<M>(a.ten12 + o, my, r0, rhs);
    }
    // (curl omega)_y = ikz om_x - ikx om_z
    double cy[2][M];
#pragma unroll
    for (int t = 0; t < M; ++t) {
      cy[0][t] = (-wx[1][t] * kz) - (-wzv[1][t] * kx);
      cy[1][t] = (wx[0][t] * kz) - (wzv[0][t] * kx);
    }
    op_dy2<M, 2>(a.T, r0, lane, cy, lap);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      const double l0 = lap[0][t] - rK * cy[0][t], l1 = lap[1][t] - rK * cy[1][t];
      rhs[0][t] = Q[0][t] - ire * l0;
      rhs[1][t] = Q[1][t] - ire * l1;
    }
    {
      double ph[2][M];
      load_col<M>(a.phi + o, my, r0, ph);
#pragma unroll
      for (int t = 0; t < M; ++t) {
        ph[0][t] = ph[0][t] + dtgamma * rhs[0][t];
        ph[1][t] = ph[1][t] + dtgamma * rhs[1][t];
      }
      store_col<M>(a.phi + o, my, r0, ph);
      store_col<M>(a.ten13 + o, my, r0, rhs);
    }
  }
  // (8) psi from the constraint                          cross.f:1027-1104
  {
    double dr[2][M], ps[2][M], dps[2][M];
    load_col<M>(sp[9], my, r0, dr);
    if (m00) {
      // rho v 00 mode = integral of drho00 + acoustic split (cross.f:1027-1072)
      double x1[1][M];
#pragma unroll
      for (int t = 0; t < M; ++t) x1[0][t] = dr[0][t];
      op_inty<M>(a.T, r0, lane, x1);
      const double trapz = -col_value<M>(x1[0], my - 1);
      const double add = trapz / (1.0 + sqrt(rc.rhotop / rc.rhobot));
#pragma unroll
      for (int t = 0; t < M; ++t)
        if (r0 + t < my) a.v00wk[r0 + t] = x1[0][t] + add;
    }
    if (rK == 0.0) {
#pragma unroll
      for (int t = 0; t < M; ++t) ps[0][t] = ps[1][t] = 0.0;
    } else {
      const double* lu = a.helm_lu + (size_t)helm_class(rc, i, k) * a.lu_rows * 5;
      solve_psi_col<M>(a.T, lu, r0, lane, dr, ps, dps);
    }
    store_col<M>(a.psi + o, my, r0, ps);
  }
  (void)dtxi;
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
int launch_lapvdvhom_m(const YTab& T, const double* lu, int lu_rows, const double* phi,
                       double* v, double* dv, const double* rK, int ncols, int mode,
                       hipStream_t st) {
  dim3 g((ncols + 3) / 4), b(256);
  hipLaunchKernelGGL((k_lapvdvhom<M>), g, b, 0, st, T, lu, lu_rows, (const double2*)phi,
                     (double2*)v, (double2*)dv, rK, ncols, mode);
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
  } else {
    if (fp32) hipLaunchKernelGGL((k_phaseC<M, float>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_phaseC<M, double>), g, b, 0, st, a);
  }
  return (int)hipGetLastError();
}

#define LOMA_INSTANTIATE_M(MV)                                                        \
  template int launch_dy_m<MV>(int, const YTab&, const double*, double*, int, hipStream_t); \
  template int launch_lapvdvhom_m<MV>(const YTab&, const double*, int, const double*,     \
                                      double*, double*, const double*, int, int,          \
                                      hipStream_t);                                       \
  template int launch_inty8_m<MV>(const YTab&, const double*, double*, hipStream_t);      \
  template int launch_phase_m<MV>(int, bool, const PhaseArgs&, hipStream_t);

}  // namespace loma
