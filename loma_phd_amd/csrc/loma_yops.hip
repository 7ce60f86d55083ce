// loma_yops.hip -- dispatch of the y-direction kernels over the chunk size M
// (rows of a y-column per lane); the kernels are in loma_yops_impl.h and are
// instantiated one M per translation unit (loma_yops_m<M>.hip).
#include <hip/hip_runtime.h>

#include "loma_kernels.h"

namespace loma {

template <int M>
int launch_dy_m(int, const YTab&, const double*, double*, int, hipStream_t);
template <int M>
int launch_lapvdvhom_m(const YTab&, const double*, int, const double*, const double*,
                       double*, double*, const double*, int, int, hipStream_t);
template <int M>
int launch_helm_bc_m(const YTab&, const double*, int, double*, double*, int, hipStream_t);
template <int M>
int launch_inty8_m(const YTab&, const double*, double*, hipStream_t);
template <int M>
int launch_visc_m(const YTab&, const double*, const double*, int, const double*, double*,
                  double*, double*, double*, double*, double*, const double*, double, double,
                  double, int, hipStream_t);
template <int M>
int launch_phase_m(int, bool, const PhaseArgs&, hipStream_t);
template <int M>
int launch_ic_m(const IcArgs&, hipStream_t);
template <int M>
int launch_thickness_m(bool, const PhaseArgs&, const double*, double, int, hipStream_t);
template <int M>
int launch_stats_m(bool, int, const StatsArgs&, hipStream_t);
template <int M>
int launch_phase_batch_m(int, bool, const PhaseArgs*, const int*, int, int, hipStream_t);

RkCoef rk_coefficients() {
  RkCoef r;
  const double p0 = 5.0 / 66.0, p1 = -1.0;
  const double g[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
  const double xi[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
  const double b3 = (p1 * (g[0] * g[2] + g[1] * g[2] + xi[1] * g[2] + g[0] * xi[2]) -
                     g[0] * g[1] * g[2] - g[0] * g[1] * xi[2]) /
                    (p1 - g[0] * g[1]);
  for (int q = 0; q < 3; ++q) { r.gama[q] = g[q]; r.xi[q] = xi[q]; }
  r.beta[0] = g[0] - p0; r.beta[1] = p1; r.beta[2] = b3;
  r.alpha[0] = p0; r.alpha[1] = g[1] + xi[1] - p1; r.alpha[2] = g[2] + xi[2] - r.beta[2];
  return r;
}


#define LOMA_DISPATCH_M(M, ...)                 \
  switch (M) {                                  \
    case 3: { constexpr int MM = 3; return __VA_ARGS__; } \
    case 4: { constexpr int MM = 4; return __VA_ARGS__; } \
    case 5: { constexpr int MM = 5; return __VA_ARGS__; } \
    case 6: { constexpr int MM = 6; return __VA_ARGS__; } \
    case 7: { constexpr int MM = 7; return __VA_ARGS__; } \
    case 8: { constexpr int MM = 8; return __VA_ARGS__; } \
    case 9: { constexpr int MM = 9; return __VA_ARGS__; } \
    case 10: { constexpr int MM = 10; return __VA_ARGS__; } \
    case 12: { constexpr int MM = 12; return __VA_ARGS__; } \
    case 14: { constexpr int MM = 14; return __VA_ARGS__; } \
    case 16: { constexpr int MM = 16; return __VA_ARGS__; } \
    default: return (int)hipErrorInvalidValue;  \
  }

int launch_dy(int M, int order, const YTab& T, const double* u, double* du, int ncols,
              hipStream_t st) {
  if (ncols <= 0) return 0;
  LOMA_DISPATCH_M(M, launch_dy_m<MM>(order, T, u, du, ncols, st));
}
int launch_lapvdvhom(int M, const YTab& T, const double* lu, const double* bc,
                     const double* phi, double* v, double* dv, const double* rK, int ncols,
                     int mode, hipStream_t st) {
  if (ncols <= 0) return 0;
  LOMA_DISPATCH_M(M, launch_lapvdvhom_m<MM>(T, lu, 5 * WAVE * MM, bc, phi, v, dv, rK, ncols,
                                            mode, st));
}
int launch_helm_bc(int M, const YTab& T, const double* lu, double* bc, double* h, int nclass,
                   hipStream_t st) {
  if (nclass <= 0) return 0;
  LOMA_DISPATCH_M(M, launch_helm_bc_m<MM>(T, lu, 5 * WAVE * MM, bc, h, nclass, st));
}
int launch_inty8(int M, const YTab& T, const double* du, double* u, hipStream_t st) {
  LOMA_DISPATCH_M(M, launch_inty8_m<MM>(T, du, u, st));
}
int launch_visc(int M, const YTab& T, const double* lu1, const double* lu2, const double* bc2,
                double* vorwk, double* vor, double* phi, double* v, double* dvdy,
                double* phiwk, const double* rk2, double dalbe, double dtri, double dtxi,
                int ncols, hipStream_t st) {
  if (ncols <= 0) return 0;
  LOMA_DISPATCH_M(M, launch_visc_m<MM>(T, lu1, lu2, 5 * WAVE * MM, bc2, vorwk, vor, phi, v,
                                       dvdy, phiwk, rk2, dalbe, dtri, dtxi, ncols, st));
}
int launch_phaseA(int M, bool fp32, const PhaseArgs& a, hipStream_t st) {
  if (a.nmodes <= 0) return 0;
  LOMA_DISPATCH_M(M, launch_phase_m<MM>(0, fp32, a, st));
}
int launch_phaseC(int M, bool fp32, const PhaseArgs& a, hipStream_t st) {
  if (a.nmodes <= 0) return 0;
  LOMA_DISPATCH_M(M, launch_phase_m<MM>(1, fp32, a, st));
}
int launch_phaseC2(int M, bool fp32, const PhaseArgs& a, hipStream_t st) {
  if (a.nmodes <= 0) return 0;
  LOMA_DISPATCH_M(M, launch_phase_m<MM>(2, fp32, a, st));
}
int launch_phase_batch(int M, bool fp32, int which, const PhaseArgs* args, const int* first,
                       int nrun, int nblocks, hipStream_t st) {
  if (nblocks <= 0 || nrun <= 0) return 0;
  LOMA_DISPATCH_M(M, launch_phase_batch_m<MM>(which, fp32, args, first, nrun, nblocks, st));
}
// the grids of launch_phase_m: 4 wavefronts (modes, or Helmholtz classes
// for phase C2) per 256-thread block
int phase_blocks(int which, const RunConst& rc) {
  const int n = which == 2 ? rc.nplanes * (rc.nz + 1) : rc.nplanes * rc.mz;
  return (n + 3) / 4;
}
int launch_ic(int M, const IcArgs& a, hipStream_t st) {
  LOMA_DISPATCH_M(M, launch_ic_m<MM>(a, st));
}
int launch_thickness(int M, bool fp32, const PhaseArgs& a, const double* hy, double ly,
                     int which, hipStream_t st) {
  LOMA_DISPATCH_M(M, launch_thickness_m<MM>(fp32, a, hy, ly, which, st));
}

// LU of (dt22 - s dt21) with identity end rows (Lapvdvhom laps.f:349-366),
// one thread per shift s.  Output per class: a lane-major table of 64*M rows
// with entries inv, up1, up2, lo1, lo2 (lower multipliers by destination
// row), zero beyond row my-1.
__global__ void k_helm_factor(const double* __restrict__ dt21,
                              const double* __restrict__ dt22, int my, int M,
                              const double* __restrict__ shifts, int nsys,
                              double* __restrict__ lu) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  const double sh = shifts[s];
  const int stride = 5 * WAVE * M;
  double* out = lu + (size_t)s * stride;
  for (int q = 0; q < stride; ++q) out[q] = 0.0;
  auto at = [&](int row, int idx) -> double& {
    const int l = row / M, t = row - l * M;
    return out[(t * 5 + idx) * WAVE + l];
  };
  auto row = [&](int j, double* r) {
    if (j >= my) { for (int d = 0; d < 5; ++d) r[d] = 0.0; return; }
    if (j == 0 || j == my - 1) {
      for (int d = 0; d < 5; ++d) r[d] = 0.0;
      r[2] = 1.0;
      return;
    }
    for (int d = 0; d < 5; ++d) r[d] = dt22[j * 5 + d] - sh * dt21[j * 5 + d];
  };
  double w[3][5];
  row(0, w[0]); row(1, w[1]); row(2, w[2]);
  for (int k = 0; k < my; ++k) {
    const double iv = 1.0 / w[0][2];
    at(k, 0) = iv;
    at(k, 1) = (k + 1 < my) ? w[0][3] : 0.0;
    at(k, 2) = (k + 2 < my) ? w[0][4] : 0.0;
    for (int i = 1; i <= 2; ++i) {
      if (k + i >= my) continue;
      const double l = w[i][2 - i] * iv;
      at(k + i, 2 + i) = l;  // multiplier of row k into row k+i
      w[i][2 - i] = 0.0;
      for (int d = 1; d <= 2; ++d) w[i][2 - i + d] -= l * w[0][2 + d];
    }
    for (int d = 0; d < 5; ++d) { w[0][d] = w[1][d]; w[1][d] = w[2][d]; }
    row(k + 3, w[2]);
  }
}

int launch_helm_factor(const double* dt21, const double* dt22, int my, int M,
                       const double* shifts, int nsys, double* lu, hipStream_t st) {
  if (nsys <= 0) return 0;
  hipLaunchKernelGGL(k_helm_factor, dim3((nsys + 63) / 64), dim3(64), 0, st, dt21, dt22,
                     my, M, shifts, nsys, lu);
  return (int)hipGetLastError();
}

// end of step: time += dt; dt <- CFL estimate (cross.f:1196-1204,
// 1769-1818) from the maxima of the x-line tiles, reduced here by one
// workgroup.  NaN flag from the CFL maxima.
constexpr int FINISH_THREADS = 1024;
__device__ __forceinline__ void finish_body(DevScalars* sc, const RunConst& rc, int icfl,
                                            const double* cflp, int ntile) {
  __shared__ double red[FINISH_THREADS / 64][4];
  double c[4] = {0.0, 0.0, 0.0, 0.0};
  if (icfl) {
    for (int t = threadIdx.x; t < ntile; t += FINISH_THREADS) {
      const double2 p = ((const double2*)cflp)[2 * (size_t)t];
      const double2 q = ((const double2*)cflp)[2 * (size_t)t + 1];
      c[0] = max_nan(c[0], p.x);
      c[1] = max_nan(c[1], p.y);
      c[2] = max_nan(c[2], q.x);
      c[3] = max_nan(c[3], q.y);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      for (int d = 32; d >= 1; d >>= 1) c[k] = max_nan(c[k], __shfl_xor(c[k], d));
    if ((threadIdx.x & 63) == 0)
      for (int k = 0; k < 4; ++k) red[threadIdx.x >> 6][k] = c[k];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int w = 1; w < FINISH_THREADS / 64; ++w)
        for (int k = 0; k < 4; ++k) c[k] = max_nan(c[k], red[w][k]);
  }
  if (threadIdx.x != 0) return;
  sc->time = sc->time + sc->dt;
  if (icfl) {
    const double cx = c[0], cy = c[1], cz = c[2], cd = c[3];
    const double conv = cx * rc.hxalp * 0.5 + cy + cz * rc.hzbet * 0.5;
    double dtr = rc.cfl / (conv + cd);
    if (rc.dt_fp32) dtr = (double)(float)dtr;
    if (!(dtr > 0.0) || isinf(dtr)) sc->nan = 1;
    sc->dtr = dtr;
    sc->dt = dtr;
  }
}
__global__ __launch_bounds__(FINISH_THREADS) void k_finish_step(DevScalars* sc, RunConst rc,
                                                                 int icfl, const double* cflp,
                                                                 int ntile) {
  finish_body(sc, rc, icfl, cflp, ntile);
}
// one workgroup per run
__global__ __launch_bounds__(FINISH_THREADS) void k_finish_batch(const FinishArgs* args) {
  const FinishArgs& f = args[blockIdx.x];
  finish_body(f.sc, f.rc, f.icfl, f.cflp, f.ntile);
}
int launch_finish_batch(const FinishArgs* args, int nrun, hipStream_t st) {
  if (nrun <= 0) return 0;
  hipLaunchKernelGGL(k_finish_batch, dim3(nrun), dim3(FINISH_THREADS), 0, st, args);
  return (int)hipGetLastError();
}


// The step's per-chunk sums into the accumulated columns: one wavefront per
// (y row, quantity) adds the row's chunk partials (contiguous in memory),
// each lane a fixed strided subset in order and then a fixed shuffle tree,
// so the sum does not depend on scheduling.  Quantities: Tp, then up, vp,
// wp, uvr, uwr, vwr, w1p, w2p, w3p, thep, theup, ep; the last index takes
// the line sums of 1/T over z (rhom).  The plane-summed intensities of this
// step are kept for uner (cross.f:802-816).
__global__ __launch_bounds__(64) void k_stats_reduce(StatsArgs a) {
  const int my = a.rc.my;
  const int j = blockIdx.x, q = blockIdx.y, lane = threadIdx.x;
  constexpr int col[STAT_NPART + 1] = {20, 11, 12, 13, 14, 15, 16, 17, 18, 19, 29, 30, 21, 28};
  double s = 0.0;
  if (q < STAT_NPART) {
    const double* p = a.part + ((size_t)q * my + j) * a.nw;
    for (int w = lane; w < a.nw; w += 64) s += p[w];
  } else {
    for (int z = lane; z < a.rc.mgalz; z += 64) s += a.rho[(size_t)z * my + j];
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d);
  if (lane == 0) {
    a.acc[(size_t)col[q] * my + j] += s;
    if (q >= 1 && q <= 9) a.last[(size_t)(q - 1) * my + j] = s;
  }
}

int launch_stats(int M, bool fp32, int which, const StatsArgs& a, hipStream_t st) {
  if (which == 3) {
    hipLaunchKernelGGL(k_stats_reduce, dim3(a.rc.my, STAT_NPART + 1), dim3(64), 0, st, a);
    return (int)hipGetLastError();
  }
  LOMA_DISPATCH_M(M, launch_stats_m<MM>(fp32, which, a, st));
}

int launch_finish_step(DevScalars* sc, RunConst rc, int icfl, const double* cflp, int ntile,
                       hipStream_t st) {
  if (icfl && (!cflp || ntile <= 0 || ntile > cfl_tiles_max(rc.my, rc.mgalz)))
    return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_finish_step, dim3(1), dim3(FINISH_THREADS), 0, st, sc, rc, icfl, cflp,
                     ntile);
  return (int)hipGetLastError();
}

}  // namespace loma
