// loma_stats_impl.h -- the statistics window on the device: the y-profile
// sums and spectra of addstats1 (trunk/lomacte3/src/cross.f:1981-2063) and
// addstats2 (cross.f:2070-2287), accumulated at rkstep 1 of the steps of a
// window.
//
// One wavefront per chunk of a kx plane: kach consecutive |kz| classes, each
// the mode pair kz = +-ka (ka = 0 alone), walked in order.  Every spectrum
// entry (kx plane, |kz|, row) therefore has one writer, and every per-y sum
// one summation order: the results do not depend on scheduling.  The per-y
// sums of a step go to a per-chunk partial; k_stats_reduce (loma_yops.hip)
// adds the chunks and planes in order.  The means of the 00 mode are added
// by the wavefront that holds it (plane 0, chunk 0).
// Included by loma_yops_impl.h (it needs the y operators and staging).
#pragma once

namespace loma {

__device__ __forceinline__ double* spe_at(const StatsArgs& a, int i, int q, int n, int ka) {
  return a.spe + (((size_t)i * 12 + q) * a.nsp + n) * (a.rc.nz + 1) + ka;
}

// Lane n of the wavefront gets the value of a column (held M rows per lane)
// at the n-th spectra row of the block n0.. (zero for lanes past the rows):
// M shuffles, one per row slot, each lane keeping the one its row sits in.
template <int M>
__device__ __forceinline__ double2 spectra_row(const StatsArgs& a, const double (&x)[2][M],
                                               int lane, int n0) {
  const int n = n0 + lane;
  const int row = n < a.nsp ? a.jsp[n] : 0;
  const int l = row / M, t = row - l * M;
  double re = 0.0, im = 0.0;
#pragma unroll
  for (int q = 0; q < M; ++q) {
    const double vr = __shfl(x[0][q], l), vi = __shfl(x[1][q], l);
    if (q == t) { re = vr; im = vi; }
  }
  return n < a.nsp ? make_double2(re, im) : make_double2(0.0, 0.0);
}

// addstats1: rho u, rho v, rho w and T as they enter the transforms (spec
// slots 0, 1, 2 and 7 after phase A, the 00 mode's FP64 profile included),
// cross.f:578-586.  Spectra 9..12, T'^2 per y (into the chunk partial), the
// means of rho u_i and T, and the phys2 sums that need no physical field
// (sigma = 0: mum = sum T^0, num = sum T over the plane, cross.f:1604-1606).
template <int M, typename TS>
__global__ __launch_bounds__(256) void k_stats_in(StatsArgs a) {
  using T2 = typename std::conditional<sizeof(TS) == 4, float2, double2>::type;
  LOMA_STAGE(M);
  const RunConst& rc = a.rc;
  const int wv = wave_uniform(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (wv >= rc.nplanes * a.nch) return;
  const int i = wv / a.nch, ch = wv - i * a.nch;
  const int ka0 = ch * a.kach, ka1 = min(rc.nz, ka0 + a.kach - 1);
  const int my = rc.my, r0 = lane * M;
  const double f = i == 0 ? 1.0 : 2.0;  // kx > 0 planes stand for their conjugates
  const double npts = (double)rc.mgalx * (double)rc.mgalz;
  double tp[M];
#pragma unroll
  for (int t = 0; t < M; ++t) tp[t] = 0.0;
  for (int km = 2 * ka0; km <= 2 * ka1 + 1; ++km) {
    const int ka = km >> 1;
    if ((km & 1) && ka == 0) continue;  // kz = 0 has no pair
    const int k = (km & 1) ? rc.mz - ka : ka;
    const int w = i * rc.mz + k;
    const bool m00 = (i == 0 && k == 0);
#pragma unroll 1
    for (int c = 0; c < 4; ++c) {
      const int slot = c < 3 ? c : 7;
      double x[2][M];
      load_col<M>((const T2*)a.spec[slot] + (size_t)w * rc.myp, my, lane, st, x);
      add_mean<M>(x, lane, my, m00, a.mean_in + (size_t)slot * my);
      for (int n0 = 0; n0 < a.nsp; n0 += WAVE) {
        const double2 v = spectra_row<M>(a, x, lane, n0);
        if (n0 + lane < a.nsp) *spe_at(a, i, 8 + c, n0 + lane, ka) += f * (v.x * v.x + v.y * v.y);
      }
      if (c == 3) {
#pragma unroll
        for (int t = 0; t < M; ++t) tp[t] += f * (x[0][t] * x[0][t] + x[1][t] * x[1][t]);
      }
      if (m00) {
#pragma unroll
        for (int t = 0; t < M; ++t) {
          const int r = r0 + t;
          if (r < my) {
            a.acc[(size_t)(3 + c) * my + r] += x[0][t];
            if (c == 3) {
              a.acc[(size_t)31 * my + r] += npts;
              a.acc[(size_t)32 * my + r] += npts * x[0][t];
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < M; ++t)
    if (r0 + t < my) a.part[((size_t)0 * my + r0 + t) * a.nw + wv] = tp[t];
}

// addstats2: u, v, w as foup2 returns them (spec slots 0, 1, 2 after the z
// transforms, the 00 mode's FP64 estimate added back), their y derivatives
// and the vorticities of cross.f:748-767 (o1 = dw/dy - ikz v, o2 = ikz u -
// ikx w, o3 = ikx v - du/dy).  The twelve per-y sums of a mode go to LDS
// (one column each, row t*64 + lane of the lane's own rows), the chunk's
// totals to its partial.  The wavefront of plane 0, chunk 0 also adds the
// 00 means and the phys2 sums of rho u_i u_j, which are the 00 mode of the
// transformed products times mgalx*mgalz (cross.f:1586-1601).
template <int M, typename TS>
__global__ __launch_bounds__(64) void k_stats_out(StatsArgs a) {
  using T2 = typename std::conditional<sizeof(TS) == 4, float2, double2>::type;
  __shared__ double2 stage_[WAVE * M];
  __shared__ double sacc[12][WAVE * M];
  double2* st = stage_;
  const RunConst& rc = a.rc;
  const int wv = blockIdx.x;
  const int i = wv / a.nch, ch = wv - i * a.nch;
  const int ka0 = ch * a.kach, ka1 = min(rc.nz, ka0 + a.kach - 1);
  const int lane = threadIdx.x;
  const int my = rc.my, r0 = lane * M;
  const double f = i == 0 ? 1.0 : 2.0;
  const double kx = rc.alp * i;
#pragma unroll
  for (int q = 0; q < 12; ++q)
#pragma unroll
    for (int t = 0; t < M; ++t) sacc[q][t * WAVE + lane] = 0.0;
  // spectrum q of this mode at the spectra rows, from a column held in
  // registers: lane n updates row n's entry (one writer per entry)
  auto spec_add = [&](int q, int ka, const double (&x)[2][M]) __attribute__((always_inline)) {
    for (int n0 = 0; n0 < a.nsp; n0 += WAVE) {
      const double2 v = spectra_row<M>(a, x, lane, n0);
      if (n0 + lane < a.nsp) *spe_at(a, i, q, n0 + lane, ka) += f * (v.x * v.x + v.y * v.y);
    }
  };
  auto mean_add = [&](int c, const double (&x)[2][M]) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < M; ++t)
      if (r0 + t < my) a.acc[(size_t)c * my + r0 + t] += x[0][t];
  };
  for (int km = 2 * ka0; km <= 2 * ka1 + 1; ++km) {
    const int ka = km >> 1;
    if ((km & 1) && ka == 0) continue;  // kz = 0 has no pair
    const int k = (km & 1) ? rc.mz - ka : ka;
    const int w = i * rc.mz + k;
    const bool m00 = (i == 0 && k == 0);
    const double kz = (k <= rc.nz) ? rc.bet * k : -rc.bet * (rc.mz - k);
    const double rK = kx * kx + kz * kz;
    double u[2][M], v[2][M], ww[2][M];
    load_col<M>((const T2*)a.spec[0] + (size_t)w * rc.myp, my, lane, st, u);
    add_mean<M>(u, lane, my, m00, a.mean_out);
    wave_sync();
    load_col<M>((const T2*)a.spec[1] + (size_t)w * rc.myp, my, lane, st, v);
    add_mean<M>(v, lane, my, m00, a.mean_out + (size_t)my);
    wave_sync();
    load_col<M>((const T2*)a.spec[2] + (size_t)w * rc.myp, my, lane, st, ww);
    add_mean<M>(ww, lane, my, m00, a.mean_out + 2 * (size_t)my);
    {
      // intensities, shear terms, o2 and the k^2 part of the dissipation
      double o2[2][M];
#pragma unroll
      for (int t = 0; t < M; ++t) {
        const double ur = u[0][t], ui = u[1][t], vr = v[0][t], vi = v[1][t];
        const double wr = ww[0][t], wi = ww[1][t];
        const double eu = ur * ur + ui * ui, ev = vr * vr + vi * vi, ew = wr * wr + wi * wi;
        const int e = t * WAVE + lane;
        sacc[0][e] += f * eu;
        sacc[1][e] += f * ev;
        sacc[2][e] += f * ew;
        sacc[3][e] += f * (ur * vr + ui * vi);  // Re(u conj v)
        sacc[4][e] += f * (ur * wr + ui * wi);  // Re(u conj w)
        sacc[5][e] += f * (wr * vr + wi * vi);  // Re(w conj v)
        o2[0][t] = -kz * ui + kx * wi;
        o2[1][t] = kz * ur - kx * wr;
        sacc[7][e] += f * (o2[0][t] * o2[0][t] + o2[1][t] * o2[1][t]);
        sacc[11][e] += f * (rK * (eu + ev + ew));
      }
      spec_add(0, ka, u);
      spec_add(1, ka, v);
      spec_add(2, ka, ww);
      for (int n0 = 0; n0 < a.nsp; n0 += WAVE) {  // Re and Im of u conj v
        const double2 uu = spectra_row<M>(a, u, lane, n0), vv = spectra_row<M>(a, v, lane, n0);
        if (n0 + lane < a.nsp) {
          *spe_at(a, i, 3, n0 + lane, ka) += f * (uu.x * vv.x + uu.y * vv.y);
          *spe_at(a, i, 4, n0 + lane, ka) += f * (uu.y * vv.x - uu.x * vv.y);
        }
      }
      spec_add(6, ka, o2);
      if (m00) {
        mean_add(0, u);
        mean_add(1, v);
        mean_add(2, ww);
        mean_add(8, o2);
      }
    }
    double d[2][M], c[2][M];
    // o3 = ikx v - du/dy; then u <- ikx u + ikz w (the divergence without dv/dy)
    op_dy1<M, 2>(a.T, lane, u, d);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      c[0][t] = -kx * v[1][t] - d[0][t];
      c[1][t] = kx * v[0][t] - d[1][t];
      const double pr = c[0][t] + kx * v[1][t], pi = c[1][t] - kx * v[0][t];  // o3 - ikx v
      const int e = t * WAVE + lane;
      sacc[8][e] += f * (c[0][t] * c[0][t] + c[1][t] * c[1][t]);
      sacc[11][e] += f * (pr * pr + pi * pi);
      const double xr = -kx * u[1][t] - kz * ww[1][t], xi = kx * u[0][t] + kz * ww[0][t];
      u[0][t] = xr;
      u[1][t] = xi;
    }
    spec_add(7, ka, c);
    if (m00) mean_add(9, c);
    // theta = ikx u + dv/dy + ikz w; thep, theup (cross.f:2167-2171, 2237-2242)
    op_dy1<M, 2>(a.T, lane, v, d);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      const double cr = u[0][t] + d[0][t], ci = u[1][t] + d[1][t];
      const int e = t * WAVE + lane;
      sacc[9][e] += f * (cr * cr + ci * ci);
      sacc[10][e] += f * (cr * v[0][t] + ci * v[1][t]);
      sacc[11][e] += f * (d[0][t] * d[0][t] + d[1][t] * d[1][t]);
    }
    if (m00) mean_add(10, d);
    // o1 = dw/dy - ikz v
    op_dy1<M, 2>(a.T, lane, ww, d);
#pragma unroll
    for (int t = 0; t < M; ++t) {
      c[0][t] = d[0][t] + kz * v[1][t];
      c[1][t] = d[1][t] - kz * v[0][t];
      const double pr = c[0][t] - kz * v[1][t], pi = c[1][t] + kz * v[0][t];  // o1 + ikz v
      const int e = t * WAVE + lane;
      sacc[6][e] += f * (c[0][t] * c[0][t] + c[1][t] * c[1][t]);
      sacc[11][e] += f * (pr * pr + pi * pi);
    }
    spec_add(5, ka, c);
    if (m00) mean_add(7, c);
    wave_sync();
  }
#pragma unroll
  for (int q = 0; q < 12; ++q)
#pragma unroll
    for (int t = 0; t < M; ++t)
      if (r0 + t < my)
        a.part[((size_t)(1 + q) * my + r0 + t) * a.nw + wv] = sacc[q][t * WAVE + lane];
  if (wv == 0) {
    // sum over the plane of rho u_i u_j = mgalx*mgalz times the 00 mode of
    // the transformed product (slot -> column: 3 ruu, 6 ruv, 7 ruw, 4 rvv,
    // 8 rvw, 5 rww)
    const double npts = (double)rc.mgalx * (double)rc.mgalz;
    const int slot[6] = {3, 6, 7, 4, 8, 5};
#pragma unroll 1
    for (int q = 0; q < 6; ++q) {
      double x[2][M];
      wave_sync();
      load_col<M>((const T2*)a.spec[slot[q]], my, lane, st, x);
      add_mean<M>(x, lane, my, true, a.mean_out + (size_t)slot[q] * my);
#pragma unroll
      for (int t = 0; t < M; ++t)
        if (r0 + t < my) a.acc[(size_t)(22 + q) * my + r0 + t] += npts * x[0][t];
    }
  }
}

template <int M>
int launch_stats_m(bool fp32, int which, const StatsArgs& a, hipStream_t st) {
  if (which == 1) {
    dim3 g((a.rc.nplanes * a.nch + 3) / 4), b(256);
    if (fp32) hipLaunchKernelGGL((k_stats_in<M, float>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_stats_in<M, double>), g, b, 0, st, a);
  } else {
    dim3 g(a.rc.nplanes * a.nch), b(64);
    if (fp32) hipLaunchKernelGGL((k_stats_out<M, float>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_stats_out<M, double>), g, b, 0, st, a);
  }
  return (int)hipGetLastError();
}

}  // namespace loma
