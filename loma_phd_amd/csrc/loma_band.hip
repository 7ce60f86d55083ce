// loma_band.hip -- batched pentadiagonal solves in isolation: LU without
// pivoting (bandec, trunk/lomacte3/src/cfdiff.f:977-1022) and the two
// triangular sweeps (banbks, cfdiff.f:954-974) for many independent systems
// with complex right-hand sides, FP64 arithmetic.
//
// One thread owns one system and runs the elimination row by row; a
// 64-thread workgroup serves 64 consecutive systems.  The inputs are
// system-major ([sys][n][5] bands, [sys][n] complex right-hand sides), so
// each workgroup stages RT rows of its 64 systems through LDS: the HBM side
// reads 64 runs of RT*5 (RT*2) contiguous doubles, the thread side reads its
// own rows.  The factors and the forward-substituted right-hand side go to a
// scratch area laid out [row][field][system] (system fastest, coalesced) for
// the backward sweep, which stages its output the same way.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "loma_internal.h"
#include "loma_kernels.h"

namespace loma {

constexpr int B5_SYS = 64;  // systems per workgroup
constexpr int B5_RT = 8;    // rows per LDS tile

__global__ __launch_bounds__(B5_SYS) void k_band5(const double* __restrict__ band,
                                                  const void* __restrict__ rhs, int rhs_fp32,
                                                  double* __restrict__ x,
                                                  double* __restrict__ scr, int n, int nsys) {
  __shared__ double tb[B5_SYS][B5_RT * 5 + 1];
  __shared__ double tr[B5_SYS][B5_RT * 2 + 1];
  const int lane = threadIdx.x;
  const int s0 = blockIdx.x * B5_SYS;
  const int s = s0 + lane;
  const int ns = min(B5_SYS, nsys - s0);
  const size_t ss = (size_t)nsys;  // scratch stride per field
  // cooperative tile load of rows [r0, r0 + RT) of the workgroup's systems
  auto load_tile = [&](int r0) {
    const int rows = min(B5_RT, n - r0);
    for (int it = lane; it < ns * B5_RT * 5; it += B5_SYS) {
      const int q = it / (B5_RT * 5), e = it - q * (B5_RT * 5);
      tb[q][e] = (e < rows * 5) ? band[((size_t)(s0 + q) * n + r0) * 5 + e] : 0.0;
    }
    for (int it = lane; it < ns * B5_RT * 2; it += B5_SYS) {
      const int q = it / (B5_RT * 2), e = it - q * (B5_RT * 2);
      double v = 0.0;
      if (e < rows * 2) {
        const size_t idx = ((size_t)(s0 + q) * n + r0) * 2 + e;
        v = rhs_fp32 ? (double)((const float*)rhs)[idx] : ((const double*)rhs)[idx];
      }
      tr[q][e] = v;
    }
    __syncthreads();
  };
  // window of three rows of the working matrix (5 entries, column j-2+d)
  double w[3][5], yr[3], yi[3];
  auto fetch = [&](int j, double* row, double& br, double& bi) {
    const int t = j % B5_RT;
#pragma unroll
    for (int d = 0; d < 5; ++d) row[d] = tb[lane][t * 5 + d];
    br = tr[lane][t * 2];
    bi = tr[lane][t * 2 + 1];
  };
  load_tile(0);
  for (int q = 0; q < 3; ++q) {
    if (q < n) fetch(q, w[q], yr[q], yi[q]);
    else {
      for (int d = 0; d < 5; ++d) w[q][d] = 0.0;
      yr[q] = yi[q] = 0.0;
    }
  }
  // forward elimination: row k eliminates column k from rows k+1, k+2
  for (int k = 0; k < n; ++k) {
    const double iv = 1.0 / w[0][2];
    for (int i = 1; i <= 2; ++i) {
      if (k + i >= n) break;
      const double l = w[i][2 - i] * iv;
      w[i][2 - i] = 0.0;
      for (int d = 1; d <= 2; ++d) w[i][2 - i + d] -= l * w[0][2 + d];
      yr[i] -= l * yr[0];
      yi[i] -= l * yi[0];
    }
    if (s < nsys) {
      double* o = scr + (size_t)k * 5 * ss + s;
      o[0] = iv;
      o[ss] = (k + 1 < n) ? w[0][3] : 0.0;
      o[2 * ss] = (k + 2 < n) ? w[0][4] : 0.0;
      o[3 * ss] = yr[0];
      o[4 * ss] = yi[0];
    }
    // shift the window; row k+3 enters
    for (int d = 0; d < 5; ++d) {
      w[0][d] = w[1][d];
      w[1][d] = w[2][d];
    }
    yr[0] = yr[1]; yi[0] = yi[1];
    yr[1] = yr[2]; yi[1] = yi[2];
    const int jn = k + 3;
    if (jn < n) {
      if (jn % B5_RT == 0) {
        __syncthreads();
        load_tile(jn);
      }
      fetch(jn, w[2], yr[2], yi[2]);
    } else {
      for (int d = 0; d < 5; ++d) w[2][d] = 0.0;
      yr[2] = yi[2] = 0.0;
    }
  }
  // backward sweep x_k = (y_k - U1 x_{k+1} - U2 x_{k+2}) / pivot, staged
  // through the right-hand-side tile for coalesced stores
  double x1r = 0.0, x1i = 0.0, x2r = 0.0, x2i = 0.0;
  __syncthreads();
  const int ntile = (n + B5_RT - 1) / B5_RT;
  for (int tl = ntile - 1; tl >= 0; --tl) {
    const int r0 = tl * B5_RT;
    const int rows = min(B5_RT, n - r0);
    for (int t = rows - 1; t >= 0; --t) {
      const int k = r0 + t;
      double xr = 0.0, xi = 0.0;
      if (s < nsys) {
        const double* o = scr + (size_t)k * 5 * ss + s;
        const double iv = o[0], u1 = o[ss], u2 = o[2 * ss];
        xr = (o[3 * ss] - u1 * x1r - u2 * x2r) * iv;
        xi = (o[4 * ss] - u1 * x1i - u2 * x2i) * iv;
      }
      tr[lane][t * 2] = xr;
      tr[lane][t * 2 + 1] = xi;
      x2r = x1r; x2i = x1i;
      x1r = xr; x1i = xi;
    }
    __syncthreads();
    for (int it = lane; it < ns * rows * 2; it += B5_SYS) {
      const int q = it / (rows * 2), e = it - q * (rows * 2);
      x[((size_t)(s0 + q) * n + r0) * 2 + e] = tr[q][e];
    }
    __syncthreads();
  }
}

int launch_band5(const double* band, const void* rhs, double* x, double* scratch, int n,
                 int nsys, int rhs_fp32, hipStream_t st) {
  if (nsys <= 0 || n <= 0) return 0;
  hipLaunchKernelGGL(k_band5, dim3((nsys + B5_SYS - 1) / B5_SYS), dim3(B5_SYS), 0, st, band,
                     rhs, rhs_fp32, x, scratch, n, nsys);
  return (int)hipGetLastError();
}

// ---- the allocation-free variant -------------------------------------------
// Scratch of loma_band5_solve_dev_noalloc: a fixed device array that every
// GPU the library runs on holds until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in
// symbol); it uses the device address that hipGetSymbolAddress returns for
// the current device, kept per device in g_scr_dev.
constexpr size_t BAND5_SCRATCH_DOUBLES = (size_t)1 << 22;  // 4194304 doubles, 32 MiB
static __device__ __align__(256) double scratch[BAND5_SCRATCH_DOUBLES];

static std::mutex g_scr_mu;
static std::vector<double*> g_scr_dev;  // device address of scratch, per device

static int scratch_address(int dev, double** out) {
  std::lock_guard<std::mutex> lk(g_scr_mu);
  if ((int)g_scr_dev.size() <= dev) g_scr_dev.resize(dev + 1, nullptr);
  if (!g_scr_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      set_error("loma_band5_solve_dev_noalloc: no device address for the scratch array");
      return LOMA_ERR_DEVICE;
    }
    g_scr_dev[dev] = (double*)p;
  }
  *out = g_scr_dev[dev];
  return LOMA_OK;
}

}  // namespace loma

extern "C" int loma_band5_solve_dev_noalloc(const double* band, void* rhs, double* x, int n,
                                            int nsys, int rhs_fp32, void* stream) {
  using namespace loma;
  if (!band || !rhs || n < 3 || nsys < 0 || (rhs_fp32 && !x)) {
    set_error("loma_band5_solve_dev_noalloc: bad arguments (n >= 3; x required for float rhs)");
    return LOMA_ERR_ARG;
  }
  if ((size_t)n * 5 * (size_t)nsys > BAND5_SCRATCH_DOUBLES) {
    set_error("loma_band5_solve_dev_noalloc: n*5*nsys exceeds the fixed scratch of "
              "4194304 doubles (32 MiB); split the batch");
    return LOMA_ERR_ARG;
  }
  if (nsys == 0) return LOMA_OK;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("loma_band5_solve_dev_noalloc: hipGetDevice failed");
    return LOMA_ERR_DEVICE;
  }
  double* scr = nullptr;
  int e = scratch_address(dev, &scr);
  if (e) return e;
  hipStream_t st = (hipStream_t)stream;
  // every scratch entry the kernel reads it wrote first in the same launch,
  // so nothing is cleared
  e = launch_band5(band, rhs, x ? x : (double*)rhs, scr, n, nsys, rhs_fp32, st);
  if (e) {
    set_error("loma_band5_solve_dev_noalloc: launch failed");
    return LOMA_ERR_DEVICE;
  }
  if (hipStreamSynchronize(st) != hipSuccess) {
    set_error("loma_band5_solve_dev_noalloc: solve failed");
    return LOMA_ERR_DEVICE;
  }
  return LOMA_OK;
}

