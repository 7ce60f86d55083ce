// loma_band.hip -- batched pentadiagonal solves in isolation: LU without
// pivoting (bandec, trunk/lomacte3/src/cfdiff.f:977-1022) and the two
// triangular sweeps (banbks, cfdiff.f:954-974) for many independent systems
// with complex right-hand sides, FP64 arithmetic.
//
// A system is solved by two threads, one per wavefront of a 128-thread
// workgroup: both run the elimination row by row on the (real) bands, wave
// 0 on the real part of the right-hand side and wave 1 on the imaginary
// part, so a workgroup serves 64 consecutive systems with two waves.  The
// inputs are system-major ([sys][n][5] bands, [sys][n] complex right-hand
// sides); tiles of RT rows of the workgroup's 64 systems go through LDS,
// the HBM side reading 64 runs of RT*5 (RT*2) contiguous doubles.  The tiles
// are double-buffered: the loads of tile t+1 are issued into registers
// before the rows of tile t are eliminated and written to LDS after, so a
// tile's load latency hides behind the previous tile's rows.  The factors
// (wave 0) and the forward-substituted right-hand sides go to a scratch area
// laid out [row][field][system] (system fastest, coalesced); the backward
// sweep prefetches the scratch rows of the next tile into registers the
// same way and stages its output through LDS for coalesced stores.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "loma_internal.h"
#include "loma_kernels.h"

namespace loma {

constexpr int B5_SYS = 64;                 // systems per workgroup
constexpr int B5_RT = 8;                   // rows per tile
constexpr int B5_THREADS = 2 * B5_SYS;     // real and imaginary wave
constexpr int B5_PB = B5_SYS * B5_RT * 5 / B5_THREADS;  // band doubles per thread per tile
constexpr int B5_PR = B5_SYS * B5_RT * 2 / B5_THREADS;  // rhs values per thread per tile

// F32: the right-hand side is complex FP32 (else FP64).  rhs and x may be
// the same buffer (an FP64 solve in place), so neither is restrict.
template <bool F32>
__global__ __launch_bounds__(B5_THREADS) void k_band5(const double* __restrict__ band,
                                                      const void* rhs, double* x,
                                                      double* __restrict__ scr, int n, int nsys) {
  __shared__ double tb[2][B5_SYS][B5_RT * 5 + 1];
  __shared__ double tr[2][B5_SYS][B5_RT * 2 + 1];
  const int tid = threadIdx.x, lane = tid & 63, part = tid >> 6;
  const int s0 = blockIdx.x * B5_SYS;
  const int s = s0 + lane;
  const int ns = min(B5_SYS, nsys - s0);
  const size_t ss = (size_t)nsys;  // scratch stride per field
  const int ntile = (n + B5_RT - 1) / B5_RT;
  // ---- tile prefetch: global -> registers -> LDS buffer ----
  double pb[B5_PB], pr[B5_PR];
  auto fetch = [&](int tl) {
    const int r0 = tl * B5_RT, rows = min(B5_RT, n - r0);
#pragma unroll
    for (int q = 0; q < B5_PB; ++q) {
      const int it = tid + q * B5_THREADS;
      const int sy = it / (B5_RT * 5), e = it - sy * (B5_RT * 5);
      pb[q] = (sy < ns && e < rows * 5) ? band[((size_t)(s0 + sy) * n + r0) * 5 + e] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < B5_PR; ++q) {
      const int it = tid + q * B5_THREADS;
      const int sy = it / (B5_RT * 2), e = it - sy * (B5_RT * 2);
      double v = 0.0;
      if (sy < ns && e < rows * 2) {
        const size_t idx = ((size_t)(s0 + sy) * n + r0) * 2 + e;
        if constexpr (F32) v = (double)((const float*)rhs)[idx];
        else v = ((const double*)rhs)[idx];
      }
      pr[q] = v;
    }
  };
  auto commit = [&](int buf) {
#pragma unroll
    for (int q = 0; q < B5_PB; ++q) {
      const int it = tid + q * B5_THREADS;
      const int sy = it / (B5_RT * 5), e = it - sy * (B5_RT * 5);
      tb[buf][sy][e] = pb[q];
    }
#pragma unroll
    for (int q = 0; q < B5_PR; ++q) {
      const int it = tid + q * B5_THREADS;
      const int sy = it / (B5_RT * 2), e = it - sy * (B5_RT * 2);
      tr[buf][sy][e] = pr[q];
    }
  };
  // row j of the working matrix (5 entries, column j-2+d) and this wave's
  // part of its right-hand side, from the tile buffer holding row j
  auto row_at = [&](int j, double* w, double& y) {
    const int buf = (j / B5_RT) & 1, t = j % B5_RT;
#pragma unroll
    for (int d = 0; d < 5; ++d) w[d] = tb[buf][lane][t * 5 + d];
    y = tr[buf][lane][t * 2 + part];
  };
  fetch(0);
  commit(0);
  __syncthreads();
  if (ntile > 1) fetch(1);
  // window of three rows
  double w[3][5], y[3];
  for (int q = 0; q < 3; ++q) {
    if (q < n) row_at(q, w[q], y[q]);
    else {
      for (int d = 0; d < 5; ++d) w[q][d] = 0.0;
      y[q] = 0.0;
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
      y[i] -= l * y[0];
    }
    if (s < nsys) {
      double* o = scr + (size_t)k * 5 * ss + s;
      if (part == 0) {
        o[0] = iv;
        o[ss] = (k + 1 < n) ? w[0][3] : 0.0;
        o[2 * ss] = (k + 2 < n) ? w[0][4] : 0.0;
      }
      o[(3 + part) * ss] = y[0];
    }
    // shift the window; row k+3 enters
    for (int d = 0; d < 5; ++d) {
      w[0][d] = w[1][d];
      w[1][d] = w[2][d];
    }
    y[0] = y[1];
    y[1] = y[2];
    const int jn = k + 3;
    if (jn < n) {
      if (jn % B5_RT == 0) {
        // tile jn/RT was prefetched; its buffer held tile jn/RT - 2, whose
        // rows every thread has left
        const int tl = jn / B5_RT;
        __syncthreads();
        commit(tl & 1);
        __syncthreads();
        if (tl + 1 < ntile) fetch(tl + 1);
      }
      row_at(jn, w[2], y[2]);
    } else {
      for (int d = 0; d < 5; ++d) w[2][d] = 0.0;
      y[2] = 0.0;
    }
  }
  // backward sweep x_k = (y_k - U1 x_{k+1} - U2 x_{k+2}) / pivot, tile by
  // tile from the last; the next tile's scratch rows are loaded into
  // registers before the current tile is solved
  double cur[B5_RT][4], nxt[B5_RT][4];
  auto sload = [&](int tl, double (&v)[B5_RT][4]) {
    const int r0 = tl * B5_RT;
#pragma unroll
    for (int t = 0; t < B5_RT; ++t) {
      const int k = r0 + t;
      if (s < nsys && k < n) {
        const double* o = scr + (size_t)k * 5 * ss + s;
        v[t][0] = o[0];
        v[t][1] = o[ss];
        v[t][2] = o[2 * ss];
        v[t][3] = o[(3 + part) * ss];
      } else {
        v[t][0] = v[t][1] = v[t][2] = v[t][3] = 0.0;
      }
    }
  };
  __syncthreads();  // the scratch rows of this workgroup are written (same threads)
  sload(ntile - 1, cur);
  double x1 = 0.0, x2 = 0.0;
  for (int tl = ntile - 1; tl >= 0; --tl) {
    if (tl > 0) sload(tl - 1, nxt);
    const int r0 = tl * B5_RT;
    const int rows = min(B5_RT, n - r0);
#pragma unroll
    for (int t = B5_RT - 1; t >= 0; --t) {
      if (t >= rows) continue;
      const double xv = (cur[t][3] - cur[t][1] * x1 - cur[t][2] * x2) * cur[t][0];
      tr[0][lane][t * 2 + part] = xv;
      x2 = x1;
      x1 = xv;
    }
    __syncthreads();
    for (int it = tid; it < ns * rows * 2; it += B5_THREADS) {
      const int q = it / (rows * 2), e = it - q * (rows * 2);
      x[((size_t)(s0 + q) * n + r0) * 2 + e] = tr[0][q][e];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < B5_RT; ++t)
#pragma unroll
      for (int c = 0; c < 4; ++c) cur[t][c] = nxt[t][c];
  }
}

int launch_band5(const double* band, const void* rhs, double* x, double* scratch, int n,
                 int nsys, int rhs_fp32, hipStream_t st) {
  if (nsys <= 0 || n <= 0) return 0;
  const dim3 g((nsys + B5_SYS - 1) / B5_SYS), b(B5_THREADS);
  if (rhs_fp32) hipLaunchKernelGGL(k_band5<true>, g, b, 0, st, band, rhs, x, scratch, n, nsys);
  else hipLaunchKernelGGL(k_band5<false>, g, b, 0, st, band, rhs, x, scratch, n, nsys);
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

