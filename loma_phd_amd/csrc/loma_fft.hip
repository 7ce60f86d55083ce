// loma_fft.hip -- x/z Fourier transforms between spectral planes and
// physical lines, and the fused physical-space kernel.
//
// Reference: foup2 (trunk/lomacte3/src/cross.f:1282-1430), fourz/fourx
// (fou3D.f:29-95), FFTPACK conventions (cftsingle.f:2-66, rftsingle.f:25-96),
// phys2 (cross.f:1440-1828).
//
// Layouts (complex T2 = float2 or double2):
//   spectral S  : [i][k][j]       (kx plane, kz mode, y)   -- y fastest
//   z-physical W: [i][z][j]       (kx plane, z point, y)   -- y fastest
//   rhst/rhsz   : [z][j][x] FP64  (physical history of RHS(T) / RHS(Z))
// Every FFT is a Stockham autosort transform in LDS with mixed radices
// 8/4/2/3/5 (Govindaraju et al., SC'08 formulation), one workgroup per tile.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "loma_kernels.h"

namespace loma {

FftPlan make_plan(int n) {
  FftPlan p;
  p.n = n;
  p.npass = 0;
  int m = n;
  while (m % 8 == 0 && m > 1) { p.rad[p.npass++] = 8; m /= 8; }
  while (m % 4 == 0 && m > 1) { p.rad[p.npass++] = 4; m /= 4; }
  while (m % 2 == 0 && m > 1) { p.rad[p.npass++] = 2; m /= 2; }
  while (m % 3 == 0 && m > 1) { p.rad[p.npass++] = 3; m /= 3; }
  while (m % 5 == 0 && m > 1) { p.rad[p.npass++] = 5; m /= 5; }
  if (m != 1) p.npass = -1;  // unsupported length
  return p;
}

template <typename T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  T2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename T2>
__device__ __forceinline__ T2 cadd(T2 a, T2 b) { T2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T2>
__device__ __forceinline__ T2 csub(T2 a, T2 b) { T2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// multiply by s*i (s = +1 or -1)
template <typename T2>
__device__ __forceinline__ T2 cmuli(T2 a, int s) { T2 r; r.x = -s * a.y; r.y = s * a.x; return r; }

// in-register DFT of size R: V_k = sum_r v_r exp(s 2 pi i r k / R)
template <typename T, typename T2>
__device__ __forceinline__ void dft2(T2* v) {
  T2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <typename T, typename T2>
__device__ __forceinline__ void dft4(T2* v, int s) {
  T2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  T2 t2 = cadd(v[1], v[3]), t3 = cmuli(csub(v[1], v[3]), s);
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}
template <typename T, typename T2>
__device__ __forceinline__ void dft8(T2* v, int s) {
  T2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
  dft4<T, T2>(e, s);
  dft4<T, T2>(o, s);
  const T h = (T)0.70710678118654752440;
  T2 w1, w3;
  w1.x = h; w1.y = s * h;     // exp(s i pi/4)
  w3.x = -h; w3.y = s * h;    // exp(s 3 i pi/4)
  T2 o1 = cmul(o[1], w1), o2 = cmuli(o[2], s), o3 = cmul(o[3], w3);
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}
template <typename T, typename T2>
__device__ __forceinline__ void dft3(T2* v, int s) {
  const T c = (T)-0.5, sn = (T)(0.86602540378443864676) * (T)s;
  T2 a = v[0], b = v[1], d = v[2];
  T2 sum = cadd(b, d), dif = csub(b, d);
  v[0] = cadd(a, sum);
  T2 m;
  m.x = a.x + c * sum.x;
  m.y = a.y + c * sum.y;
  T2 r;
  r.x = -sn * dif.y;
  r.y = sn * dif.x;
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}
template <typename T, typename T2>
__device__ __forceinline__ void dft5(T2* v, int s) {
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212 * (T)s, s2 = (T)0.58778525229247312917 * (T)s;
  T2 a = v[0];
  T2 b1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
  T2 b2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
  v[0].x = a.x + b1.x + b2.x;
  v[0].y = a.y + b1.y + b2.y;
  T2 m1, m2, r1, r2;
  m1.x = a.x + c1 * b1.x + c2 * b2.x; m1.y = a.y + c1 * b1.y + c2 * b2.y;
  m2.x = a.x + c2 * b1.x + c1 * b2.x; m2.y = a.y + c2 * b1.y + c1 * b2.y;
  // r = i*(s1 d1 + s2 d2), i*(s2 d1 - s1 d2)
  r1.x = -(s1 * d1.y + s2 * d2.y); r1.y = s1 * d1.x + s2 * d2.x;
  r2.x = -(s2 * d1.y - s1 * d2.y); r2.y = s2 * d1.x - s1 * d2.x;
  v[1] = cadd(m1, r1);
  v[4] = csub(m1, r1);
  v[2] = cadd(m2, r2);
  v[3] = csub(m2, r2);
}

// LDS layout of FFT work buffers: one pad entry after every 8, so that the
// stride-R butterfly writes of the first passes spread over the banks.
#ifndef LOMA_LDS_PAD
#define LOMA_LDS_PAD 0
#endif
// waves per SIMD the x-line and z-transform kernels are compiled for
#ifndef LOMA_XL_WPE
#define LOMA_XL_WPE 2
#endif
#ifndef LOMA_ZF_WPE
#define LOMA_ZF_WPE 4
#endif
template <int PAD = LOMA_LDS_PAD>
__device__ __forceinline__ int lds_i(int i) { return PAD ? i + (i >> 3) : i; }
__host__ __device__ inline int lds_entries(int n, int pad = LOMA_LDS_PAD) {
  return pad ? n + (n >> 3) + 1 : n;
}
// Row-major LDS rows of a z-transform tile: entry i of row `row` at
// row * rs + lds_i<PAD>(i).  The in-row pad spreads the stride-R butterfly
// writes over the banks; the row stride is odd (in entries) so that the
// tile's global loads and stores, which walk down a column (consecutive
// rows, same i), hit distinct banks.
__host__ __device__ inline int lds_row_stride(int n, int pad) { return n + (pad ? n / 8 : 0) + 1; }
#ifndef LOMA_ZF_PAD32
#define LOMA_ZF_PAD32 1
#endif
template <typename T>
constexpr int zf_pad() { return sizeof(T) == 8 ? 1 : LOMA_ZF_PAD32; }

// Twiddles of one butterfly: w[r] *= tw[jm * r * tstep] (conjugated for
// s > 0), r = 1..R-1.  With LOMA_TW_REC = 1 (default) only tw[jm * tstep]
// is loaded and its powers are formed by products (at most three
// multiplications deep), which trades six table loads per butterfly for
// complex multiplications: 5.75 against 5.97 ms per FP64 x-line launch and
// 1.64 against 1.74 ms per FP64 forward z launch at 512x513x256.
#ifndef LOMA_TW_REC
#define LOMA_TW_REC 1
#endif
// the powers of a loaded first twiddle p1 (LOMA_TW_REC's products)
template <int R, typename T2>
__device__ __forceinline__ void twiddle_pow(T2* w, T2 p1, int s) {
  T2 p[R];
  p[1] = p1;
  if (s > 0) p[1].y = -p[1].y;
  if constexpr (R >= 3) p[2] = cmul(p[1], p[1]);
  if constexpr (R >= 4) p[3] = cmul(p[2], p[1]);
  if constexpr (R >= 5) p[4] = cmul(p[2], p[2]);
  if constexpr (R >= 8) {
    p[5] = cmul(p[4], p[1]);
    p[6] = cmul(p[3], p[3]);
    p[7] = cmul(p[4], p[3]);
  }
#pragma unroll
  for (int r = 1; r < R; ++r) w[r] = cmul(w[r], p[r]);
}
template <int R, typename T2>
__device__ __forceinline__ void twiddle(T2* w, const T2* __restrict__ tw, int jm, int tstep,
                                        int s) {
  if constexpr (LOMA_TW_REC) {
    twiddle_pow<R>(w, tw[jm * tstep], s);
  } else {
#pragma unroll
    for (int r = 1; r < R; ++r) {
      T2 t = tw[jm * r * tstep];
      if (s > 0) t.y = -t.y;
      w[r] = cmul(w[r], t);
    }
  }
}

template <int R, typename T, typename T2, int PAD = LOMA_LDS_PAD>
__device__ __forceinline__ void stockham_item(const T2* __restrict__ src, T2* __restrict__ dst,
                                              int rowbase, int j, int nb, int Ns, int tstep,
                                              const T2* __restrict__ tw, int s) {
  T2 v[R];
  const int jm = j % Ns;
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = src[lds_i<PAD>(rowbase + j + r * nb)];
  if (Ns > 1) twiddle<R>(v, tw, jm, tstep, s);
  if (R == 2) dft2<T, T2>(v);
  if (R == 3) dft3<T, T2>(v, s);
  if (R == 4) dft4<T, T2>(v, s);
  if (R == 5) dft5<T, T2>(v, s);
  if (R == 8) dft8<T, T2>(v, s);
  const int idxD = rowbase + (j / Ns) * Ns * R + jm;
#pragma unroll
  for (int r = 0; r < R; ++r) dst[lds_i<PAD>(idxD + r * Ns)] = v[r];
}

// The same butterfly with its first twiddle p1 = tw[(j % Ns) * tstep]
// already in a register (no global load inside the pass)
template <int R, typename T, typename T2, int PAD = LOMA_LDS_PAD>
__device__ __forceinline__ void stockham_item_p1(const T2* __restrict__ src,
                                                 T2* __restrict__ dst, int rowbase, int j,
                                                 int nb, int Ns, T2 p1, int s) {
  static_assert(LOMA_TW_REC, "register twiddles are the first power only");
  T2 v[R];
  const int jm = j % Ns;
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = src[lds_i<PAD>(rowbase + j + r * nb)];
  if (Ns > 1) twiddle_pow<R>(v, p1, s);
  if (R == 2) dft2<T, T2>(v);
  if (R == 3) dft3<T, T2>(v, s);
  if (R == 4) dft4<T, T2>(v, s);
  if (R == 5) dft5<T, T2>(v, s);
  if (R == 8) dft8<T, T2>(v, s);
  const int idxD = rowbase + (j / Ns) * Ns * R + jm;
#pragma unroll
  for (int r = 0; r < R; ++r) dst[lds_i<PAD>(idxD + r * Ns)] = v[r];
}

// nrows independent length-N transforms of rows stored with stride N in a
// (LDS, padded layout lds_i); b is scratch of the same size.  Returns the
// buffer holding the result.  s = -1 forward (e^{-i}), +1 inverse; unscaled.
// Ends with a barrier.
template <typename T, typename T2, int PAD = LOMA_LDS_PAD>
__device__ T2* fft_rows(T2* a, T2* b, int N, int RS, int nrows, const FftPlan& pl,
                        const T2* __restrict__ tw, int s) {
  int Ns = 1;
  for (int p = 0; p < pl.npass; ++p) {
    const int R = pl.rad[p];
    const int nb = N / R;
    const int tstep = N / (Ns * R);
    const int total = nrows * nb;
    for (int it = threadIdx.x; it < total; it += blockDim.x) {
      const int row = it / nb, j = it - row * nb;
      const int rb = row * RS;
      switch (R) {
        case 2: stockham_item<2, T, T2, PAD>(a, b, rb, j, nb, Ns, tstep, tw, s); break;
        case 3: stockham_item<3, T, T2, PAD>(a, b, rb, j, nb, Ns, tstep, tw, s); break;
        case 4: stockham_item<4, T, T2, PAD>(a, b, rb, j, nb, Ns, tstep, tw, s); break;
        case 5: stockham_item<5, T, T2, PAD>(a, b, rb, j, nb, Ns, tstep, tw, s); break;
        default: stockham_item<8, T, T2, PAD>(a, b, rb, j, nb, Ns, tstep, tw, s); break;
      }
    }
    __syncthreads();
    T2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// In-place variant for one LDS buffer: each pass loads its items into
// registers, synchronises, writes the butterflies back, synchronises.
// nrows * N <= blockDim.x * E (E = complex entries per thread).
template <int R, typename T, typename T2, int PAD>
__device__ __forceinline__ void radix_pass_ip(T2* a, int N, int RS, int nrows, int Ns, int tstep,
                                              const T2* __restrict__ tw, int s, T2* v,
                                              int nitems) {
  const int nb = N / R;
  const int total = nrows * nb;
#pragma unroll
  for (int k = 0; k < nitems; ++k) {
    const int it = threadIdx.x + k * blockDim.x;
    if (it < total) {
      const int row = it / nb, j = it - row * nb;
#pragma unroll
      for (int r = 0; r < R; ++r) v[k * R + r] = a[row * RS + lds_i<PAD>(j + r * nb)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < nitems; ++k) {
    const int it = threadIdx.x + k * blockDim.x;
    if (it < total) {
      const int row = it / nb, j = it - row * nb;
      const int jm = j % Ns;
      T2* w = v + k * R;
      if (Ns > 1) twiddle<R>(w, tw, jm, tstep, s);
      if (R == 2) dft2<T, T2>(w);
      if (R == 3) dft3<T, T2>(w, s);
      if (R == 4) dft4<T, T2>(w, s);
      if (R == 5) dft5<T, T2>(w, s);
      if (R == 8) dft8<T, T2>(w, s);
      const int idxD = (j / Ns) * Ns * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) a[row * RS + lds_i<PAD>(idxD + r * Ns)] = w[r];
    }
  }
  __syncthreads();
}

template <typename T, typename T2, int E, int PAD>
__device__ void fft_rows_ip(T2* a, int N, int nrows, const FftPlan& pl,
                            const T2* __restrict__ tw, int s) {
  T2 v[E + 8];
  const int RS = lds_row_stride(N, PAD);
  int Ns = 1;
  for (int p = 0; p < pl.npass; ++p) {
    const int R = pl.rad[p];
    const int tstep = N / (Ns * R);
    switch (R) {
      case 2: radix_pass_ip<2, T, T2, PAD>(a, N, RS, nrows, Ns, tstep, tw, s, v, (E + 1) / 2); break;
      case 3: radix_pass_ip<3, T, T2, PAD>(a, N, RS, nrows, Ns, tstep, tw, s, v, (E + 2) / 3); break;
      case 4: radix_pass_ip<4, T, T2, PAD>(a, N, RS, nrows, Ns, tstep, tw, s, v, (E + 3) / 4); break;
      case 5: radix_pass_ip<5, T, T2, PAD>(a, N, RS, nrows, Ns, tstep, tw, s, v, (E + 4) / 5); break;
      default: radix_pass_ip<8, T, T2, PAD>(a, N, RS, nrows, Ns, tstep, tw, s, v, (E + 7) / 8); break;
    }
    Ns *= R;
  }
}

// Compile-time plans for the production lengths: the same radix sequence as
// make_plan (8s, then 4, 2, 3, 5), with every stride, index split and
// twiddle step a constant, so the index arithmetic folds into shifts and
// multiplies.  Results are identical to the run-time plan.
constexpr int ct_radix(int m) {
  return m % 8 == 0 ? 8 : (m % 4 == 0 ? 4 : (m % 2 == 0 ? 2 : (m % 3 == 0 ? 3 : 5)));
}

// WR: every pass has 64 items per row (length 512 in radix 8), so the items
// of a row stay in one wavefront from pass to pass; the passes are then
// separated by wavefront barriers and only the last ends with a workgroup
// barrier (its result is read across waves).
#ifndef LOMA_XL_WR
#define LOMA_XL_WR 1
#endif
// TWR (with WR, N = 512): the first twiddles of the passes come from a
// copy in LDS, twl[j] = tw[j] (Ns = 64) and twl[64 + j % 8] = tw[(j % 8) * 8]
// (Ns = 8); the passes then issue no global loads, so they do not wait
// (vmcnt counts loads and stores in order) for the previous field's stores.
template <typename T, typename T2, int N, int Ns, int NT, int RS, int PAD = LOMA_LDS_PAD,
          bool WR = false, bool TWR = false>
__device__ __forceinline__ T2* fft_rows_ct(T2* a, T2* b, int nrows, const T2* __restrict__ tw,
                                           int s, const T2* twl = nullptr) {
  if constexpr (Ns >= N) {
    return a;
  } else {
    constexpr int R = ct_radix(N / Ns);
    constexpr int nb = N / R;
    constexpr int tstep = N / (Ns * R);
    static_assert(!TWR || (WR && N == 512 && R == 8), "register twiddles: 512 = 8 x 8 x 8");
    // the pass's index arithmetic starts here, not hoisted into earlier
    // passes where it would hold registers across them
    int t0 = threadIdx.x;
    asm volatile("" : "+v"(t0));
    const int total = nrows * nb;
    for (int it = t0; it < total; it += NT) {
      const int row = it / nb, j = it - row * nb;
      if constexpr (TWR)
        stockham_item_p1<R, T, T2, PAD>(a, b, row * RS, j, nb, Ns,
                                        Ns == 8 ? twl[64 + (j & 7)] : twl[j], s);
      else
        stockham_item<R, T, T2, PAD>(a, b, row * RS, j, nb, Ns, tstep, tw, s);
    }
    static_assert(!WR || (nb == 64 && NT % 64 == 0), "wave-local rows need 64 items per row");
    if constexpr (WR && Ns * R < N) wave_sync();
    else __syncthreads();
    return fft_rows_ct<T, T2, N, Ns * R, NT, RS, PAD, WR, TWR>(b, a, nrows, tw, s, twl);
  }
}

template <typename T, typename T2, int E, int N, int Ns, int NT, int PAD, int NSTOP = N>
__device__ __forceinline__ void fft_rows_ip_ct(T2* a, int nrows, const T2* __restrict__ tw,
                                               int s, T2* v) {
  if constexpr (Ns < NSTOP) {
    constexpr int R = ct_radix(N / Ns);
    constexpr int nb = N / R;
    constexpr int tstep = N / (Ns * R);
    constexpr int nitems = (E + R - 1) / R;
    constexpr int RS = N + (PAD ? N / 8 : 0) + 1;
    const int total = nrows * nb;
    int t0 = threadIdx.x;
    asm volatile("" : "+v"(t0));
#pragma unroll
    for (int k = 0; k < nitems; ++k) {
      const int it = t0 + k * NT;
      if (it < total) {
        const int row = it / nb, j = it - row * nb;
#pragma unroll
        for (int r = 0; r < R; ++r) v[k * R + r] = a[row * RS + lds_i<PAD>(j + r * nb)];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < nitems; ++k) {
      const int it = t0 + k * NT;
      if (it < total) {
        const int row = it / nb, j = it - row * nb;
        const int jm = j % Ns;
        T2* w = v + k * R;
        if (Ns > 1) twiddle<R>(w, tw, jm, tstep, s);
        if (R == 2) dft2<T, T2>(w);
        if (R == 3) dft3<T, T2>(w, s);
        if (R == 4) dft4<T, T2>(w, s);
        if (R == 5) dft5<T, T2>(w, s);
        if (R == 8) dft8<T, T2>(w, s);
        const int idxD = (j / Ns) * Ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) a[row * RS + lds_i<PAD>(idxD + r * Ns)] = w[r];
      }
    }
    __syncthreads();
    fft_rows_ip_ct<T, T2, E, N, Ns * R, NT, PAD, NSTOP>(a, nrows, tw, s, v);
  }
}

// Ns of the last pass of the compile-time plan of length N
constexpr int ct_last_ns(int N) {
  int Ns = 1;
  while (Ns * ct_radix(N / Ns) < N) Ns *= ct_radix(N / Ns);
  return Ns;
}
template <int R, typename T, typename T2>
__device__ __forceinline__ void dft_r(T2* w, int s) {
  if constexpr (R == 2) dft2<T, T2>(w);
  if constexpr (R == 3) dft3<T, T2>(w, s);
  if constexpr (R == 4) dft4<T, T2>(w, s);
  if constexpr (R == 5) dft5<T, T2>(w, s);
  if constexpr (R == 8) dft8<T, T2>(w, s);
}

// NN > 0: compile-time length NN; NN = 0: the run-time plan
// LOMA_XL_RS1 = 1 stores x-line rows with an odd stride (in entries); it
// was measured slower than the even stride at 512x513x256, so it is off
#ifndef LOMA_XL_RS1
#define LOMA_XL_RS1 0
#endif
template <int NN>
__device__ __forceinline__ int xl_row_stride(int N) { return (NN ? NN : N) + LOMA_XL_RS1; }

// TWR: the 512-point plan's first twiddles from the LDS copy twl
// (fft_rows_ct)
template <typename T, typename T2, int NN, int NT, int PAD = LOMA_LDS_PAD, bool TWR = false>
__device__ __forceinline__ T2* fft_any(T2* a, T2* b, int N, int nrows, const FftPlan& pl,
                                       const T2* __restrict__ tw, int s,
                                       const T2* twl = nullptr) {
  if constexpr (NN > 0)
    return fft_rows_ct<T, T2, NN, 1, NT, NN + LOMA_XL_RS1, PAD,
                       (NN == 512 && LOMA_XL_WR != 0), TWR>(a, b, nrows, tw, s, twl);
  else
    return fft_rows<T, T2, PAD>(a, b, N, N + LOMA_XL_RS1, nrows, pl, tw, s);
}
template <typename T, typename T2, int E, int NN, int NT, int PAD>
__device__ __forceinline__ void fft_any_ip(T2* a, int N, int nrows, const FftPlan& pl,
                                           const T2* __restrict__ tw, int s) {
  if constexpr (NN > 0) {
    T2 v[E + 8];
    fft_rows_ip_ct<T, T2, E, NN, 1, NT, PAD>(a, nrows, tw, s, v);
  } else {
    fft_rows_ip<T, T2, E, PAD>(a, N, nrows, pl, tw, s);
  }
}

// ------------------------------------------------------------------------
// z transforms: spectral [i][k][j] <-> z-physical [i][z][j]
// (localyz2zy + fourz(+1), fou3D.f:103-146 / fourz(-1) + localzy2yz)
// ------------------------------------------------------------------------
// z transform of a J-row tile with register-resident ends (compile-time
// length N).  The first radix pass takes its items row-fastest (lanes
// 0..J-1 are the J y rows at one z), so its R inputs per item come straight
// from HBM in runs of J entries; it has no twiddles and writes LDS.  The
// middle passes run in LDS; the last pass, again row-fastest, writes its
// outputs z = j + r*Ns straight to HBM.  This saves the staging of the tile
// through LDS on the way in and on the way out (two LDS passes, two
// barriers).  INV: spectral [i][k][j] -> z-physical [i][z][j] (the dealiased
// band reads as zero); else the reverse, scaled by 1/N, dealiased modes
// dropped.
#ifndef LOMA_ZF_REG
#define LOMA_ZF_REG 1
#endif
template <typename T, typename T2, int E, int N, int PAD, bool INV>
__device__ __forceinline__ void zfft_reg(const T2* __restrict__ src, T2* __restrict__ dst,
                                         T2* A, int i, int j0, int J, int my, int myp, int mz,
                                         int nz, const T2* __restrict__ tw) {
  constexpr int RS = N + (PAD ? N / 8 : 0) + 1;
  constexpr int R1 = ct_radix(N), nb1 = N / R1;
  constexpr int NSL = ct_last_ns(N), RL = N / NSL, nbL = N / RL;
  static_assert(NSL > 1, "at least two passes");
  const int sgn = INV ? +1 : -1;
  T2 v[E + 8];
  // first pass
  {
    constexpr int nit = (E + R1 - 1) / R1;
    const int total = J * nb1;
#pragma unroll
    for (int q = 0; q < nit; ++q) {
      const int it = threadIdx.x + q * 256;
      if (it < total) {
        const int jj = it % J, j = it / J, jy = j0 + jj;
        T2* w = v + q * R1;
#pragma unroll
        for (int r = 0; r < R1; ++r) {
          const int z = j + r * nb1;
          T2 x = T2{0, 0};
          if (jy < my) {
            if constexpr (INV) {
              const int k = z <= nz ? z : (z >= N - nz ? z - N + mz : -1);
              if (k >= 0) x = src[((size_t)i * mz + k) * myp + jy];
            } else {
              x = src[((size_t)i * N + z) * myp + jy];
            }
          }
          w[r] = x;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < nit; ++q) {
      const int it = threadIdx.x + q * 256;
      if (it < total) {
        const int jj = it % J, j = it / J;
        T2* w = v + q * R1;
        dft_r<R1, T, T2>(w, sgn);
#pragma unroll
        for (int r = 0; r < R1; ++r) A[jj * RS + lds_i<PAD>(j * R1 + r)] = w[r];
      }
    }
  }
  __syncthreads();
  // middle passes (Ns = R1 .. NSL)
  fft_rows_ip_ct<T, T2, E, N, R1, 256, PAD, NSL>(A, J, tw, sgn, v);
  // last pass.  FP64: every item's first twiddle is loaded before the first
  // store (a load after a store waits for it: vmcnt counts both, in order)
  // and pinned there so that the compiler does not sink the loads below the
  // earlier items' stores: 0.949 -> 0.920 ms (inverse) and 1.224 -> 1.214 ms
  // (forward) per launch at 512x513x256.  FP32 lost with it (inverse 0.458
  // -> 0.478 ms) and loads each item's twiddle in the item.
  {
    constexpr int nit = (E + RL - 1) / RL;
    const int total = J * nbL;
    const T sc = INV ? (T)1 : (T)(1.0 / (double)N);
    constexpr bool TWA = sizeof(T) == 8 && LOMA_TW_REC;
    T2 tw1[nit];
    if constexpr (TWA) {
#pragma unroll
      for (int q = 0; q < nit; ++q) {
        const int it = threadIdx.x + q * 256;
        tw1[q] = it < total ? tw[it / J] : T2{1, 0};
      }
#pragma unroll
      for (int q = 0; q < nit; ++q) asm volatile("" : "+v"(tw1[q].x), "+v"(tw1[q].y));
    }
#pragma unroll
    for (int q = 0; q < nit; ++q) {
      const int it = threadIdx.x + q * 256;
      if (it < total) {
        const int jj = it % J, j = it / J, jy = j0 + jj;
        T2* w = v + q * RL;
#pragma unroll
        for (int r = 0; r < RL; ++r) w[r] = A[jj * RS + lds_i<PAD>(j + r * nbL)];
        if constexpr (TWA) twiddle_pow<RL>(w, tw1[q], sgn);
        else twiddle<RL>(w, tw, j, 1, sgn);
        dft_r<RL, T, T2>(w, sgn);
        if (jy < my) {
#pragma unroll
          for (int r = 0; r < RL; ++r) {
            const int z = j + r * NSL;
            if constexpr (INV) {
              dst[((size_t)i * N + z) * myp + jy] = w[r];
            } else {
              const int k = z <= nz ? z : (z >= N - nz ? z - N + mz : -1);
              if (k >= 0) {
                T2 y = w[r];
                y.x *= sc;
                y.y *= sc;
                dst[((size_t)i * mz + k) * myp + jy] = y;
              }
            }
          }
        }
      }
    }
  }
}

// One workgroup: J consecutive y entries of one kx plane of one field, all
// z; one LDS buffer, in-place FFT (J * mgalz <= 256 * E).
template <typename T, int E, int NN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LOMA_ZF_WPE))) void k_zfft_inv(FftArgs a) {
  const int bx = blockIdx.x, by = blockIdx.y;
#include "loma_zfft_inv_body.inc"
}
template <typename T, int E, int NN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LOMA_ZF_WPE))) void k_zfft_inv_batch(
    const FftArgs* __restrict__ args, const int* __restrict__ first, int nrun) {
  const int r = batch_run(first, nrun, blockIdx.x);
  const FftArgs& a = args[r];
  const int bx = blockIdx.x - first[r], by = blockIdx.y;
#include "loma_zfft_inv_body.inc"
}

template <typename T, int E, int NN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LOMA_ZF_WPE))) void k_zfft_fwd(FftArgs a) {
  const int bx = blockIdx.x, by = blockIdx.y;
#include "loma_zfft_fwd_body.inc"
}
template <typename T, int E, int NN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LOMA_ZF_WPE))) void k_zfft_fwd_batch(
    const FftArgs* __restrict__ args, const int* __restrict__ first, int nrun) {
  const int r = batch_run(first, nrun, blockIdx.x);
  const FftArgs& a = args[r];
  const int bx = blockIdx.x - first[r], by = blockIdx.y;
#include "loma_zfft_fwd_body.inc"
}

// ------------------------------------------------------------------------
// x lines: inverse x transform of the 8 (12 with Z) transform inputs, phys2
// algebra, forward x transform of the 11 (12) outputs, for 2*pairs lines
// (fixed z, consecutive y) per 256-thread workgroup.  Two real lines share
// one complex FFT.  The algebra runs in two passes so that only rho*u_i, T
// and one accumulator per point stay in registers:
//   pass 1: T, rho u, rho v, rho w -> u_i = rho u_i T and the six rho u_i u_j
//   pass 2: lap T, T_x, T_y, T_z (and Z's) accumulated into
//           r = T (lap T / Pe - rho u . grad T), then DT, drho (and DZ).
// Every global access of a field is a run of 2*PAIRS consecutive y entries.
// ------------------------------------------------------------------------
// 256 threads per x-line workgroup (512 measured slower on MI355X: more
// spills at the 128-VGPR budget that 8 waves per SIMD pair impose)
#ifndef LOMA_XL_THREADS
#define LOMA_XL_THREADS 256
#endif
constexpr int XL_THREADS = LOMA_XL_THREADS;

// Workgroups are dealt to the 8 XCDs round-robin; this maps the launch so
// that tiles t, t+1, ... (neighbouring y runs that share cache lines) run on
// one XCD and share its L2.  Launch 8*ceil(n/8) groups; tiles >= n idle.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int per = (n + 7) / 8;
  return (b & 7) * per + (b >> 3);
}
__host__ inline int xcd_groups(int n) { return 8 * ((n + 7) / 8); }

// Per-thread share of one input field's lines: XR (kx, line pair) items of
// two complex coefficients, all loaded before the first is packed.  The
// FP32 512-point path fetches the next field's share while the current one
// is transformed (LOMA_XL_PF); elsewhere the held registers cost more than
// the hidden latency (FP32 at 3 waves per SIMD) or gained nothing (FP64).
template <typename T2, int XR>
struct XlIn {
  T2 a[XR], b[XR];
};

template <typename T2, int XR, int NN>
__device__ __forceinline__ void xl_fetch(const FftArgs& a, const T2* __restrict__ src, int z,
                                         int j0, int pairs, XlIn<T2, XR>& v) {
  const int np = NN ? NN / 3 : a.rc.nplanes, my = a.rc.my, mgz = a.rc.mgalz;
  const int lp = __builtin_ctz(pairs);  // pairs is a power of two
#pragma unroll
  for (int r = 0; r < XR; ++r) {
    const int it = threadIdx.x + r * XL_THREADS;
    T2 ca = T2{0, 0}, cb = T2{0, 0};
    if (it < np * pairs) {
      const int ii = it >> lp, p = it & (pairs - 1);
      const int ja = j0 + 2 * p, jb = ja + 1;
      const size_t base = ((size_t)ii * mgz + z) * a.rc.myp;
      if (ja < my) ca = src[base + ja];
      if (jb < my) cb = src[base + jb];
    }
    v.a[r] = ca;
    v.b[r] = cb;
  }
}

// Fields per x transform.  The rows of field f are f*pairs ..
// f*pairs+pairs-1, so one transform of 2*pairs rows keeps every thread of
// a radix pass busy where pairs alone is too few rows: FP64 at 512 points
// has 2 pairs, 128 radix-8 butterflies for 256 threads (two fields per
// transform measured 5.13 against 5.58 ms per x-line launch at
// 512x513x256).  FP32 has 4 pairs and gains nothing from two fields.
#ifndef LOMA_XL_DUAL64
#define LOMA_XL_DUAL64 1
#endif
#ifndef LOMA_XL_DUAL32
#define LOMA_XL_DUAL32 0
#endif
template <typename T>
constexpr int xl_nf() {
  return (sizeof(T) == 8 ? LOMA_XL_DUAL64 : LOMA_XL_DUAL32) ? 2 : 1;
}
// x-line LDS rows: one pad entry after every eight (lds_i<1>) spreads the
// strided butterfly writes of the first two radix-8 passes over the banks.
// FP64 (16-byte entries): 4.91 against 5.20 ms per launch at 512x513x256;
// FP32 measured slower padded (3.03 against 2.90 ms) and stays unpadded.
#ifndef LOMA_XL_PAD64
#define LOMA_XL_PAD64 1
#endif
#ifndef LOMA_XL_PAD32
#define LOMA_XL_PAD32 0
#endif
template <typename T>
constexpr int xl_pad() {
  return sizeof(T) == 8 ? LOMA_XL_PAD64 : LOMA_XL_PAD32;
}

// fetched lines of NF fields (xl_fetch) -> Hermitian-packed complex rows in
// A (row f*pairs + p = lines j0+2p, j0+2p+1 of field f); no barrier.
template <typename T, typename T2, int XR, int NN, int NF>
__device__ __forceinline__ void xl_pack_from(const FftArgs& a, const XlIn<T2, XR>* v, T2* A,
                                             int pairs) {
  constexpr int XP = xl_pad<T>();
  const int N = NN ? NN : a.rc.mgalx, np = NN ? NN / 3 : a.rc.nplanes;
  const int lp = __builtin_ctz(pairs);
  const int RS = xl_row_stride<NN>(N);
  // the dealiased band np..N-np is zero
  const int nzero = N - 2 * np + 1;
  for (int it = threadIdx.x; it < NF * pairs * nzero; it += XL_THREADS) {
    const int p = it / nzero, q = it - p * nzero;
    A[lds_i<XP>(p * RS + np + q)] = T2{0, 0};
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) {
#pragma unroll
    for (int r = 0; r < XR; ++r) {
      const int it = threadIdx.x + r * XL_THREADS;
      if (it < np * pairs) {
        const int ii = it >> lp, p = it & (pairs - 1);
        const T2 ca = v[f].a[r], cb = v[f].b[r];
        const int rb = (f * pairs + p) * RS;
        if (ii == 0) {
          // the imaginary part of the kx=0 coefficient is ignored (rft, rftsingle.f:87-91)
          A[lds_i<XP>(rb)] = T2{ca.x, cb.x};
        } else {
          A[lds_i<XP>(rb + ii)] = T2{ca.x - cb.y, ca.y + cb.x};           // X_a + i X_b
          A[lds_i<XP>(rb + N - ii)] = T2{ca.x + cb.y, -ca.y + cb.x};      // conj(X_a) + i conj(X_b)
        }
      }
    }
  }
}
// lines of NF fields (W layout [i][z][j]; src1 unused for NF = 1) -> the
// packed rows in A; no barrier.
template <typename T, typename T2, int XR, int NN, int NF>
__device__ __forceinline__ void xl_pack(const FftArgs& a, const T2* __restrict__ src0,
                                        const T2* __restrict__ src1, T2* A, int z, int j0,
                                        int pairs) {
  XlIn<T2, XR> v[NF];
  xl_fetch<T2, XR, NN>(a, src0, z, j0, pairs, v[0]);
  if constexpr (NF > 1) xl_fetch<T2, XR, NN>(a, src1, z, j0, pairs, v[1]);
  xl_pack_from<T, T2, XR, NN, NF>(a, v, A, pairs);
}

// Hermitian unpacking of the transformed pairs in R into fields dst0 (rows
// 0..pairs-1) and dst1 (rows pairs..2*pairs-1), scaled by 1/N; no barrier.
template <typename T, typename T2, int NN, int NF>
__device__ __forceinline__ void xl_unpack(const FftArgs& a, const T2* R, T2* __restrict__ dst0,
                                          T2* __restrict__ dst1, int z, int j0, int pairs) {
  constexpr int XP = xl_pad<T>();
  const int N = NN ? NN : a.rc.mgalx, np = NN ? NN / 3 : a.rc.nplanes, my = a.rc.my,
            mgz = a.rc.mgalz;
  const int lp = __builtin_ctz(pairs);
  const int RS = xl_row_stride<NN>(N);
  const T hs = (T)(0.5 / (double)N);
  const int per = np * pairs;
  for (int it = threadIdx.x; it < NF * per; it += XL_THREADS) {
    const int f = NF > 1 && it >= per ? 1 : 0, jt = it - f * per;
    const int ii = jt >> lp, p = jt & (pairs - 1);
    const int ja = j0 + 2 * p, jb = ja + 1;
    const int rb = (f * pairs + p) * RS;
    const T2 zk = R[lds_i<XP>(rb + ii)];
    T2 zc = R[lds_i<XP>(rb + (N - ii) % N)];
    zc.y = -zc.y;
    T2 ya{(zk.x + zc.x) * hs, (zk.y + zc.y) * hs};
    T2 yb{(zk.y - zc.y) * hs, -(zk.x - zc.x) * hs};  // (zk - zc) / (2i)
    if (ii == 0) { ya.y = 0; yb.y = 0; }
    T2* dst = f ? dst1 : dst0;
    const size_t base = ((size_t)ii * mgz + z) * a.rc.myp;
    if (ja < my) dst[base + ja] = ya;
    if (jb < my) dst[base + jb] = yb;
  }
}

// first twiddles of the 512-point passes (radix 8, 8, 8) for lane j:
// tw[(j % 8) * 8] (the Ns = 8 pass) and tw[j] (Ns = 64)
template <typename T2>
struct XlTw {
  T2 w8, w64;
};
template <typename T2>
__device__ __forceinline__ XlTw<T2> xl_tw(const FftArgs& a) {
  const T2* tw = (const T2*)a.twx;
  const int j = threadIdx.x & 63;
  return XlTw<T2>{tw[(j & 7) * 8], tw[j]};
}
// LOMA_XL_TWR = 1: the 512-point x lines outside the register-resident path
// (FP64) take those twiddles from a 72-entry LDS copy (fft_rows_ct's TWR;
// in registers they held 8 VGPRs across the kernel and made it spill)
#ifndef LOMA_XL_TWR
#define LOMA_XL_TWR 1
#endif

// pack, then the inverse FFT.  Returns the buffer holding the physical pairs.
// TWR: the 512-point plan's first twiddles from the 72-entry LDS copy twl;
// TWF: every twiddle from twl holding the whole table (other lengths)
template <typename T, typename T2, int XR, int NN, int NF, bool TWR = false, bool TWF = false>
__device__ __forceinline__ T2* xl_inverse(const FftArgs& a, const T2* __restrict__ src0,
                          const T2* __restrict__ src1, T2* A, T2* B, int z, int j0, int pairs,
                          const T2* twl = nullptr) {
  xl_pack<T, T2, XR, NN, NF>(a, src0, src1, A, z, j0, pairs);
  __syncthreads();
  const int N = NN ? NN : a.rc.mgalx;
  const T2* tw = TWF ? twl : (const T2*)a.twx;
  return fft_any<T, T2, NN, XL_THREADS, xl_pad<T>(), TWR>(A, B, N, NF * pairs, a.plx, tw, +1,
                                                           twl);
}

// forward FFT of A (rows of NF outputs, filled by the caller, then
// synchronised here) and the unpacking
template <typename T, typename T2, int NN, int NF, bool TWR = false, bool TWF = false>
__device__ __forceinline__ void xl_forward(const FftArgs& a, T2* __restrict__ dst0, T2* __restrict__ dst1,
                           T2* A, T2* B, int z, int j0, int pairs, const T2* twl = nullptr) {
  const int N = NN ? NN : a.rc.mgalx;
  __syncthreads();
  const T2* tw = TWF ? twl : (const T2*)a.twx;
  const T2* R = fft_any<T, T2, NN, XL_THREADS, xl_pad<T>(), TWR>(A, B, N, NF * pairs, a.plx,
                                                                  tw, -1, twl);
  xl_unpack<T, T2, NN, NF>(a, R, dst0, dst1, z, j0, pairs);
  __syncthreads();
}

// Register-resident ends of the 512-point transform (radix 8, 8, 8) for 4
// pairs (4 rows, 256 items per pass: one per thread, row tid/64, item
// j = tid%64).  The last inverse pass reads and writes entries j + 64r of its
// row, which are exactly the entries x0 + 64m the physics assigns to the
// thread; so the inverse keeps its last pass in registers (no store, no
// barrier, no reload) and the forward takes its first pass from registers.
// Barriers of the register-resident path.  Its radix passes are local to a
// wavefront (wave p owns line pair p), so LOMA_XL_WAVE = 2 (default) puts
// only a wavefront barrier between them and keeps workgroup barriers around
// the pack and unpack, whose coalesced lane mapping crosses waves: 2.74
// against 2.77 ms per launch at 512x513x256.  LOMA_XL_WAVE = 1 also packs
// and unpacks per wave (16 bytes per lane instead of 64-byte runs, no
// workgroup barrier at all): 2.80 ms.  0: workgroup barriers throughout.
#ifndef LOMA_XL_WAVE
#define LOMA_XL_WAVE 2
#endif
// barrier between two wave-local steps (LOMA_XL_WAVE >= 1: the lanes of
// one wavefront only) and between steps that cross waves (a workgroup
// barrier unless everything is wave-local, LOMA_XL_WAVE = 1)
template <typename T, typename T2>
__device__ __forceinline__ void xl_sync() {
  if constexpr (LOMA_XL_WAVE >= 1) wave_sync();
  else __syncthreads();
}
template <typename T, typename T2>
__device__ __forceinline__ void xl_sync_x() {
  if constexpr (LOMA_XL_WAVE == 1) wave_sync();
  else __syncthreads();
}

// LOMA_XL_PF = 1: the register-resident path takes the first twiddles of
// its passes from registers (loaded once per workgroup) and fetches the next
// input field while the current one is transformed.  vmcnt counts loads and
// stores in order, so a twiddle load inside a pass would make the pass wait
// for the prefetch (and for the previous field's stores).
#ifndef LOMA_XL_PF
#define LOMA_XL_PF 1
#endif
static_assert(!LOMA_XL_PF || LOMA_XL_WAVE == 2, "the prefetch feeds the workgroup pack");
// inverse transform of a fetched field (LOMA_XL_PF)
template <typename T, typename T2, int XR>
__device__ __forceinline__ void xl_inverse_reg_in(const FftArgs& a, const XlIn<T2, XR>& in,
                                                  T2* A, T2* B, int pairs, const XlTw<T2>& tw,
                                                  T (&v)[8][2]) {
  constexpr int RS = 512 + LOMA_XL_RS1;
  const int row = threadIdx.x >> 6, j = threadIdx.x & 63;
  xl_pack_from<T, T2, XR, 512, 1>(a, &in, A, pairs);
  xl_sync_x<T, T2>();
  stockham_item_p1<8, T, T2, 0>(A, B, row * RS, j, 64, 1, tw.w8, +1);  // Ns = 1: no twiddles
  xl_sync<T, T2>();
  stockham_item_p1<8, T, T2, 0>(B, A, row * RS, j, 64, 8, tw.w8, +1);
  xl_sync<T, T2>();
  T2 w[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) w[r] = A[row * RS + j + r * 64];
  twiddle_pow<8>(w, tw.w64, +1);
  dft8<T, T2>(w, +1);
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    v[r][0] = w[r].x;
    v[r][1] = w[r].y;
  }
  xl_sync_x<T, T2>();  // A is packed again by the next field
}

template <typename T, typename T2, int XR>
__device__ __forceinline__ void xl_inverse_reg(const FftArgs& a, const T2* __restrict__ src,
                                               T2* A, T2* B, int z, int j0, int pairs,
                                               T (&v)[8][2]) {
  static_assert(xl_pad<T>() == 0, "register ends assume unpadded rows");
  constexpr int RS = 512 + LOMA_XL_RS1;
  const T2* tw = (const T2*)a.twx;
  const int row = threadIdx.x >> 6, j = threadIdx.x & 63;
  if constexpr (LOMA_XL_WAVE == 1) {
    // wave `row` packs line pair `row`: lane j takes coefficients ii = j + 64k
    constexpr int np = 512 / 3;
    const int ja = j0 + 2 * row, jb = ja + 1, my = a.rc.my, mgz = a.rc.mgalz;
    T2* Ar = A + row * RS;
    T2 ca[3], cb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int ii = j + 64 * k;
      ca[k] = T2{0, 0};
      cb[k] = T2{0, 0};
      if (ii < np) {
        const size_t base = ((size_t)ii * mgz + z) * a.rc.myp;
        if (ja < my) ca[k] = src[base + ja];
        if (jb < my) cb[k] = src[base + jb];
      }
    }
    for (int q = j; q < 512 - 2 * np + 1; q += 64) Ar[np + q] = T2{0, 0};  // dealiased band
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int ii = j + 64 * k;
      if (ii < np) {
        if (ii == 0) {
          Ar[0] = T2{ca[k].x, cb[k].x};  // kx = 0: imaginary part ignored (rftsingle.f:87-91)
        } else {
          Ar[ii] = T2{ca[k].x - cb[k].y, ca[k].y + cb[k].x};
          Ar[512 - ii] = T2{ca[k].x + cb[k].y, -ca[k].y + cb[k].x};
        }
      }
    }
  } else {
    xl_pack<T, T2, XR, 512, 1>(a, src, nullptr, A, z, j0, pairs);
  }
  xl_sync_x<T, T2>();
  stockham_item<8, T, T2, 0>(A, B, row * RS, j, 64, 1, 64, tw, +1);
  xl_sync<T, T2>();
  stockham_item<8, T, T2, 0>(B, A, row * RS, j, 64, 8, 8, tw, +1);
  xl_sync<T, T2>();
  T2 w[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) w[r] = A[row * RS + j + r * 64];
  twiddle<8>(w, tw, j, 1, +1);
  dft8<T, T2>(w, +1);
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    v[r][0] = w[r].x;
    v[r][1] = w[r].y;
  }
  xl_sync_x<T, T2>();  // A is packed again by the next field
}

template <typename T, typename T2, bool PF>
__device__ __forceinline__ void xl_forward_reg(const FftArgs& a, T2* __restrict__ dst, T2* A,
                                               T2* B, int z, int j0, int pairs,
                                               const T (&v)[8][2], const XlTw<T2>& twr) {
  static_assert(xl_pad<T>() == 0, "register ends assume unpadded rows");
  constexpr int RS = 512 + LOMA_XL_RS1;
  const T2* tw = (const T2*)a.twx;
  const int row = threadIdx.x >> 6, j = threadIdx.x & 63;
  T2 w[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) w[r] = T2{v[r][0], v[r][1]};
  dft8<T, T2>(w, -1);  // first pass (no twiddles): entries j + 64r -> 8j + r
#pragma unroll
  for (int r = 0; r < 8; ++r) A[row * RS + 8 * j + r] = w[r];
  xl_sync<T, T2>();
  if constexpr (PF) {
    stockham_item_p1<8, T, T2, 0>(A, B, row * RS, j, 64, 8, twr.w8, -1);
    xl_sync<T, T2>();
    stockham_item_p1<8, T, T2, 0>(B, A, row * RS, j, 64, 64, twr.w64, -1);
  } else {
    stockham_item<8, T, T2, 0>(A, B, row * RS, j, 64, 8, 8, tw, -1);
    xl_sync<T, T2>();
    stockham_item<8, T, T2, 0>(B, A, row * RS, j, 64, 64, 1, tw, -1);
  }
  xl_sync_x<T, T2>();
  if constexpr (LOMA_XL_WAVE == 1) {
    // wave `row` unpacks and stores line pair `row`
    constexpr int np = 512 / 3;
    const int ja = j0 + 2 * row, jb = ja + 1, my = a.rc.my, mgz = a.rc.mgalz;
    const T2* Ar = A + row * RS;
    const T hs = (T)(0.5 / 512.0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int ii = j + 64 * k;
      if (ii < np) {
        const T2 zk = Ar[ii];
        T2 zc = Ar[(512 - ii) & 511];
        zc.y = -zc.y;
        T2 ya{(zk.x + zc.x) * hs, (zk.y + zc.y) * hs};
        T2 yb{(zk.y - zc.y) * hs, -(zk.x - zc.x) * hs};  // (zk - zc) / (2i)
        if (ii == 0) { ya.y = 0; yb.y = 0; }
        const size_t base = ((size_t)ii * mgz + z) * a.rc.myp;
        if (ja < my) dst[base + ja] = ya;
        if (jb < my) dst[base + jb] = yb;
      }
    }
  } else {
    xl_unpack<T, T2, 512, 1>(a, A, dst, nullptr, z, j0, pairs);
  }
  xl_sync_x<T, T2>();
}

// Mean split: the transform inputs carry no 00 mode; its y-profile comes in
// FP64 (mean_in) and is added back here, per line.  Each output has an FP64
// per-y estimate g (the same algebra on the profiles) subtracted before the
// forward transform and stored in mean_out for phase C to add back, so the
// transforms (FP32 in mixed mode) only carry perturbation-sized values.
constexpr int XL_MAXL = 32;  // lines per workgroup (2 * pairs, pairs <= 16)

// Makes register arrays opaque to the optimiser at a loop head, so that
// per-output expressions are not hoisted out of the output loop and kept
// live for all outputs at once.
template <typename T, int E>
__device__ __forceinline__ void opaque_regs(T (&v)[E][2]) {
#pragma unroll
  for (int m = 0; m < E; ++m)
#pragma unroll
    for (int h = 0; h < 2; ++h) asm volatile("" : "+v"(v[m][h]));
}

// RK3 coefficients of a substep times dt; dtalpha = ab * dtbeta, so that
// the constraint source drho = d/(T T+ dtbeta) - ab r/T^2 takes one division:
// (d T - dtalpha r T+) / (T^2 T+ dtbeta)
struct XlRk {
  double dtxi, dtgamma, dtbeta, ab, dtalpha;
};
__device__ __forceinline__ XlRk xl_rk(double dt, int rkstep) {
  const double gam[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
  const double xis[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
  const double p0 = 5.0 / 66.0, p1 = -1.0;
  const double b3 = (p1 * (gam[0] * gam[2] + gam[1] * gam[2] + xis[1] * gam[2] + gam[0] * xis[2]) -
                     gam[0] * gam[1] * gam[2] - gam[0] * gam[1] * xis[2]) /
                    (p1 - gam[0] * gam[1]);
  const double betas[3] = {gam[0] - p0, p1, b3};
  const double alphas[3] = {p0, gam[1] + xis[1] - p1, gam[2] + xis[2] - b3};
  const int rk = rkstep - 1;
  return XlRk{dt * xis[rk], dt * gam[rk], dt * betas[rk], alphas[rk] / betas[rk],
              dt * alphas[rk]};
}

// Compiled for two waves per SIMD (LOMA_XL_WPE; FP64 always): at three,
// FP64 spills heavily (and its Z-transport build gave wrong results), and
// FP32 spills enough to measure slower (3.33 vs 3.22 ms at 512x513x256).
// XREG: the register-resident transform ends (FP32, 512 points, 4 pairs)
template <typename T, bool WZ, int EPT, int NN, bool XREG>
__global__ __launch_bounds__(XL_THREADS)
__attribute__((amdgpu_waves_per_eu(XL_THREADS == 512 ? 4 : (sizeof(T) == 8 ? 2 : LOMA_XL_WPE)))) void k_xline(FftArgs a) {
  const int bx = blockIdx.x;
#include "loma_xline_body.inc"
}
template <typename T, bool WZ, int EPT, int NN, bool XREG>
__global__ __launch_bounds__(XL_THREADS)
__attribute__((amdgpu_waves_per_eu(XL_THREADS == 512 ? 4 : (sizeof(T) == 8 ? 2 : LOMA_XL_WPE)))) void k_xline_batch(
    const FftArgs* __restrict__ args, const int* __restrict__ first, int nrun) {
  const int r = batch_run(first, nrun, blockIdx.x);
  const FftArgs& a = args[r];
  const int bx = blockIdx.x - first[r];
#include "loma_xline_body.inc"
}

// fourz on standalone rows (FFTPACK convention), for the per-operator API
__global__ __launch_bounds__(256) void k_fourz(FftPlan pl, const double2* tw, double2* c,
                                               int nrows, int iopt) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = pl.n;
  double2* A = (double2*)smem;
  double2* B = A + lds_entries(N);
  const int row = blockIdx.x;
  if (row >= nrows) return;
  for (int it = threadIdx.x; it < N; it += blockDim.x) A[lds_i(it)] = c[(size_t)row * N + it];
  __syncthreads();
  double2* R = fft_rows<double, double2>(A, B, N, N, 1, pl, tw, iopt < 0 ? -1 : 1);
  const double sc = iopt < 0 ? 1.0 / N : 1.0;
  for (int it = threadIdx.x; it < N; it += blockDim.x) {
    double2 v = R[lds_i(it)];
    v.x *= sc;
    v.y *= sc;
    c[(size_t)row * N + it] = v;
  }
}

// ------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------
static size_t zfft_lds(bool fp32, const FftArgs& a) {
  return (size_t)a.J * lds_row_stride(a.rc.mgalz, fp32 ? LOMA_ZF_PAD32 : 1) * (fp32 ? 8 : 16);
}
static size_t xline_lds(bool fp32, const FftArgs& a) {
  const int nf = fp32 ? xl_nf<float>() : xl_nf<double>();
  const int pad = fp32 ? xl_pad<float>() : xl_pad<double>();
  return (size_t)2 * lds_entries(nf * (a.L / 2) * (a.rc.mgalx + LOMA_XL_RS1), pad) *
         (fp32 ? 8 : 16);
}

// kernel instance for a transform length: compile-time plans for the
// production lengths, the run-time plan otherwise
#define LOMA_BY_LENGTH(n, LAUNCH)              \
  switch (n) {                                 \
    case 64: { constexpr int NN = 64; LAUNCH; break; }   \
    case 96: { constexpr int NN = 96; LAUNCH; break; }   \
    case 128: { constexpr int NN = 128; LAUNCH; break; } \
    case 192: { constexpr int NN = 192; LAUNCH; break; } \
    case 256: { constexpr int NN = 256; LAUNCH; break; } \
    case 384: { constexpr int NN = 384; LAUNCH; break; } \
    case 512: { constexpr int NN = 512; LAUNCH; break; } \
    default: { constexpr int NN = 0; LAUNCH; break; }    \
  }

int launch_zfft_inv(bool fp32, const FftArgs& a, hipStream_t st) {
  const int nyb = (a.rc.my + a.J - 1) / a.J;
  dim3 g(a.rc.nplanes * nyb, a.nfields), b(256);
  const size_t lds = zfft_lds(fp32, a);
  // J * mgalz <= 256 * E (checked at init)
  if (fp32) {
    LOMA_BY_LENGTH(a.rc.mgalz, hipLaunchKernelGGL((k_zfft_inv<float, LOMA_ZF_E32, NN>), g, b, lds, st, a));
  } else {
    LOMA_BY_LENGTH(a.rc.mgalz, hipLaunchKernelGGL((k_zfft_inv<double, LOMA_ZF_E64, NN>), g, b, lds, st, a));
  }
  return (int)hipGetLastError();
}

int launch_zfft_fwd(bool fp32, const FftArgs& a, hipStream_t st) {
  const int nyb = (a.rc.my + a.J - 1) / a.J;
  dim3 g(a.rc.nplanes * nyb, a.nfields), b(256);
  const size_t lds = zfft_lds(fp32, a);
  if (fp32) {
    LOMA_BY_LENGTH(a.rc.mgalz, hipLaunchKernelGGL((k_zfft_fwd<float, LOMA_ZF_E32, NN>), g, b, lds, st, a));
  } else {
    LOMA_BY_LENGTH(a.rc.mgalz, hipLaunchKernelGGL((k_zfft_fwd<double, LOMA_ZF_E64, NN>), g, b, lds, st, a));
  }
  return (int)hipGetLastError();
}

// FP32 x lines at 512 points with 4 pairs take the register-resident
// transform ends (LOMA_XL_REG32 = 0 turns them off)
#ifndef LOMA_XL_REG32
#define LOMA_XL_REG32 1
#endif
template <typename T, bool WZ, int E, int NN>
static void launch_xline_nn(const FftArgs& a, dim3 g, dim3 b, size_t lds, hipStream_t st) {
  if constexpr (sizeof(T) == 4 && NN == 512 && E == 8 && LOMA_XL_REG32 && xl_pad<T>() == 0 &&
                xl_nf<T>() == 1) {
    if (a.L == 8) {
      hipLaunchKernelGGL((k_xline<T, WZ, E, NN, true>), g, b, lds, st, a);
      return;
    }
  }
  hipLaunchKernelGGL((k_xline<T, WZ, E, NN, false>), g, b, lds, st, a);
}

int launch_xline(bool fp32, const FftArgs& a, hipStream_t st) {
  const int nyb = (a.rc.my + a.L - 1) / a.L;
  dim3 g(xcd_groups(a.rc.mgalz * nyb)), b(XL_THREADS);
  const size_t lds = xline_lds(fp32, a);
  // entries per thread: pairs * mgalx <= XL_THREADS * EPT, i.e. at most
  // 1024 (FP64) / 2048 (FP32) entries per workgroup (checked at init)
  constexpr int E64 = 1024 / XL_THREADS, E32 = 2048 / XL_THREADS;
  if (fp32) {
    if (a.with_z) {
      LOMA_BY_LENGTH(a.rc.mgalx, (launch_xline_nn<float, true, E32, NN>(a, g, b, lds, st)));
    } else {
      LOMA_BY_LENGTH(a.rc.mgalx, (launch_xline_nn<float, false, E32, NN>(a, g, b, lds, st)));
    }
  } else {
    if (a.with_z) {
      LOMA_BY_LENGTH(a.rc.mgalx, (launch_xline_nn<double, true, E64, NN>(a, g, b, lds, st)));
    } else {
      LOMA_BY_LENGTH(a.rc.mgalx, (launch_xline_nn<double, false, E64, NN>(a, g, b, lds, st)));
    }
  }
  return (int)hipGetLastError();
}

// ---- batched launches (one class of runs: same length, tiles, fields) ----
int zfft_blocks(const FftArgs& a) { return a.rc.nplanes * ((a.rc.my + a.J - 1) / a.J); }
int xline_blocks(const FftArgs& a) {
  return xcd_groups(a.rc.mgalz * ((a.rc.my + a.L - 1) / a.L));
}

int launch_zfft_batch(bool fp32, bool inv, const FftArgs& proto, const FftArgs* args,
                      const int* first, int nrun, int nblocks, hipStream_t st) {
  if (nblocks <= 0 || nrun <= 0) return 0;
  dim3 g(nblocks, proto.nfields), b(256);
  const size_t lds = zfft_lds(fp32, proto);
  if (fp32) {
    if (inv) {
      LOMA_BY_LENGTH(proto.rc.mgalz, hipLaunchKernelGGL((k_zfft_inv_batch<float, LOMA_ZF_E32, NN>), g, b, lds, st, args, first, nrun));
    } else {
      LOMA_BY_LENGTH(proto.rc.mgalz, hipLaunchKernelGGL((k_zfft_fwd_batch<float, LOMA_ZF_E32, NN>), g, b, lds, st, args, first, nrun));
    }
  } else {
    if (inv) {
      LOMA_BY_LENGTH(proto.rc.mgalz, hipLaunchKernelGGL((k_zfft_inv_batch<double, LOMA_ZF_E64, NN>), g, b, lds, st, args, first, nrun));
    } else {
      LOMA_BY_LENGTH(proto.rc.mgalz, hipLaunchKernelGGL((k_zfft_fwd_batch<double, LOMA_ZF_E64, NN>), g, b, lds, st, args, first, nrun));
    }
  }
  return (int)hipGetLastError();
}

template <typename T, bool WZ, int E, int NN>
static void launch_xline_batch_nn(const FftArgs& proto, dim3 g, dim3 b, size_t lds,
                                  hipStream_t st, const FftArgs* args, const int* first,
                                  int nrun) {
  if constexpr (sizeof(T) == 4 && NN == 512 && E == 8 && LOMA_XL_REG32 && xl_pad<T>() == 0 &&
                xl_nf<T>() == 1) {
    if (proto.L == 8) {
      hipLaunchKernelGGL((k_xline_batch<T, WZ, E, NN, true>), g, b, lds, st, args, first, nrun);
      return;
    }
  }
  hipLaunchKernelGGL((k_xline_batch<T, WZ, E, NN, false>), g, b, lds, st, args, first, nrun);
}

int launch_xline_batch(bool fp32, const FftArgs& proto, const FftArgs* args, const int* first,
                       int nrun, int nblocks, hipStream_t st) {
  if (nblocks <= 0 || nrun <= 0) return 0;
  dim3 g(nblocks), b(XL_THREADS);
  const size_t lds = xline_lds(fp32, proto);
  constexpr int E64 = 1024 / XL_THREADS, E32 = 2048 / XL_THREADS;
  if (fp32) {
    if (proto.with_z) {
      LOMA_BY_LENGTH(proto.rc.mgalx, (launch_xline_batch_nn<float, true, E32, NN>(proto, g, b, lds, st, args, first, nrun)));
    } else {
      LOMA_BY_LENGTH(proto.rc.mgalx, (launch_xline_batch_nn<float, false, E32, NN>(proto, g, b, lds, st, args, first, nrun)));
    }
  } else {
    if (proto.with_z) {
      LOMA_BY_LENGTH(proto.rc.mgalx, (launch_xline_batch_nn<double, true, E64, NN>(proto, g, b, lds, st, args, first, nrun)));
    } else {
      LOMA_BY_LENGTH(proto.rc.mgalx, (launch_xline_batch_nn<double, false, E64, NN>(proto, g, b, lds, st, args, first, nrun)));
    }
  }
  return (int)hipGetLastError();
}

int launch_fourz(const FftPlan& pl, const double2* tw, double2* c, int nrows, int iopt,
                 hipStream_t st) {
  if (nrows <= 0) return 0;
  hipLaunchKernelGGL(k_fourz, dim3(nrows), dim3(256), (size_t)2 * lds_entries(pl.n) * 16, st, pl, tw, c,
                     nrows, iopt);
  return (int)hipGetLastError();
}


}  // namespace loma
