// loma_api.hip -- C ABI and host orchestration of the RK3 time advance.
//
// A run handle owns all device state of one mixing-layer instance (the
// reference keeps it in module globals, trunk/lomacte3/src/modulos.f, and in
// cross1's work arrays, main.f:128-206).  State stays on the device between
// calls; host transfers happen only at upload/download (restart boundary).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "loma_internal.h"
#include "loma_kernels.h"

using namespace loma;

namespace loma {
static thread_local std::string g_err;
// Serialises the calls that allocate, free, use the null stream or capture a
// graph (init, free, upload, download, graph builds): HIP invalidates a
// stream capture when another thread makes such calls while it records, and
// runs may be set up on one thread while another advances others.
static std::recursive_mutex g_setup_mu;
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace loma

#define HIPCHK(x)                                                               \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      set_error(std::string(#x) + ": " + hipGetErrorString(e_));                \
      return LOMA_ERR_DEVICE;                                                   \
    }                                                                           \
  } while (0)
#define LCHK(x)                                                                 \
  do {                                                                          \
    int e_ = (x);                                                               \
    if (e_ != 0) {                                                              \
      set_error(std::string(#x) + ": " + hipGetErrorString((hipError_t)e_));    \
      return LOMA_ERR_DEVICE;                                                   \
    }                                                                           \
  } while (0)

struct loma_run {
  int mgalx = 0, my = 0, mgalz = 0, mx = 0, mz = 0, nplanes = 0, nz = 0;
  int M = 0;            // rows per lane of the y-column solver
  int lu_stride = 0;    // doubles per Helmholtz class table (5 * 64 * M)
  int scan_levels[6] = {6, 6, 6, 6, 6, 6};  // L1 U1 L2 U2 Li Ui
  int precision = 0, flags = 0;
  bool fp32 = false, with_z = false;
  double Re = 0, alp = 0, bet = 0, pr = 0.7, cfl = 0.5, dt0 = 1e-4;
  int ncfl = 1;
  FdOps fd;
  std::vector<double> y;  // the y grid (host copy)
  RunConst rc{};
  hipStream_t stream = nullptr;
  // device memory: one allocation (arena) carved by dalloc
  void* arena = nullptr;
  size_t arena_size = 0;
  size_t bytes = 0;
  double* hstage = nullptr;  // pinned host staging of one complex field
  size_t hstage_n = 0;       // doubles
  double* tab = nullptr;  // all per-row tables
  YTab T{};
  double* dt21 = nullptr;  // plain (my,5) row-major copies for the factor kernel
  double* dt22 = nullptr;
  double* helm_lu = nullptr;
  double* helm_bc = nullptr;  // per class Neumann constants
  double* helm_h = nullptr;   // per class homogeneous solutions h1, h2 (lane-major)
  int nclass = 0;
  double2 *vor = nullptr, *phi = nullptr, *psi = nullptr, *scal = nullptr;
  double2 *ten12 = nullptr, *ten13 = nullptr, *zf = nullptr;
  void* spec[12] = {};
  void* wbuf[12] = {};
  void* wout[12] = {};
  double *rhst = nullptr, *rhsz = nullptr, *hy = nullptr;
  double *u00wk = nullptr, *v00wk = nullptr, *w00wk = nullptr, *rf0u = nullptr,
         *rf0w = nullptr;
  double *mean_in = nullptr, *mean_out = nullptr, *rbar = nullptr, *rzbar = nullptr;
  DevScalars* sc = nullptr;
  double* cflp = nullptr;  // [x-line tile][4] CFL maxima
  FftPlan plz{}, plx{};
  void *twz = nullptr, *twx = nullptr;
  double2* twz64 = nullptr;  // double twiddles for the per-operator fourz
  int J = 8, L = 4;
  int istep = 0;
  bool uploaded = false;
  bool owns_stream = false;  // stream created by the library
  // graph
  bool use_graph = false;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  // per-kernel profiling with HIP events on the run's stream
  bool prof = false;
  struct ProfEv { int cls; hipEvent_t a, b; };
  std::vector<ProfEv> prof_ev;
  // one step in flight for a request loop (loma_step_begin / loma_step_poll)
  hipEvent_t sev[2] = {nullptr, nullptr};
  DevScalars* hsc = nullptr;  // pinned copy of the scalars after the step
  bool inflight = false;
  // device checkpoint of the time-advance state (loma_snapshot / loma_restore)
  void* snap = nullptr;
  size_t snap_bytes = 0;
  int snap_istep = -1;
  // statistics window (loma_stats_*): device accumulators, the spectra rows
  // and the host-side switch and counter; allocated by loma_stats_config
  double *st_acc = nullptr, *st_spe = nullptr, *st_part = nullptr, *st_rho = nullptr,
         *st_last = nullptr;
  int* st_jsp = nullptr;
  int st_nsp = 0;
  bool st_on = false;
  int st_nacum = 0;
};

// kernel classes for profiling
enum { KC_PHASEA = 0, KC_ZINV, KC_XLINE, KC_ZFWD, KC_PHASEC, KC_FINISH, KC_PHASEC2, KC_COUNT };

static int prof_begin(loma_run* r, int cls, hipEvent_t* a) {
  if (!r->prof) return 0;
  HIPCHK(hipEventCreate(a));
  HIPCHK(hipEventRecord(*a, r->stream));
  (void)cls;
  return 0;
}
static int prof_end(loma_run* r, int cls, hipEvent_t a) {
  if (!r->prof) return 0;
  hipEvent_t b;
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(b, r->stream));
  r->prof_ev.push_back({cls, a, b});
  return 0;
}

// Every buffer of a run is carved from one device allocation, 256-byte
// aligned: one hipMalloc (and one hipFree) per run instead of dozens, so
// that setting runs up and tearing them down holds the runtime briefly while
// other runs step.
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

template <typename P>
static int dalloc(loma_run* r, P** p, size_t bytes) {
  const size_t a = align256(bytes == 0 ? 16 : bytes);
  if (!r->arena || r->bytes + a > r->arena_size) {
    set_error("loma_init: device arena exhausted (allocation plan out of step)");
    return LOMA_ERR_OOM;
  }
  *p = (P*)((char*)r->arena + r->bytes);
  r->bytes += a;
  return LOMA_OK;
}

static int pick_M(int my) {
  int m = (my + 63) / 64;
  if (m < 3) m = 3;
  static const int ok[] = {3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 16};
  for (int v : ok)
    if (v >= m) return v;
  return -1;
}

// Device bytes loma_init requests for a grid: the sum of its allocations,
// each rounded to 256 bytes (kept in step with loma_init; a test compares
// this with the byte count of a live run).
static size_t plan_bytes(int mgalx, int my, int mgalz, bool fp32, bool with_z) {
  const int M = pick_M(my);
  const size_t R = 64 * (size_t)M;
  const size_t mx = 2 * (mgalx / 3), np = mx / 2, mz = 2 * (mgalz / 3) - 1;
  const size_t nz = (mz - 1) / 2, nclass = np * (nz + 1);
  const size_t myp = (my + 15) / 16 * 16;
  const size_t S = np * mz * my, Sp = np * mz * myp, Wp = np * mgalz * myp;
  const size_t P = (size_t)mgalx * my * mgalz, es = fp32 ? 8 : 16;
  const size_t nsp = with_z ? 12 : 11, nst = with_z ? 7 : 6;
  const size_t tw = fp32 ? 8 : 16;
  size_t b = align256(12 * (size_t)ytab_offset(YT_COUNT, M));  // y tables, FP64 and FP32
  b += 2 * align256(8 * 5 * (size_t)my);                       // band copies
  b += align256(8 * nclass) + align256(8 * 5 * R * nclass) + align256(32 * nclass);
  b += align256(8 * 2 * R * nclass);                           // homogeneous solutions
  b += nst * align256(S * 16);                                 // state and histories
  b += nsp * (align256(Sp * es) + 2 * align256(Wp * es));      // transform buffers
  b += (with_z ? 2 : 1) * align256(P * 8);                     // physical RK histories
  b += 6 * align256(8 * (size_t)my);                           // 00-mode work
  b += 2 * align256(8 * 12 * (size_t)my) + 2 * align256(8 * 2 * (size_t)my);  // mean split
  b += align256(sizeof(DevScalars));
  b += align256(32 * (size_t)cfl_tiles_max(my, mgalz));           // CFL maxima per x-line tile
  b += align256(tw * mgalz) + align256(tw * mgalx) + align256(16 * (size_t)mgalz);  // twiddles
  return b;
}

static void* pinned_get(size_t bytes) {
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}
static void pinned_put(void* p, size_t) {
  if (p) hipHostFree(p);
}

// Pinned host staging for transfers of one complex field, grown on demand;
// copies go on the run's own stream, so they hold the runtime only while
// they are enqueued.
static int stage_host(loma_run* r, size_t ndoubles) {
  if (r->hstage_n >= ndoubles) return LOMA_OK;
  pinned_put(r->hstage, r->hstage_n * sizeof(double));
  r->hstage = nullptr;
  r->hstage_n = 0;
  r->hstage = (double*)pinned_get(ndoubles * sizeof(double));
  if (!r->hstage) { set_error("pinned host allocation failed"); return LOMA_ERR_OOM; }
  r->hstage_n = ndoubles;
  return LOMA_OK;
}

extern "C" {

const char* loma_version(void) { return "loma_amd 0.1 (gfx950)"; }
const char* loma_last_error(void) { return g_err.c_str(); }

int loma_device_count(int* n) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  if (n) *n = c;
  return LOMA_OK;
}

int loma_fd_matrices(int my, const double* y, double* dt11, double* dt12, double* dt21,
                     double* dt22, double* fmap) {
  if (my < 8 || !y) { set_error("loma_fd_matrices: bad arguments"); return LOMA_ERR_ARG; }
  FdOps op;
  int rc = fd_build(my, y, op);
  if (rc) { set_error("loma_fd_matrices: singular stencil system"); return rc; }
  // (nb,my) Fortran order == row-major [j][d]
  if (dt11) std::memcpy(dt11, op.a1.data(), sizeof(double) * 7 * my);
  if (dt12) std::memcpy(dt12, op.b1.data(), sizeof(double) * 7 * my);
  if (dt21) std::memcpy(dt21, op.a2.data(), sizeof(double) * 5 * my);
  if (dt22) std::memcpy(dt22, op.b2.data(), sizeof(double) * 5 * my);
  if (fmap) std::memcpy(fmap, op.fmap.data(), sizeof(double) * my);
  return LOMA_OK;
}

}  // extern "C"

// ---- ensembles: one graph launch per step for many runs -----------------
// The steps of all runs are captured into one hipGraph whose branches (one
// per run, on the run's own stream) are independent, so the device overlaps
// the runs' kernels; graphs are cached by run set and CFL pattern.
struct EnsembleGraph {
  std::vector<loma_run*> runs;
  std::vector<int> icfl;
  hipStream_t master = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<hipEvent_t> ready;  // per run: work already queued on its stream
  void* desc = nullptr;           // device argument blocks of the batched launches
  std::vector<hipStream_t> lanes; // one capture stream per group of runs
  int cap = 0;                    // most runs per batched launch
};
// Batch shape of an ensemble whose runs all carry LOMA_FLAG_AUTOTUNE: its
// first steps alternate between one launch per group and at most two runs
// per launch, timed on the device, and the faster is kept.  Every run's
// result is the same either way, so the trial steps are ordinary steps.
struct EnsShape {
  std::vector<loma_run*> runs;
  int trial = 0;          // trial steps done
  float best[2] = {1e30f, 1e30f};
  int cap = 0;            // chosen (0 while trying)
};
static std::vector<EnsShape> g_shape;
constexpr int ENS_CAPS[2] = {1 << 20, 2};
constexpr int ENS_TRIALS = 4;   // alternating, two per shape
static std::mutex g_ens_mu;
static std::vector<EnsembleGraph> g_ens;

static void ensemble_forget(loma_run* r) {
  std::lock_guard<std::mutex> lk(g_ens_mu);
  for (size_t q = 0; q < g_ens.size();) {
    bool has = false;
    for (loma_run* x : g_ens[q].runs) has = has || x == r;
    if (has) {
      // a launch of the graph may still read the argument blocks
      if (g_ens[q].master) hipStreamSynchronize(g_ens[q].master);
      if (g_ens[q].exec) hipGraphExecDestroy(g_ens[q].exec);
      if (g_ens[q].master) hipStreamDestroy(g_ens[q].master);
      for (hipEvent_t e : g_ens[q].ready) hipEventDestroy(e);
      if (g_ens[q].desc) hipFree(g_ens[q].desc);
      for (hipStream_t l : g_ens[q].lanes) hipStreamDestroy(l);
      g_ens.erase(g_ens.begin() + q);
    } else {
      ++q;
    }
  }
  for (size_t q = 0; q < g_shape.size();) {
    bool has = false;
    for (loma_run* x : g_shape[q].runs) has = has || x == r;
    if (has) g_shape.erase(g_shape.begin() + q);
    else ++q;
  }
}

extern "C" {

int loma_free(loma_run* r) {
  if (!r) return LOMA_OK;
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  ensemble_forget(r);
  if (r->stream) hipStreamSynchronize(r->stream);
  for (int q = 0; q < 2; ++q) {
    if (r->gexec[q]) hipGraphExecDestroy(r->gexec[q]);
    if (r->sev[q]) hipEventDestroy(r->sev[q]);
  }
  pinned_put(r->hsc, sizeof(DevScalars));
  pinned_put(r->hstage, r->hstage_n * sizeof(double));
  if (r->arena) hipFree(r->arena);
  if (r->snap) hipFree(r->snap);
  for (void* p : {(void*)r->st_acc, (void*)r->st_spe, (void*)r->st_part, (void*)r->st_rho,
                  (void*)r->st_last, (void*)r->st_jsp})
    if (p) hipFree(p);
  if (r->owns_stream && r->stream) hipStreamDestroy(r->stream);
  delete r;
  return LOMA_OK;
}

static int autotune_tiles(loma_run* r);

int loma_init(loma_run** out, int mgalx, int my, int mgalz, const double* y, double Re,
              double alp, double bet, double pr, double cfl, int ncfl, double dt0,
              int precision, int flags) {
  if (!out || !y) { set_error("loma_init: null argument"); return LOMA_ERR_ARG; }
  *out = nullptr;
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  if (mgalx < 8 || mgalz < 8 || my < 16 || my > 1024) {
    set_error("loma_init: grid out of range (my in [16,1024], mgalx,mgalz >= 8)");
    return LOMA_ERR_ARG;
  }
  if (precision != LOMA_FP64 && precision != LOMA_MIXED) {
    set_error("loma_init: unknown precision");
    return LOMA_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("loma_init: no GPU visible");
    return LOMA_ERR_NODEV;
  }
  loma_run* r = new loma_run();
  r->mgalx = mgalx; r->my = my; r->mgalz = mgalz;
  r->mx = 2 * (mgalx / 3);
  r->nplanes = r->mx / 2;
  r->mz = 2 * (mgalz / 3) - 1;
  r->nz = (r->mz - 1) / 2;
  r->Re = Re; r->alp = alp; r->bet = bet; r->pr = pr; r->cfl = cfl;
  r->ncfl = ncfl > 0 ? ncfl : 1; r->dt0 = dt0;
  r->precision = precision; r->flags = flags;
  r->fp32 = precision == LOMA_MIXED;
  r->with_z = (flags & LOMA_FLAG_SCALAR_Z) != 0;
  r->plz = make_plan(mgalz);
  r->plx = make_plan(mgalx);
  if (r->plz.npass < 0 || r->plx.npass < 0 || (mgalx & 1)) {
    set_error("loma_init: mgalx/mgalz must factor into 2,3,5 (mgalx even)");
    delete r;
    return LOMA_ERR_ARG;
  }
  r->M = pick_M(my);
  if (r->M < 0) { set_error("loma_init: my too large"); delete r; return LOMA_ERR_ARG; }
  int rc = fd_build(my, y, r->fd);
  if (rc) { set_error("loma_init: FD setup failed"); delete r; return rc; }
  r->y.assign(y, y + my);
  // the run's own non-blocking stream carries its set-up as well as its steps
  if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("loma_init: stream creation failed"); delete r; return LOMA_ERR_DEVICE;
  }
  r->owns_stream = true;
  r->arena_size = plan_bytes(mgalx, my, mgalz, r->fp32, r->with_z);
  {
    const hipError_t e = hipMalloc(&r->arena, r->arena_size);
    if (e != hipSuccess) {
      r->arena = nullptr;
      set_error(std::string("loma_init: hipMalloc failed: ") + hipGetErrorString(e));
      loma_free(r);
      return LOMA_ERR_OOM;
    }
  }
  hipStream_t sst = r->stream;
  // run constants
  RunConst& c = r->rc;
  c.mgalx = mgalx; c.my = my; c.mgalz = mgalz; c.mx = r->mx; c.mz = r->mz;
  c.nplanes = r->nplanes; c.nz = r->nz;
  c.myp = (my + 15) / 16 * 16;  // 128-byte aligned y runs in the transform buffers
  c.alp = alp; c.bet = bet; c.ire = 1.0 / Re; c.peclet = Re * pr;
  c.hxalp = alp * (mgalx - 1) / M_PI;  // cross.f:1500
  c.hzbet = bet * (mgalz - 1) / M_PI;
  c.cfl = cfl;
  c.dt_fp32 = (flags & LOMA_FLAG_DT_FP32) ? 1 : 0;

  // ---- per-row tables, lane-major over 64*M rows, zero padded ----
  const int R = 64 * r->M;
  const int M = r->M;
  const FdOps& f = r->fd;
  // boundary-derivative functionals: w0[j] = d/dy(e_j) at row 0, w1[j] at
  // row my-1 (fmap included), used by the Neumann solve
  std::vector<double> wbd((size_t)2 * my, 0.0);
  {
    std::vector<double> e(my), d(my);
    for (int j = 0; j < my; ++j) {
      std::fill(e.begin(), e.end(), 0.0);
      e[j] = 1.0;
      band_apply(f.b1.data(), my, 3, e.data(), d.data());
      band_lu_solve(f.inv1.data(), f.up1.data(), f.lo1.data(), my, 3, d.data());
      wbd[(size_t)j * 2 + 0] = d[0] * f.fmap[0];
      wbd[(size_t)j * 2 + 1] = d[my - 1] * f.fmap[my - 1];
    }
  }
  // per-row tables and scan compositions at ytab_offset(id, M), the layout
  // the kernels address with constant offsets
  const size_t tot = ytab_offset(YT_COUNT, M);
  std::vector<double> h(tot, 0.0);
  auto lm = [&](int row, int w, int idx) -> size_t {
    const int l = row / M, t = row - l * M;
    return (size_t)(t * w + idx) * 64 + l;
  };
  // row-indexed table -> lane-major
  auto put = [&](int id, const std::vector<double>& src) {
    const int w = ytab_width(id);
    const size_t o = ytab_offset(id, M);
    for (int j = 0; j < my; ++j)
      for (int d = 0; d < w; ++d) h[o + lm(j, w, d)] = src[(size_t)j * w + d];
  };
  // lower multipliers: lo[k][i-1] (row k into row k+i) -> destination row k+i
  auto put_lo = [&](int id, const std::vector<double>& src) {
    const int w = ytab_width(id);
    const size_t o = ytab_offset(id, M);
    for (int j = 0; j < my; ++j)
      for (int i = 1; i <= w; ++i)
        if (j - i >= 0) h[o + lm(j, w, i - 1)] = src[(size_t)(j - i) * w + i - 1];
  };
  put(YT_B1, f.b1); put(YT_A1, f.a1); put(YT_INV1, f.inv1); put(YT_UP1, f.up1);
  put_lo(YT_LO1, f.lo1); put(YT_FMAP, f.fmap); put(YT_B2, f.b2); put(YT_A2, f.a2);
  put(YT_INV2, f.inv2); put(YT_UP2, f.up2); put_lo(YT_LO2, f.lo2); put(YT_INVI, f.invi);
  put(YT_UPI, f.upi); put_lo(YT_LOI, f.loi); put(YT_WBD, wbd);
  auto H = [&](int id) { return h.data() + ytab_offset(id, M); };
  int lev[6];
  lev[0] = scan_table(my, M, 3, false, f.lo1.data(), nullptr, nullptr, H(YT_HL1));
  lev[1] = scan_table(my, M, 3, true, nullptr, f.up1.data(), f.inv1.data(), H(YT_HU1));
  lev[2] = scan_table(my, M, 2, false, f.lo2.data(), nullptr, nullptr, H(YT_HL2));
  lev[3] = scan_table(my, M, 2, true, nullptr, f.up2.data(), f.inv2.data(), H(YT_HU2));
  lev[4] = scan_table(my, M, 3, false, f.loi.data(), nullptr, nullptr, H(YT_HLI));
  lev[5] = scan_table(my, M, 3, true, nullptr, f.upi.data(), f.invi.data(), H(YT_HUI));
  // the FP64 tables, then their FP32 rounding in the same layout (the
  // mixed form's FP32 y-operators read it)
  std::vector<float> h32(h.begin(), h.end());
  if ((rc = dalloc(r, &r->tab, tot * (sizeof(double) + sizeof(float))))) {
    loma_free(r); return rc;
  }
  if (hipMemcpyAsync(r->tab, h.data(), tot * sizeof(double), hipMemcpyHostToDevice, sst) !=
          hipSuccess ||
      hipMemcpyAsync(r->tab + tot, h32.data(), tot * sizeof(float), hipMemcpyHostToDevice,
                     sst) != hipSuccess ||
      hipStreamSynchronize(sst) != hipSuccess) {
    set_error("loma_init: table upload failed"); loma_free(r); return LOMA_ERR_DEVICE;
  }
  YTab& T = r->T;
  T.base = r->tab;
  T.base32 = reinterpret_cast<const float*>(r->tab + tot);
  T.kL1 = lev[0]; T.kU1 = lev[1]; T.kL2 = lev[2]; T.kU2 = lev[3]; T.kLi = lev[4];
  T.kUi = lev[5];
  std::memcpy(r->scan_levels, lev, sizeof(lev));
  T.my = my;
  // plain band copies for the Helmholtz factor kernel
  if ((rc = dalloc(r, &r->dt21, sizeof(double) * 5 * my))) { loma_free(r); return rc; }
  if ((rc = dalloc(r, &r->dt22, sizeof(double) * 5 * my))) { loma_free(r); return rc; }
  hipMemcpyAsync(r->dt21, f.a2.data(), sizeof(double) * 5 * my, hipMemcpyHostToDevice, sst);
  hipMemcpyAsync(r->dt22, f.b2.data(), sizeof(double) * 5 * my, hipMemcpyHostToDevice, sst);
  // ---- Helmholtz LU per (kx, |kz|) class: rK = kx^2 + kz^2 ----
  r->nclass = r->nplanes * (r->nz + 1);
  r->lu_stride = 5 * R;
  {
    std::vector<double> shifts(r->nclass);
    for (int i = 0; i < r->nplanes; ++i)
      for (int ka = 0; ka <= r->nz; ++ka) {
        const double kx = alp * i, kz = bet * ka;
        shifts[i * (r->nz + 1) + ka] = kx * kx + kz * kz;
      }
    double* dsh = nullptr;
    if ((rc = dalloc(r, &dsh, sizeof(double) * r->nclass))) { loma_free(r); return rc; }
    hipMemcpyAsync(dsh, shifts.data(), sizeof(double) * r->nclass, hipMemcpyHostToDevice, sst);
    if ((rc = dalloc(r, &r->helm_lu, sizeof(double) * (size_t)r->lu_stride * r->nclass))) {
      loma_free(r); return rc;
    }
    if ((rc = dalloc(r, &r->helm_bc, sizeof(double) * 4 * (size_t)r->nclass))) {
      loma_free(r); return rc;
    }
    if ((rc = dalloc(r, &r->helm_h, sizeof(double) * 2 * R * (size_t)r->nclass))) {
      loma_free(r); return rc;
    }
    if (launch_helm_factor(r->dt21, r->dt22, my, M, dsh, r->nclass, r->helm_lu, sst) ||
        launch_helm_bc(M, r->T, r->helm_lu, r->helm_bc, r->helm_h, r->nclass, sst) ||
        hipStreamSynchronize(sst) != hipSuccess) {
      set_error("loma_init: factor launch failed"); loma_free(r); return LOMA_ERR_DEVICE;
    }
  }
  // ---- state and work buffers ----
  const size_t S = (size_t)r->nplanes * r->mz * my;
  const size_t P = (size_t)mgalx * my * mgalz;
  // transform buffers use the padded y stride myp
  const size_t Sp = (size_t)r->nplanes * r->mz * c.myp;
  const size_t Wp = (size_t)r->nplanes * mgalz * c.myp;
  const size_t es = r->fp32 ? 8 : 16;
  double2** st[6] = {&r->vor, &r->phi, &r->psi, &r->scal, &r->ten12, &r->ten13};
  for (auto p : st)
    if ((rc = dalloc(r, p, S * 16))) { loma_free(r); return rc; }
  if (r->with_z && (rc = dalloc(r, &r->zf, S * 16))) { loma_free(r); return rc; }
  const int nsp = r->with_z ? 12 : 11;
  for (int q = 0; q < nsp; ++q) {
    if ((rc = dalloc(r, &r->spec[q], Sp * es))) { loma_free(r); return rc; }
    if ((rc = dalloc(r, &r->wbuf[q], Wp * es))) { loma_free(r); return rc; }
    if ((rc = dalloc(r, &r->wout[q], Wp * es))) { loma_free(r); return rc; }
  }
  if ((rc = dalloc(r, &r->rhst, P * 8))) { loma_free(r); return rc; }
  if (r->with_z && (rc = dalloc(r, &r->rhsz, P * 8))) { loma_free(r); return rc; }
  double** v00[6] = {&r->u00wk, &r->v00wk, &r->w00wk, &r->rf0u, &r->rf0w, &r->hy};
  for (auto p : v00)
    if ((rc = dalloc(r, p, sizeof(double) * my))) { loma_free(r); return rc; }
  {
    double** mb[4] = {&r->mean_in, &r->mean_out, &r->rbar, &r->rzbar};
    const size_t nmb[4] = {(size_t)12 * my, (size_t)12 * my, (size_t)2 * my, (size_t)2 * my};
    for (int q = 0; q < 4; ++q) {
      if ((rc = dalloc(r, mb[q], sizeof(double) * nmb[q]))) { loma_free(r); return rc; }
      if (hipMemsetAsync(*mb[q], 0, sizeof(double) * nmb[q], sst) != hipSuccess) {
        set_error("loma_init: memset failed"); loma_free(r); return LOMA_ERR_DEVICE;
      }
    }
  }
  {
    std::vector<double> hyv(my);
    for (int j = 1; j < my - 1; ++j) hyv[j] = (y[j + 1] - y[j - 1]) / 2.0;  // main.f:689-696
    hyv[0] = (y[1] - y[0]) / 2.0;
    hyv[my - 1] = (y[my - 1] - y[my - 2]) / 2.0;
    hipMemcpyAsync(r->hy, hyv.data(), sizeof(double) * my, hipMemcpyHostToDevice, sst);
    hipStreamSynchronize(sst);  // hyv is a local
  }
  if ((rc = dalloc(r, &r->sc, sizeof(DevScalars)))) { loma_free(r); return rc; }
  if ((rc = dalloc(r, &r->cflp, 32 * (size_t)cfl_tiles_max(my, mgalz)))) { loma_free(r); return rc; }
  // twiddles
  auto mk_tw = [&](int n, void** dst) -> int {
    std::vector<double> tw(2 * n);
    for (int t = 0; t < n; ++t) {
      tw[2 * t] = std::cos(2.0 * M_PI * t / n);
      tw[2 * t + 1] = -std::sin(2.0 * M_PI * t / n);
    }
    if (r->fp32) {
      std::vector<float> tf(2 * n);
      for (int q = 0; q < 2 * n; ++q) tf[q] = (float)tw[q];
      int e = dalloc(r, dst, sizeof(float) * 2 * n);
      if (e) return e;
      hipMemcpyAsync(*dst, tf.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, sst);
      hipStreamSynchronize(sst);
    } else {
      int e = dalloc(r, dst, sizeof(double) * 2 * n);
      if (e) return e;
      hipMemcpyAsync(*dst, tw.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, sst);
      hipStreamSynchronize(sst);
    }
    return 0;
  };
  if ((rc = mk_tw(mgalz, &r->twz)) || (rc = mk_tw(mgalx, &r->twx))) { loma_free(r); return rc; }
  {
    std::vector<double> tw(2 * mgalz);
    for (int t = 0; t < mgalz; ++t) {
      tw[2 * t] = std::cos(2.0 * M_PI * t / mgalz);
      tw[2 * t + 1] = -std::sin(2.0 * M_PI * t / mgalz);
    }
    if ((rc = dalloc(r, &r->twz64, sizeof(double) * 2 * mgalz))) { loma_free(r); return rc; }
    hipMemcpyAsync(r->twz64, tw.data(), sizeof(double) * 2 * mgalz, hipMemcpyHostToDevice, sst);
    hipStreamSynchronize(sst);
  }
  // tile sizes: keep dynamic LDS per workgroup <= 64 KiB
  {
    const int e2 = r->fp32 ? 8 : 16;
    // z transforms: J y entries per workgroup, J * mgalz <= 256 * E
    // (E = LOMA_ZF_E32 in FP32, LOMA_ZF_E64 in FP64): one LDS buffer
    const int capz = 256 * (r->fp32 ? LOMA_ZF_E32 : LOMA_ZF_E64);
    int J = 32;
    while (J > 1 && J * mgalz > capz) J /= 2;
    if (J * mgalz > capz) {
      set_error("loma_init: mgalz too large for the z-transform kernel");
      loma_free(r);
      return LOMA_ERR_ARG;
    }
    r->J = J;
    // x lines: 2*pairs lines per 256-thread workgroup, pairs*mgalx entries
    // at most 256*EPT (EPT = 4 in FP64, 8 in FP32): two LDS buffers of
    // 32 KiB together, 64 KiB in FP64 where each transform carries two
    // fields (MI355X allows 160 KiB per workgroup)
    const int cap = r->fp32 ? 2048 : 1024;
    int pairs = 16;  // a power of two: each thread serves one line pair
    while (pairs > 1 && pairs * mgalx > cap) pairs /= 2;
    if (pairs * mgalx > cap) {
      set_error("loma_init: mgalx too large for the x-line kernel");
      loma_free(r);
      return LOMA_ERR_ARG;
    }
    r->L = 2 * pairs;
  }
  if (hipStreamSynchronize(sst) != hipSuccess) {
    set_error("loma_init: device error during set-up"); loma_free(r); return LOMA_ERR_DEVICE;
  }
  if (flags & LOMA_FLAG_AUTOTUNE) {
    if ((rc = autotune_tiles(r))) { loma_free(r); return rc; }
  }
  *out = r;
  return LOMA_OK;
}

int loma_dims(const loma_run* r, long long* out) {
  if (!r || !out) return LOMA_ERR_ARG;
  out[0] = r->mx; out[1] = r->mz; out[2] = r->nplanes; out[3] = r->nz;
  out[4] = (long long)r->nplanes * r->mz * r->my;
  out[5] = (long long)r->nplanes * r->mgalz * r->my;
  out[6] = (long long)r->mgalx * r->my * r->mgalz;
  out[7] = (long long)r->bytes;
  return LOMA_OK;
}

int loma_set_stream(loma_run* r, void* stream) {
  if (!r) return LOMA_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  if (r->owns_stream && r->stream && r->stream != (hipStream_t)stream) {
    hipStreamSynchronize(r->stream);
    hipStreamDestroy(r->stream);
  }
  r->owns_stream = false;
  r->stream = (hipStream_t)stream;
  for (int q = 0; q < 2; ++q)
    if (r->gexec[q]) { hipGraphExecDestroy(r->gexec[q]); r->gexec[q] = nullptr; }
  return LOMA_OK;
}

// host complex field (interleaved float or double pairs) <-> doubles
static void to_doubles(const void* src, int fp64, size_t n2, double* dst) {
  if (fp64) {
    std::memcpy(dst, src, sizeof(double) * n2);
  } else {
    const float* f = (const float*)src;
    for (size_t q = 0; q < n2; ++q) dst[q] = f[q];
  }
}
static void from_doubles(const double* src, int fp64, size_t n2, void* dst) {
  if (fp64) {
    std::memcpy(dst, src, sizeof(double) * n2);
  } else {
    float* f = (float*)dst;
    for (size_t q = 0; q < n2; ++q) f[q] = (float)src[q];
  }
}
// One complex field host -> device (and back) through the pinned staging,
// on the run's stream.  The staging is reused field after field.
static int put_field(loma_run* r, const void* src, int fp64, size_t S, double2* dst) {
  to_doubles(src, fp64, 2 * S, r->hstage);
  HIPCHK(hipMemcpyAsync(dst, r->hstage, S * 16, hipMemcpyHostToDevice, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  return LOMA_OK;
}
static int get_field(loma_run* r, const double2* src, size_t S, int fp64, void* dst) {
  HIPCHK(hipMemcpyAsync(r->hstage, src, S * 16, hipMemcpyDeviceToHost, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  from_doubles(r->hstage, fp64, 2 * S, dst);
  return LOMA_OK;
}

// Shared end of loma_upload and loma_seed, once the spectral fields are on
// the device: history and work buffers zeroed, 00 modes and scalars set,
// the step interface's events made, captured step graphs dropped.
static int upload_tail(loma_run* r, const double* u00, const double* v00, const double* w00) {
  const size_t S = (size_t)r->nplanes * r->mz * r->my;
  const int my = r->my;
  // history and work buffers start at zero (main.f:162-176)
  hipStream_t st = r->stream;
  const size_t P8 = sizeof(double) * (size_t)r->mgalx * my * r->mgalz;
  HIPCHK(hipMemsetAsync(r->ten12, 0, S * 16, st));
  HIPCHK(hipMemsetAsync(r->ten13, 0, S * 16, st));
  HIPCHK(hipMemsetAsync(r->rhst, 0, P8, st));
  if (r->rhsz) HIPCHK(hipMemsetAsync(r->rhsz, 0, P8, st));
  HIPCHK(hipMemsetAsync(r->rf0u, 0, sizeof(double) * my, st));
  HIPCHK(hipMemsetAsync(r->rf0w, 0, sizeof(double) * my, st));
  HIPCHK(hipMemsetAsync(r->rbar, 0, sizeof(double) * 2 * my, st));
  HIPCHK(hipMemsetAsync(r->rzbar, 0, sizeof(double) * 2 * my, st));
  std::memcpy(r->hstage, u00, sizeof(double) * my);
  std::memcpy(r->hstage + my, v00, sizeof(double) * my);
  std::memcpy(r->hstage + 2 * my, w00, sizeof(double) * my);
  HIPCHK(hipMemcpyAsync(r->u00wk, r->hstage, sizeof(double) * my, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(r->v00wk, r->hstage + my, sizeof(double) * my, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(r->w00wk, r->hstage + 2 * my, sizeof(double) * my, hipMemcpyHostToDevice,
                        st));
  DevScalars s{};
  s.dt = (r->flags & LOMA_FLAG_DT_FP32) ? (double)(float)r->dt0 : r->dt0;
  s.time = 0.0;
  HIPCHK(hipMemcpyAsync(r->sc, &s, sizeof(s), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  // events and pinned scalars of the one-step-in-flight interface, made here
  // (under the set-up lock) rather than in the stepping loop
  if (!r->sev[0]) {
    HIPCHK(hipEventCreate(&r->sev[0]));
    HIPCHK(hipEventCreate(&r->sev[1]));
  }
  if (!r->hsc) {
    r->hsc = (DevScalars*)pinned_get(sizeof(DevScalars));
    if (!r->hsc) { set_error("loma_upload: pinned host allocation failed"); return LOMA_ERR_OOM; }
  }
  r->inflight = false;
  r->istep = 0;
  r->uploaded = true;
  // the step graphs and every ensemble graph with this run hold its
  // constants (boundary densities and scalar values) by value
  for (int q = 0; q < 2; ++q)
    if (r->gexec[q]) { hipGraphExecDestroy(r->gexec[q]); r->gexec[q] = nullptr; }
  ensemble_forget(r);
  return LOMA_OK;
}

int loma_upload(loma_run* r, const void* vor, const void* phi, const void* psi,
                const void* scal, const double* u00, const double* v00, const double* w00,
                int fp64) {
  if (!r || !vor || !phi || !psi || !scal || !u00 || !v00 || !w00) {
    set_error("loma_upload: null argument");
    return LOMA_ERR_ARG;
  }
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  const size_t S = (size_t)r->nplanes * r->mz * r->my;
  const int my = r->my;
  int e = stage_host(r, std::max(2 * S, (size_t)3 * my));
  if (e) return e;
  HIPCHK(hipStreamSynchronize(r->stream));  // nothing of the run may still read the state
  // boundary densities from the 00 mode of T (cross.f:208-213)
  {
    const double T0 = fp64 ? ((const double*)scal)[0] : ((const float*)scal)[0];
    const double T1 = fp64 ? ((const double*)scal)[2 * (my - 1)]
                           : ((const float*)scal)[2 * (my - 1)];
    if (!(T0 > 0.0) || !(T1 > 0.0)) {
      set_error("loma_upload: 00 mode of T must be positive at both ends");
      return LOMA_ERR_ARG;
    }
    r->rc.rhobot = 1.0 / T0;
    r->rc.rhotop = 1.0 / T1;
  }
  const void* src[4] = {vor, phi, psi, scal};
  double2* dst[4] = {r->vor, r->phi, r->psi, r->scal};
  for (int q = 0; q < 4; ++q)
    if ((e = put_field(r, src[q], fp64, S, dst[q]))) return e;
  if (r->zf) {
    // Z starts at zero until loma_upload_z sets it, never stale arena memory
    HIPCHK(hipMemsetAsync(r->zf, 0, S * 16, r->stream));
    r->rc.zbot = r->rc.ztop = 0.0;
  }
  return upload_tail(r, u00, v00, w00);
}

int loma_seed(loma_run* r, const double* coef, int nymax, double s, double amp,
              double env_width, double k0) {
  if (!r || !coef) { set_error("loma_seed: null argument"); return LOMA_ERR_ARG; }
  if (nymax < 1 || nymax > 64 || !(s > 0.0) || !(env_width > 0.0) || !(k0 > 0.0) ||
      !(amp >= 0.0)) {
    set_error("loma_seed: nymax in [1,64], s, env_width, k0 > 0 and amp >= 0 required");
    return LOMA_ERR_ARG;
  }
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  const int my = r->my;
  const size_t nmodes = (size_t)r->nplanes * r->mz;
  const size_t es = r->fp32 ? 8 : 16;
  const size_t wbytes = (size_t)r->nplanes * r->mgalz * r->rc.myp * es;
  const size_t cbytes = 3 * nmodes * nymax * 16;
  // the coefficients, the per-mode energies and y go to transform work
  // buffers, which the next step overwrites anyway
  if (cbytes > wbytes || nmodes * 3 * 8 > wbytes || (size_t)my * 8 > wbytes) {
    set_error("loma_seed: transform work buffer too small for the draws (my >= 48 always fits)");
    return LOMA_ERR_ARG;
  }
  int e = stage_host(r, std::max({cbytes / 8, nmodes * 3, (size_t)3 * my}));
  if (e) return e;
  HIPCHK(hipStreamSynchronize(r->stream));  // nothing of the run may still read the state
  hipStream_t st = r->stream;
  double2* dcoef = (double2*)r->wbuf[0];
  double* dpart = (double*)r->wbuf[1];
  double* dy = (double*)r->wbuf[2];
  std::memcpy(r->hstage, coef, cbytes);
  HIPCHK(hipMemcpyAsync(dcoef, r->hstage, cbytes, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  std::memcpy(r->hstage, r->y.data(), sizeof(double) * my);
  HIPCHK(hipMemcpyAsync(dy, r->hstage, sizeof(double) * my, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  IcArgs a{};
  a.rc = r->rc;
  a.T = r->T;
  a.helm_lu = r->helm_lu;
  a.lu_stride = r->lu_stride;
  a.coef = dcoef;
  a.y = dy;
  a.partial = dpart;
  a.vor = r->vor; a.phi = r->phi; a.psi = r->psi; a.scal = r->scal; a.zf = r->zf;
  a.s = s; a.env_width = env_width; a.k0 = k0;
  a.scale[0] = a.scale[1] = a.scale[2] = 1.0;
  a.nymax = nymax;
  a.with_z = r->with_z ? 1 : 0;
  // pass 1: centre-plane energy of each mode, summed here in mode order
  a.pass = 1;
  LCHK(launch_ic(r->M, a, st));
  HIPCHK(hipMemcpyAsync(r->hstage, dpart, nmodes * 3 * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int c = 0; c < 3; ++c) {
    double sum = 0.0;
    for (size_t w = 0; w < nmodes; ++w) sum += r->hstage[w * 3 + c];
    const double rms = std::sqrt(sum);
    a.scale[c] = rms > 0.0 ? amp / rms : 1.0;
  }
  // pass 2: the state
  a.pass = 2;
  LCHK(launch_ic(r->M, a, st));
  HIPCHK(hipStreamSynchronize(st));
  // mean profiles: rho00 = 1 + (s-1)/(s+1) tanh(-2y), U = tanh(2y)/2,
  // Z = (1 + tanh(2y))/2 (initloma3 main.f:360-373); boundary densities from
  // the 00 mode of T = 1/rho00 as loma_upload takes them (cross.f:208-213)
  std::vector<double> u00(my), zero(my, 0.0);
  const double* yy = r->y.data();
  auto rho = [&](double y) { return 1.0 + (s - 1.0) / (s + 1.0) * std::tanh(-2.0 * y); };
  for (int j = 0; j < my; ++j) u00[j] = rho(yy[j]) * (0.5 * std::tanh(2.0 * yy[j]));
  r->rc.rhobot = 1.0 / (1.0 / rho(yy[0]));
  r->rc.rhotop = 1.0 / (1.0 / rho(yy[my - 1]));
  if (r->with_z) {
    r->rc.zbot = 0.5 * (1.0 + std::tanh(2.0 * yy[0]));
    r->rc.ztop = 0.5 * (1.0 + std::tanh(2.0 * yy[my - 1]));
  }
  return upload_tail(r, u00.data(), zero.data(), zero.data());
}

int loma_upload_z(loma_run* r, const void* z, int fp64) {
  if (!r || !z || !r->with_z) { set_error("loma_upload_z: run has no scalar Z"); return LOMA_ERR_ARG; }
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  const size_t S = (size_t)r->nplanes * r->mz * r->my;
  int e = stage_host(r, 2 * S);
  if (e) return e;
  HIPCHK(hipStreamSynchronize(r->stream));
  r->rc.zbot = fp64 ? ((const double*)z)[0] : ((const float*)z)[0];
  r->rc.ztop = fp64 ? ((const double*)z)[2 * (r->my - 1)] : ((const float*)z)[2 * (r->my - 1)];
  if ((e = put_field(r, z, fp64, S, r->zf))) return e;
  for (int q = 0; q < 2; ++q)
    if (r->gexec[q]) { hipGraphExecDestroy(r->gexec[q]); r->gexec[q] = nullptr; }
  ensemble_forget(r);
  return LOMA_OK;
}

int loma_download(loma_run* r, void* vor, void* phi, void* psi, void* scal, double* u00,
                  double* v00, double* w00, int fp64) {
  if (!r || !r->uploaded) { set_error("loma_download: nothing uploaded"); return LOMA_ERR_STATE; }
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  const size_t S = (size_t)r->nplanes * r->mz * r->my;
  const int my = r->my;
  int e = stage_host(r, std::max(2 * S, (size_t)3 * my));
  if (e) return e;
  HIPCHK(hipStreamSynchronize(r->stream));
  void* dst[4] = {vor, phi, psi, scal};
  const double2* src[4] = {r->vor, r->phi, r->psi, r->scal};
  for (int q = 0; q < 4; ++q)
    if (dst[q] && (e = get_field(r, src[q], S, fp64, dst[q]))) return e;
  hipStream_t st = r->stream;
  HIPCHK(hipMemcpyAsync(r->hstage, r->u00wk, sizeof(double) * my, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(r->hstage + my, r->v00wk, sizeof(double) * my, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(r->hstage + 2 * my, r->w00wk, sizeof(double) * my, hipMemcpyDeviceToHost,
                        st));
  HIPCHK(hipStreamSynchronize(st));
  if (u00) std::memcpy(u00, r->hstage, sizeof(double) * my);
  if (v00) std::memcpy(v00, r->hstage + my, sizeof(double) * my);
  if (w00) std::memcpy(w00, r->hstage + 2 * my, sizeof(double) * my);
  return LOMA_OK;
}

int loma_download_z(loma_run* r, void* z, int fp64) {
  if (!r || !r->with_z || !z) { set_error("loma_download_z: run has no scalar Z"); return LOMA_ERR_ARG; }
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  const size_t S = (size_t)r->nplanes * r->mz * r->my;
  int e = stage_host(r, 2 * S);
  if (e) return e;
  HIPCHK(hipStreamSynchronize(r->stream));
  return get_field(r, r->zf, S, fp64, z);
}

static PhaseArgs phase_args(loma_run* r, int rkstep) {
  PhaseArgs a{};
  a.rc = r->rc;
  a.T = r->T;
  a.helm_lu = r->helm_lu;
  a.helm_bc = r->helm_bc;
  a.helm_h = r->helm_h;
  a.lu_stride = r->lu_stride;
  a.vor = r->vor; a.phi = r->phi; a.psi = r->psi; a.scal = r->scal;
  a.ten12 = r->ten12; a.ten13 = r->ten13; a.zf = r->zf;
  for (int q = 0; q < 12; ++q) a.spec[q] = r->spec[q];
  a.u00wk = r->u00wk; a.v00wk = r->v00wk; a.w00wk = r->w00wk;
  a.rf0u = r->rf0u; a.rf0w = r->rf0w;
  a.mean_in = r->mean_in; a.mean_out = r->mean_out;
  a.hy = r->hy;
  a.sc = r->sc;
  a.rkstep = rkstep;
  a.with_z = r->with_z ? 1 : 0;
  a.mode_base = 0;
  a.nmodes = r->nplanes * r->mz;
  return a;
}

static FftArgs fft_args(loma_run* r, int rkstep, int do_cfl) {
  FftArgs a{};
  a.rc = r->rc;
  a.plz = r->plz;
  a.plx = r->plx;
  a.twz = r->twz;
  a.twx = r->twx;
  for (int q = 0; q < 12; ++q) {
    a.spec[q] = r->spec[q];
    a.wbuf[q] = r->wbuf[q];
    a.wout[q] = r->wout[q];
  }
  a.rhst = r->rhst;
  a.rhsz = r->rhsz;
  a.mean_in = r->mean_in;
  a.mean_out = r->mean_out;
  a.rbar = r->rbar;
  a.rzbar = r->rzbar;
  a.hy = r->hy;
  a.sc = r->sc;
  a.cflp = r->cflp;
  a.rkstep = rkstep;
  a.do_cfl = do_cfl;
  a.with_z = r->with_z ? 1 : 0;
  a.J = r->J;
  a.L = r->L;
  return a;
}

#define PROF(cls, call)                      \
  do {                                       \
    hipEvent_t ev_ = nullptr;                \
    int pe_ = prof_begin(r, cls, &ev_);      \
    if (pe_) return pe_;                     \
    LCHK(call);                              \
    pe_ = prof_end(r, cls, ev_);             \
    if (pe_) return pe_;                     \
  } while (0)

static StatsArgs stats_args(loma_run* r) {
  StatsArgs a{};
  a.rc = r->rc;
  a.T = r->T;
  for (int q = 0; q < 12; ++q) a.spec[q] = r->spec[q];
  a.mean_in = r->mean_in;
  a.mean_out = r->mean_out;
  a.jsp = r->st_jsp;
  a.nsp = r->st_nsp;
  a.kach = STAT_KACH;
  a.nch = (r->nz + STAT_KACH) / STAT_KACH;
  a.nw = r->nplanes * a.nch;
  a.acc = r->st_acc;
  a.spe = r->st_spe;
  a.part = r->st_part;
  a.rho = r->st_rho;
  a.last = r->st_last;
  return a;
}

static int enqueue_substep(loma_run* r, int rkstep, int icfl) {
  hipStream_t st = r->stream;
  // the statistics window accumulates at rkstep 1 (cross.f:579, 795, 1586)
  const bool stats = r->st_on && rkstep == 1;
  PhaseArgs pa = phase_args(r, rkstep);
  PROF(KC_PHASEA, launch_phaseA(r->M, r->fp32, pa, st));
  if (stats) LCHK(launch_stats(r->M, r->fp32, 1, stats_args(r), st));
  FftArgs fa = fft_args(r, rkstep, (rkstep == 1 && icfl) ? 1 : 0);
  fa.stat_rho = stats ? r->st_rho : nullptr;
  fa.nfields = r->with_z ? 12 : 8;
  PROF(KC_ZINV, launch_zfft_inv(r->fp32, fa, st));
  PROF(KC_XLINE, launch_xline(r->fp32, fa, st));
  fa.nfields = r->with_z ? 12 : 11;
  PROF(KC_ZFWD, launch_zfft_fwd(r->fp32, fa, st));
  PROF(KC_PHASEC, launch_phaseC(r->M, r->fp32, pa, st));
  PROF(KC_PHASEC2, launch_phaseC2(r->M, r->fp32, pa, st));
  if (stats) {
    const StatsArgs sa = stats_args(r);
    LCHK(launch_stats(r->M, r->fp32, 2, sa, st));
    LCHK(launch_stats(r->M, r->fp32, 3, sa, st));
    r->st_nacum++;
  }
  if ((r->flags & LOMA_FLAG_HISTORY) && rkstep != 2)
    LCHK(launch_thickness(r->M, r->fp32, pa, r->hy, r->y[r->my - 1] - r->y[0],
                         rkstep == 1 ? 1 : 3, st));
  return LOMA_OK;
}

// ---- tile autotuner ------------------------------------------------------
// Times the transform kernels for the admissible tile shapes of a run's grid
// (y entries per z-transform workgroup J, line pairs per x-line workgroup)
// and keeps the fastest; results are cached per (grid, precision, Z).  Runs
// on the run's own (not yet uploaded) work buffers.
struct TileKey {
  int mgalx, my, mgalz, fp32, z;
  bool operator==(const TileKey& o) const {
    return mgalx == o.mgalx && my == o.my && mgalz == o.mgalz && fp32 == o.fp32 && z == o.z;
  }
};
static std::mutex g_tile_mu;
static std::vector<std::pair<TileKey, std::pair<int, int>>> g_tiles;

static int time_launch(loma_run* r, int which, float* best_ms) {
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  float best = 1e30f;
  for (int rep = 0; rep < 3; ++rep) {
    FftArgs fa = fft_args(r, 1, 0);
    HIPCHK(hipEventRecord(a, r->stream));
    int e = 0;
    if (which == 0) {
      fa.nfields = r->with_z ? 12 : 8;
      e = launch_zfft_inv(r->fp32, fa, r->stream);
      fa.nfields = r->with_z ? 12 : 11;
      if (!e) e = launch_zfft_fwd(r->fp32, fa, r->stream);
    } else {
      e = launch_xline(r->fp32, fa, r->stream);
    }
    if (e) { hipEventDestroy(a); hipEventDestroy(b); return LOMA_ERR_DEVICE; }
    HIPCHK(hipEventRecord(b, r->stream));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    if (rep > 0) best = std::min(best, ms);  // first launch warms up
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  *best_ms = best;
  return LOMA_OK;
}

static int autotune_tiles(loma_run* r) {
  const TileKey key{r->mgalx, r->my, r->mgalz, r->fp32 ? 1 : 0, r->with_z ? 1 : 0};
  {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    for (auto& e : g_tiles)
      if (e.first == key) { r->J = e.second.first; r->L = e.second.second; return LOMA_OK; }
  }
  const int J0 = r->J, L0 = r->L;
  int bestJ = J0, bestL = L0;
  float best = 1e30f, ms = 0.0f;
  for (int J = J0; J >= 1 && J >= J0 / 4; J /= 2) {
    r->J = J;
    int e = time_launch(r, 0, &ms);
    if (e) return e;
    if (ms < best) { best = ms; bestJ = J; }
  }
  r->J = bestJ;
  best = 1e30f;
  for (int L = L0; L >= 2 && L >= L0 / 4; L /= 2) {
    r->L = L;
    int e = time_launch(r, 1, &ms);
    if (e) return e;
    if (ms < best) { best = ms; bestL = L; }
  }
  r->L = bestL;
  std::lock_guard<std::mutex> lk(g_tile_mu);
  g_tiles.push_back({key, {bestJ, bestL}});
  return LOMA_OK;
}

static int enqueue_step(loma_run* r, int icfl) {
  for (int rk = 1; rk <= 3; ++rk) {
    int e = enqueue_substep(r, rk, icfl);
    if (e) return e;
  }
  PROF(KC_FINISH, launch_finish_step(r->sc, r->rc, icfl, r->cflp,
                                     xl_tiles(r->my, r->mgalz, r->L), r->stream));
  return LOMA_OK;
}

// ---- batched ensemble step ------------------------------------------------
// The runs of an ensemble are grouped by everything that selects a kernel
// instantiation: M and precision (phase kernels), mgalz and the z tile J, mgalx
// and the x tile L, and Z.  Each group takes one launch per kernel and
// substage; the launch's blocks are the runs' own grids one after another and
// each block reads its run's argument block from device memory, so a run's
// blocks compute exactly what its own launch computes.  Each group's step
// runs on a stream of its own inside the capture (forked from the ensemble's
// stream and joined back), so the groups' kernels overlap on the device as
// the runs' own kernels did.
struct EnsLaunch {
  enum Kind { PHASE, ZINV, XLINE, ZFWD, THICK, FINISH } kind;
  int which = 0;      // phase: 0 A, 1 C, 2 C2; thickness: 1 or 3
  int M = 0;
  bool fp32 = false;
  FftArgs proto{};    // host copy of the group's first run (lengths, tiles)
  size_t args_off = 0, first_off = 0;
  int nrun = 0, nblocks = 0;
  loma_run* run = nullptr;  // thickness
  PhaseArgs pa{};           // thickness
};
struct EnsPlan {
  std::vector<unsigned char> bytes;
  std::vector<std::vector<EnsLaunch>> lanes;  // one launch list per group
  size_t put(const void* p, size_t n) {
    const size_t off = (bytes.size() + 255) / 256 * 256;
    bytes.resize(off + n);
    std::memcpy(bytes.data() + off, p, n);
    return off;
  }
};

// LOMA_ENS_BATCH=n fixes the most runs per batched launch (0: not set)
static int ens_batch_env() {
  const char* e = std::getenv("LOMA_ENS_BATCH");
  const int v = e ? std::atoi(e) : 0;
  return v > 0 ? v : 0;
}

static void ensemble_plan(loma_run* const* runs, int nruns, const std::vector<int>& icfl,
                          int cap, EnsPlan& plan) {
  std::vector<std::array<int, 7>> keys;
  std::vector<std::vector<int>> groups;
  for (int q = 0; q < nruns; ++q) {
    loma_run* r = runs[q];
    const std::array<int, 7> k{r->M, r->fp32 ? 1 : 0, r->mgalz, r->J, r->mgalx, r->L,
                               r->with_z ? 1 : 0};
    size_t c = 0;
    while (c < keys.size() && keys[c] != k) ++c;
    if (c == keys.size()) { keys.push_back(k); groups.emplace_back(); }
    groups[c].push_back(q);
  }
  // at most `cap` runs per launch: a group larger than that is dealt round
  // robin to several lanes, which then overlap
  std::vector<std::vector<int>> split;
  for (const auto& g : groups) {
    const int nl = ((int)g.size() + cap - 1) / cap;
    for (int l = 0; l < nl; ++l) {
      split.emplace_back();
      for (size_t n = l; n < g.size(); n += nl) split.back().push_back(g[n]);
    }
  }
  groups.swap(split);
  for (const auto& g : groups) {
    std::vector<EnsLaunch> lane;
    loma_run* r0 = runs[g[0]];
    for (int rk = 1; rk <= 3; ++rk) {
      std::vector<PhaseArgs> pa;
      std::vector<FftArgs> fa;
      for (int q : g) {
        pa.push_back(phase_args(runs[q], rk));
        fa.push_back(fft_args(runs[q], rk, (rk == 1 && icfl[q]) ? 1 : 0));
      }
      const size_t pa_off = plan.put(pa.data(), pa.size() * sizeof(PhaseArgs));
      const size_t fa_off = plan.put(fa.data(), fa.size() * sizeof(FftArgs));
      auto launch = [&](EnsLaunch::Kind kind, int which) {
        std::vector<int> first;
        int nb = 0;
        for (size_t n = 0; n < g.size(); ++n) {
          first.push_back(nb);
          loma_run* r = runs[g[n]];
          nb += kind == EnsLaunch::PHASE ? phase_blocks(which, r->rc)
                : kind == EnsLaunch::XLINE ? xline_blocks(fa[n]) : zfft_blocks(fa[n]);
        }
        EnsLaunch L;
        L.kind = kind; L.which = which;
        L.M = r0->M; L.fp32 = r0->fp32;
        L.proto = fa[0];
        L.proto.nfields = kind == EnsLaunch::ZINV ? (r0->with_z ? 12 : 8) : (r0->with_z ? 12 : 11);
        L.args_off = kind == EnsLaunch::PHASE ? pa_off : fa_off;
        L.first_off = plan.put(first.data(), first.size() * sizeof(int));
        L.nrun = (int)g.size(); L.nblocks = nb;
        lane.push_back(L);
      };
      launch(EnsLaunch::PHASE, 0);
      launch(EnsLaunch::ZINV, 0);
      launch(EnsLaunch::XLINE, 0);
      launch(EnsLaunch::ZFWD, 0);
      launch(EnsLaunch::PHASE, 1);
      launch(EnsLaunch::PHASE, 2);
      if (rk != 2)
        for (size_t n = 0; n < g.size(); ++n)
          if (runs[g[n]]->flags & LOMA_FLAG_HISTORY) {
            EnsLaunch L;
            L.kind = EnsLaunch::THICK; L.which = rk == 1 ? 1 : 3;
            L.run = runs[g[n]]; L.pa = pa[n];
            lane.push_back(L);
          }
    }
    std::vector<FinishArgs> fin;
    for (int q : g) {
      loma_run* r = runs[q];
      fin.push_back(FinishArgs{r->sc, r->rc, r->cflp, icfl[q], xl_tiles(r->my, r->mgalz, r->L)});
    }
    EnsLaunch L;
    L.kind = EnsLaunch::FINISH;
    L.args_off = plan.put(fin.data(), fin.size() * sizeof(FinishArgs));
    L.nrun = (int)g.size();
    lane.push_back(L);
    plan.lanes.push_back(std::move(lane));
  }
}

static int enqueue_lane(hipStream_t st, const void* desc, const std::vector<EnsLaunch>& lane) {
  const char* d = (const char*)desc;
  for (const EnsLaunch& L : lane) {
    const int* first = (const int*)(d + L.first_off);
    switch (L.kind) {
      case EnsLaunch::PHASE:
        LCHK(launch_phase_batch(L.M, L.fp32, L.which, (const PhaseArgs*)(d + L.args_off), first,
                                L.nrun, L.nblocks, st));
        break;
      case EnsLaunch::ZINV:
      case EnsLaunch::ZFWD:
        LCHK(launch_zfft_batch(L.fp32, L.kind == EnsLaunch::ZINV, L.proto,
                               (const FftArgs*)(d + L.args_off), first, L.nrun, L.nblocks, st));
        break;
      case EnsLaunch::XLINE:
        LCHK(launch_xline_batch(L.fp32, L.proto, (const FftArgs*)(d + L.args_off), first, L.nrun,
                                L.nblocks, st));
        break;
      case EnsLaunch::THICK: {
        loma_run* r = L.run;
        LCHK(launch_thickness(r->M, r->fp32, L.pa, r->hy, r->y[r->my - 1] - r->y[0], L.which, st));
        break;
      }
      case EnsLaunch::FINISH:
        LCHK(launch_finish_batch((const FinishArgs*)(d + L.args_off), L.nrun, st));
        break;
    }
  }
  return LOMA_OK;
}

int loma_substep(loma_run* r, int rkstep) {
  if (!r || !r->uploaded) { set_error("loma_substep: nothing uploaded"); return LOMA_ERR_STATE; }
  if (rkstep < 1 || rkstep > 3) { set_error("loma_substep: rkstep in 1..3"); return LOMA_ERR_ARG; }
  const int icfl = ((r->istep) % r->ncfl == 0) ? 1 : 0;
  int e = enqueue_substep(r, rkstep, icfl);
  if (e) return e;
  HIPCHK(hipStreamSynchronize(r->stream));
  return LOMA_OK;
}

static int ensure_stream(loma_run* r);

// Capture one full step (with or without the CFL pass) into the run's graph.
static int build_graph(loma_run* r, int icfl) {
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  hipGraphExec_t& ge = r->gexec[icfl];
  if (ge) return LOMA_OK;
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
  int e = enqueue_step(r, icfl);
  hipError_t ce = hipStreamEndCapture(r->stream, &g);
  if (e) return e;
  HIPCHK(ce);
  hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  HIPCHK(ie);
  return LOMA_OK;
}

// The pieces of a run that carry the time advance from one step to the next
// (spectral state, RK histories, 00-mode work arrays, scalars); everything
// else is rebuilt within a step or constant.
static int snapshot_parts(loma_run* r, std::vector<std::pair<void*, size_t>>& parts) {
  const size_t S = (size_t)r->nplanes * r->mz * r->my;
  const size_t P8 = sizeof(double) * (size_t)r->mgalx * r->my * r->mgalz;
  const size_t m8 = sizeof(double) * (size_t)r->my;
  parts = {{r->vor, S * 16}, {r->phi, S * 16}, {r->psi, S * 16}, {r->scal, S * 16},
           {r->ten12, S * 16}, {r->ten13, S * 16}, {r->rhst, P8},
           {r->u00wk, m8}, {r->v00wk, m8}, {r->w00wk, m8}, {r->rf0u, m8}, {r->rf0w, m8},
           {r->rbar, 2 * m8}, {r->rzbar, 2 * m8}, {r->sc, sizeof(DevScalars)}};
  if (r->zf) parts.push_back({r->zf, S * 16});
  if (r->rhsz) parts.push_back({r->rhsz, P8});
  return LOMA_OK;
}

int loma_snapshot(loma_run* r) {
  if (!r || !r->uploaded) { set_error("loma_snapshot: nothing uploaded"); return LOMA_ERR_STATE; }
  if (r->inflight) { set_error("loma_snapshot: a step is in flight"); return LOMA_ERR_STATE; }
  std::vector<std::pair<void*, size_t>> parts;
  snapshot_parts(r, parts);
  size_t need = 0;
  for (auto& p : parts) need += align256(p.second);
  if (!r->snap) {
    if (hipMalloc(&r->snap, need) != hipSuccess) {
      r->snap = nullptr;
      set_error("loma_snapshot: device allocation failed");
      return LOMA_ERR_OOM;
    }
    r->snap_bytes = need;
  }
  char* d = (char*)r->snap;
  for (auto& p : parts) {
    HIPCHK(hipMemcpyAsync(d, p.first, p.second, hipMemcpyDeviceToDevice, r->stream));
    d += align256(p.second);
  }
  HIPCHK(hipStreamSynchronize(r->stream));
  r->snap_istep = r->istep;
  return LOMA_OK;
}

int loma_restore(loma_run* r) {
  if (!r || !r->snap || r->snap_istep < 0) {
    set_error("loma_restore: no snapshot taken");
    return LOMA_ERR_STATE;
  }
  if (r->inflight) { set_error("loma_restore: a step is in flight"); return LOMA_ERR_STATE; }
  std::vector<std::pair<void*, size_t>> parts;
  snapshot_parts(r, parts);
  const char* d = (const char*)r->snap;
  for (auto& p : parts) {
    HIPCHK(hipMemcpyAsync(p.first, d, p.second, hipMemcpyDeviceToDevice, r->stream));
    d += align256(p.second);
  }
  HIPCHK(hipStreamSynchronize(r->stream));
  r->istep = r->snap_istep;
  return LOMA_OK;
}

int loma_set_cfl(loma_run* r, double cfl) {
  if (!r || !(cfl > 0.0)) { set_error("loma_set_cfl: cfl must be positive"); return LOMA_ERR_ARG; }
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  if (r->stream) HIPCHK(hipStreamSynchronize(r->stream));
  r->cfl = cfl;
  r->rc.cfl = cfl;
  // the step graphs and ensemble graphs hold the run constants by value
  for (int q = 0; q < 2; ++q)
    if (r->gexec[q]) { hipGraphExecDestroy(r->gexec[q]); r->gexec[q] = nullptr; }
  ensemble_forget(r);
  return LOMA_OK;
}

int loma_set_physics(loma_run* r, double Re, double pr, double cfl, double dt0) {
  if (!r || !(Re > 0.0) || !(pr > 0.0) || !(cfl > 0.0) || !(dt0 > 0.0)) {
    set_error("loma_set_physics: Re, pr, cfl and dt0 must be positive");
    return LOMA_ERR_ARG;
  }
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  if (r->stream) HIPCHK(hipStreamSynchronize(r->stream));
  r->Re = Re; r->pr = pr; r->cfl = cfl; r->dt0 = dt0;
  r->rc.ire = 1.0 / Re;
  r->rc.peclet = Re * pr;
  r->rc.cfl = cfl;
  // the step graphs hold the run constants by value
  for (int q = 0; q < 2; ++q)
    if (r->gexec[q]) { hipGraphExecDestroy(r->gexec[q]); r->gexec[q] = nullptr; }
  ensemble_forget(r);
  r->uploaded = false;  // a new state must follow
  return LOMA_OK;
}

int loma_use_graph(loma_run* r, int enable) {
  if (!r) return LOMA_ERR_ARG;
  r->use_graph = enable != 0;
  if (r->use_graph && r->uploaded && !r->prof) {
    // capture now, so that a later step only launches
    int e = ensure_stream(r);
    for (int icfl = 0; icfl < 2 && !e; ++icfl) e = build_graph(r, icfl);
    if (e) return e;
  }
  return LOMA_OK;
}

// A run without a caller-set stream gets one of its own, non-blocking with
// respect to the null stream, so that set-up and transfers of other runs on
// the null stream neither wait for nor delay its steps.
static int ensure_stream(loma_run* r) {
  if (r->stream) return LOMA_OK;
  HIPCHK(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
  r->owns_stream = true;
  return LOMA_OK;
}

static int run_steps(loma_run* r, int nsteps) {
  for (int n = 0; n < nsteps; ++n) {
    // cross.f:238: the step size is re-estimated every ncfl steps
    const int icfl = ((r->istep) % r->ncfl == 0) ? 1 : 0;
    // steps of an open statistics window run their launches directly: the
    // captured step graphs hold no statistics kernels
    if (r->use_graph && !r->prof && !r->st_on) {
      if (!r->gexec[icfl]) {
        int e = build_graph(r, icfl);
        if (e) return e;
      }
      HIPCHK(hipGraphLaunch(r->gexec[icfl], r->stream));
    } else {
      int e = enqueue_step(r, icfl);
      if (e) return e;
    }
    r->istep++;
  }
  return LOMA_OK;
}

// device scalars -> host on the run's own stream (through the pinned copy
// when the run has one), so the read waits for this run's work only
static int read_scalars(loma_run* r, DevScalars* s) {
  if (r->hsc) {
    HIPCHK(hipMemcpyAsync(r->hsc, r->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    *s = *r->hsc;
  } else {
    HIPCHK(hipMemcpyAsync(s, r->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
  }
  return LOMA_OK;
}

int loma_advance(loma_run* r, int nsteps) {
  if (!r || !r->uploaded) { set_error("loma_advance: nothing uploaded"); return LOMA_ERR_STATE; }
  if (nsteps < 0) { set_error("loma_advance: nsteps < 0"); return LOMA_ERR_ARG; }
  if (r->use_graph) {  // graphs need a non-default stream
    int e = ensure_stream(r);
    if (e) return e;
  }
  int e = run_steps(r, nsteps);
  if (e) return e;
  DevScalars s;
  if ((e = read_scalars(r, &s))) return e;
  if (s.nan) { set_error("loma_advance: non-finite state"); return LOMA_ERR_NAN; }
  return LOMA_OK;
}

int loma_advance_async(loma_run* r, int nsteps) {
  if (!r || !r->uploaded) { set_error("loma_advance_async: nothing uploaded"); return LOMA_ERR_STATE; }
  if (nsteps < 0) { set_error("loma_advance_async: nsteps < 0"); return LOMA_ERR_ARG; }
  if (r->use_graph) {
    int e = ensure_stream(r);
    if (e) return e;
  }
  return run_steps(r, nsteps);
}

int loma_sync(loma_run* r) {
  if (!r) return LOMA_ERR_ARG;
  DevScalars s;
  int e = read_scalars(r, &s);
  if (e) return e;
  if (s.nan) { set_error("loma_sync: non-finite state"); return LOMA_ERR_NAN; }
  return LOMA_OK;
}

int loma_profile(loma_run* r, int enable) {
  if (!r) return LOMA_ERR_ARG;
  r->prof = enable != 0;
  return LOMA_OK;
}

int loma_profile_read(loma_run* r, double* total_ms, int* counts, int nclasses) {
  if (!r || !total_ms || !counts) return LOMA_ERR_ARG;
  HIPCHK(hipStreamSynchronize(r->stream));
  for (int c = 0; c < nclasses; ++c) { total_ms[c] = 0.0; counts[c] = 0; }
  for (auto& e : r->prof_ev) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
    if (e.cls < nclasses) { total_ms[e.cls] += ms; counts[e.cls] += 1; }
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  r->prof_ev.clear();
  return LOMA_OK;
}

int loma_stream(loma_run* r, void** stream) {
  if (!r || !stream) return LOMA_ERR_ARG;
  *stream = (void*)r->stream;
  return LOMA_OK;
}

int loma_advance_many(loma_run** runs, int nruns, int nsteps) {
  if (!runs || nruns <= 0 || nsteps < 0) return LOMA_ERR_ARG;
  bool graph = true, all_tune = true;
  for (int q = 0; q < nruns; ++q) {
    loma_run* r = runs[q];
    if (!r || !r->uploaded) { set_error("loma_advance_many: run not uploaded"); return LOMA_ERR_STATE; }
    if (r->inflight) { set_error("loma_advance_many: a run has a step in flight"); return LOMA_ERR_STATE; }
    if (r->st_on) {
      set_error("loma_advance_many: a run has an open statistics window (advance it alone)");
      return LOMA_ERR_STATE;
    }
    for (int p = 0; p < q; ++p)
      if (runs[p] == r) { set_error("loma_advance_many: run listed twice"); return LOMA_ERR_ARG; }
    int e = ensure_stream(r);  // every run needs a stream of its own
    if (e) return e;
    graph = graph && r->use_graph && !r->prof;
    all_tune = all_tune && (r->flags & LOMA_FLAG_AUTOTUNE);
  }
  for (int n = 0; n < nsteps; ++n) {
    std::vector<int> icfl(nruns);
    for (int q = 0; q < nruns; ++q) icfl[q] = (runs[q]->istep % runs[q]->ncfl == 0) ? 1 : 0;
    if (!graph) {
      // independent streams: the runs still overlap on the device
      for (int q = 0; q < nruns; ++q) {
        int e = enqueue_step(runs[q], icfl[q]);
        if (e) return e;
      }
    } else {
      // lock order everywhere: set-up, then ensembles
      std::lock_guard<std::recursive_mutex> lk2(g_setup_mu);
      std::lock_guard<std::mutex> lk(g_ens_mu);
      // batch shape: fixed by the environment, tried and kept (autotune), or
      // one launch per group
      int cap = ens_batch_env();
      EnsShape* shape = nullptr;
      if (!cap && all_tune) {
        for (auto& e : g_shape)
          if ((int)e.runs.size() == nruns && std::equal(e.runs.begin(), e.runs.end(), runs))
            shape = &e;
        if (!shape) {
          g_shape.push_back(EnsShape{});
          shape = &g_shape.back();
          shape->runs.assign(runs, runs + nruns);
        }
        cap = shape->cap ? shape->cap : ENS_CAPS[shape->trial & 1];
        if (shape->cap) shape = nullptr;  // decided: no timing
      }
      if (!cap) cap = ENS_CAPS[0];
      EnsembleGraph* eg = nullptr;
      for (auto& c : g_ens)
        if ((int)c.runs.size() == nruns && c.icfl == icfl && c.cap == cap &&
            std::equal(c.runs.begin(), c.runs.end(), runs))
          eg = &c;
      if (!eg) {
        EnsembleGraph c;
        c.runs.assign(runs, runs + nruns);
        c.icfl = icfl;
        c.cap = cap;
        // argument blocks of the batched launches, on the device before the
        // capture starts (nothing is allocated while it records)
        EnsPlan plan;
        ensemble_plan(runs, nruns, icfl, cap, plan);
        HIPCHK(hipMalloc(&c.desc, plan.bytes.size()));
        if (hipMemcpy(c.desc, plan.bytes.data(), plan.bytes.size(), hipMemcpyHostToDevice) !=
            hipSuccess) {
          hipFree(c.desc);
          set_error("loma_advance_many: descriptor upload failed");
          return LOMA_ERR_DEVICE;
        }
        HIPCHK(hipStreamCreate(&c.master));
        c.lanes.resize(plan.lanes.size(), nullptr);
        for (auto& l : c.lanes) HIPCHK(hipStreamCreateWithFlags(&l, hipStreamNonBlocking));
        hipEvent_t fork;
        std::vector<hipEvent_t> join(plan.lanes.size());
        HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        for (auto& e : join) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        // one level of fork and join: every lane waits on the ensemble stream
        // and is joined back to it before the capture ends
        HIPCHK(hipStreamBeginCapture(c.master, hipStreamCaptureModeThreadLocal));
        int err = hipEventRecord(fork, c.master);
        for (size_t k = 0; k < plan.lanes.size() && !err; ++k) {
          err = hipStreamWaitEvent(c.lanes[k], fork, 0);
          if (!err) err = enqueue_lane(c.lanes[k], c.desc, plan.lanes[k]);
          if (!err) err = hipEventRecord(join[k], c.lanes[k]);
          if (!err) err = hipStreamWaitEvent(c.master, join[k], 0);
        }
        hipGraph_t g;
        hipError_t ce = hipStreamEndCapture(c.master, &g);
        hipEventDestroy(fork);
        for (auto& e : join) hipEventDestroy(e);
        if (err) {
          for (auto& l : c.lanes) hipStreamDestroy(l);
          hipStreamDestroy(c.master);
          hipFree(c.desc);
          set_error("loma_advance_many: capture failed");
          return LOMA_ERR_DEVICE;
        }
        HIPCHK(ce);
        hipError_t ie = hipGraphInstantiate(&c.exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        HIPCHK(ie);
        c.ready.resize(nruns, nullptr);
        for (auto& e : c.ready) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        g_ens.push_back(c);
        eg = &g_ens.back();
      }
      // the graph runs on its own stream: order it after whatever is already
      // queued on each member run's stream (an earlier asynchronous advance)
      for (int q = 0; q < nruns; ++q) {
        HIPCHK(hipEventRecord(eg->ready[q], runs[q]->stream));
        HIPCHK(hipStreamWaitEvent(eg->master, eg->ready[q], 0));
      }
      if (shape) {
        // a trial step of the batch shape: timed on the device
        hipEvent_t t0, t1;
        HIPCHK(hipEventCreate(&t0));
        HIPCHK(hipEventCreate(&t1));
        HIPCHK(hipEventRecord(t0, eg->master));
        HIPCHK(hipGraphLaunch(eg->exec, eg->master));
        HIPCHK(hipEventRecord(t1, eg->master));
        HIPCHK(hipEventSynchronize(t1));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, t0, t1));
        hipEventDestroy(t0);
        hipEventDestroy(t1);
        const int v = shape->trial & 1;
        shape->best[v] = std::min(shape->best[v], ms);
        if (++shape->trial >= ENS_TRIALS)
          shape->cap = ENS_CAPS[shape->best[1] < shape->best[0] ? 1 : 0];
      } else {
        HIPCHK(hipGraphLaunch(eg->exec, eg->master));
      }
    }
    for (int q = 0; q < nruns; ++q) runs[q]->istep++;
  }
  // wait for every run (graph launches complete on their master streams)
  HIPCHK(hipDeviceSynchronize());
  for (int q = 0; q < nruns; ++q) {
    DevScalars s;
    HIPCHK(hipMemcpy(&s, runs[q]->sc, sizeof(s), hipMemcpyDeviceToHost));
    if (s.nan) { set_error("loma_advance_many: non-finite state in a run"); return LOMA_ERR_NAN; }
  }
  return LOMA_OK;
}

int loma_tiles(const loma_run* r, int* J, int* L) {
  if (!r || !J || !L) return LOMA_ERR_ARG;
  *J = r->J;
  *L = r->L;
  return LOMA_OK;
}

int loma_footprint(int mgalx, int my, int mgalz, int precision, int flags,
                   long long* bytes) {
  if (!bytes || mgalx < 8 || mgalz < 8 || my < 16 || my > 1024 || pick_M(my) < 0 ||
      (precision != LOMA_FP64 && precision != LOMA_MIXED)) {
    set_error("loma_footprint: grid or precision out of range");
    return LOMA_ERR_ARG;
  }
  *bytes = (long long)plan_bytes(mgalx, my, mgalz, precision == LOMA_MIXED,
                                 (flags & LOMA_FLAG_SCALAR_Z) != 0);
  return LOMA_OK;
}

int loma_step_begin(loma_run* r) {
  if (!r || !r->uploaded) { set_error("loma_step_begin: nothing uploaded"); return LOMA_ERR_STATE; }
  if (r->inflight) { set_error("loma_step_begin: a step is in flight"); return LOMA_ERR_STATE; }
  if (!r->stream) {
    std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
    int e = ensure_stream(r);
    if (e) return e;
  }
  HIPCHK(hipEventRecord(r->sev[0], r->stream));
  int e = run_steps(r, 1);
  if (e) return e;
  HIPCHK(hipMemcpyAsync(r->hsc, r->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, r->stream));
  HIPCHK(hipEventRecord(r->sev[1], r->stream));
  r->inflight = true;
  return LOMA_OK;
}

int loma_step_poll(loma_run* r, int* done, double* device_ms, int* nonfinite) {
  if (!r || !done) return LOMA_ERR_ARG;
  if (!r->inflight) { set_error("loma_step_poll: no step in flight"); return LOMA_ERR_STATE; }
  const hipError_t q = hipEventQuery(r->sev[1]);
  if (q == hipErrorNotReady) { *done = 0; return LOMA_OK; }
  HIPCHK(q);
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, r->sev[0], r->sev[1]));
  *done = 1;
  if (device_ms) *device_ms = ms;
  if (nonfinite) *nonfinite = r->hsc->nan;
  r->inflight = false;
  return LOMA_OK;
}

int loma_scalars(loma_run* r, double* out) {
  if (!r || !out) return LOMA_ERR_ARG;
  DevScalars s;
  int e = read_scalars(r, &s);
  if (e) return e;
  out[0] = s.dt; out[1] = s.time; out[2] = s.dtr; out[3] = r->istep; out[4] = s.nan;
  return LOMA_OK;
}

int loma_thickness(loma_run* r, double* out) {
  if (!r || !out) return LOMA_ERR_ARG;
  if (!(r->flags & LOMA_FLAG_HISTORY)) {
    set_error("loma_thickness: the run was made without LOMA_FLAG_HISTORY");
    return LOMA_ERR_STATE;
  }
  DevScalars s;
  int e = read_scalars(r, &s);
  if (e) return e;
  out[0] = s.dm;
  out[1] = s.dw;
  return LOMA_OK;
}

int loma_history(loma_run* r, double* out) {
  if (!r || !out) return LOMA_ERR_ARG;
  DevScalars s;
  int e = read_scalars(r, &s);
  if (e) return e;
  const bool h = (r->flags & LOMA_FLAG_HISTORY) != 0;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  out[0] = s.time;
  out[1] = h ? s.dm : nan;
  out[2] = h ? s.dw : nan;
  out[3] = h ? s.dw * r->Re : nan;
  out[4] = s.trapz;
  out[5] = s.mass[0];
  out[6] = s.mass[1];
  out[7] = s.mass[2];
  return LOMA_OK;
}

// ---- statistics window ----------------------------------------------------
int loma_stats_config(loma_run* r, int nspec, const int* jspecy) {
  if (!r || nspec < 0 || (nspec > 0 && !jspecy)) {
    set_error("loma_stats_config: bad arguments");
    return LOMA_ERR_ARG;
  }
  if (r->inflight) { set_error("loma_stats_config: a step is in flight"); return LOMA_ERR_STATE; }
  const int my = r->my;
  // jsptot: the given rows, then the centre row (my+1)/2 (main.f:698-703)
  std::vector<int> rows;
  for (int n = 0; n < nspec; ++n) {
    if (jspecy[n] < 1 || jspecy[n] > my) {
      set_error("loma_stats_config: spectra row outside 1..my");
      return LOMA_ERR_ARG;
    }
    rows.push_back(jspecy[n] - 1);
  }
  rows.push_back((my + 1) / 2 - 1);
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  if (r->stream) HIPCHK(hipStreamSynchronize(r->stream));
  for (void** p : {(void**)&r->st_acc, (void**)&r->st_spe, (void**)&r->st_part,
                   (void**)&r->st_rho, (void**)&r->st_last, (void**)&r->st_jsp})
    if (*p) { hipFree(*p); *p = nullptr; }
  const int nsp = (int)rows.size();
  const size_t nacc = (size_t)STAT_NCOL * my;
  const size_t nspe = (size_t)r->nplanes * 12 * nsp * (r->nz + 1);
  const size_t npart = (size_t)r->nplanes * ((r->nz + STAT_KACH) / STAT_KACH) * STAT_NPART * my;
  const size_t nrho = (size_t)r->mgalz * my;
  struct { void** p; size_t bytes; } al[] = {
      {(void**)&r->st_acc, nacc * 8}, {(void**)&r->st_spe, nspe * 8},
      {(void**)&r->st_part, npart * 8}, {(void**)&r->st_rho, nrho * 8},
      {(void**)&r->st_last, (size_t)9 * my * 8}, {(void**)&r->st_jsp, (size_t)nsp * 4}};
  for (auto& a : al) {
    if (hipMalloc(a.p, a.bytes) != hipSuccess) {
      *a.p = nullptr;
      set_error("loma_stats_config: device allocation failed");
      return LOMA_ERR_OOM;
    }
  }
  r->st_nsp = nsp;
  HIPCHK(hipMemcpy(r->st_jsp, rows.data(), (size_t)nsp * 4, hipMemcpyHostToDevice));
  r->st_on = false;
  return loma_stats_reset(r);
}

int loma_stats_accumulate(loma_run* r, int on) {
  if (!r) return LOMA_ERR_ARG;
  if (on && !r->st_acc) {
    set_error("loma_stats_accumulate: no statistics window (loma_stats_config)");
    return LOMA_ERR_STATE;
  }
  r->st_on = on != 0;
  return LOMA_OK;
}

int loma_stats_reset(loma_run* r) {
  if (!r) return LOMA_ERR_ARG;
  if (!r->st_acc) { set_error("loma_stats_reset: no statistics window"); return LOMA_ERR_STATE; }
  hipStream_t st = r->stream;
  const int my = r->my;
  HIPCHK(hipMemsetAsync(r->st_acc, 0, sizeof(double) * STAT_NCOL * my, st));
  HIPCHK(hipMemsetAsync(r->st_spe, 0,
                        sizeof(double) * r->nplanes * 12 * r->st_nsp * (r->nz + 1), st));
  HIPCHK(hipMemsetAsync(r->st_last, 0, sizeof(double) * 9 * my, st));
  HIPCHK(hipStreamSynchronize(st));
  r->st_nacum = 0;
  return LOMA_OK;
}

int loma_stats_read(loma_run* r, double* sta, float* spe, double* info) {
  if (!r) return LOMA_ERR_ARG;
  if (!r->st_acc) { set_error("loma_stats_read: no statistics window"); return LOMA_ERR_STATE; }
  if (r->inflight) { set_error("loma_stats_read: a step is in flight"); return LOMA_ERR_STATE; }
  const int my = r->my;
  const size_t nspe = (size_t)r->nplanes * 12 * r->st_nsp * (r->nz + 1);
  std::vector<double> acc((size_t)STAT_NCOL * my), last((size_t)9 * my), sp(spe ? nspe : 0);
  hipStream_t st = r->stream;
  HIPCHK(hipMemcpyAsync(acc.data(), r->st_acc, acc.size() * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(last.data(), r->st_last, last.size() * 8, hipMemcpyDeviceToHost, st));
  if (spe) HIPCHK(hipMemcpyAsync(sp.data(), r->st_spe, nspe * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // writestats (escru.f:366-414): columns 22.. are sums over the x-z plane,
  // written times 1/(mgalx mgalz)
  const double fac = 1.0 / ((double)r->mgalx * r->mgalz);
  if (sta)
    for (int j = 0; j < my; ++j)
      for (int c = 0; c < STAT_NCOL; ++c)
        sta[(size_t)j * STAT_NCOL + c] = acc[(size_t)c * my + j] * (c >= 22 ? fac : 1.0);
  if (spe)
    for (size_t q = 0; q < nspe; ++q) spe[q] = (float)sp[q];
  if (info) {
    std::vector<double> hy(my);
    const std::vector<double>& y = r->y;
    for (int j = 1; j < my - 1; ++j) hy[j] = (y[j + 1] - y[j - 1]) / 2.0;  // main.f:689-696
    hy[0] = (y[1] - y[0]) / 2.0;
    hy[my - 1] = (y[my - 1] - y[my - 2]) / 2.0;
    info[0] = r->st_nacum;
    info[1] = r->st_nacum;
    // masst (escru.f:372-376) sums wkstats(j,18) hy(j+1) for j = 0..my-1;
    // wkstats is declared (my,22), so its j = 0 term reads the element before
    // column 18 in memory, wkstats(my,17): rww of the top row.  Kept as the
    // reference computes it, so the .cf column is the same.
    double masst = fac * acc[(size_t)27 * my + (my - 1)] * hy[0];
    for (int j = 1; j < my; ++j) masst += fac * acc[(size_t)28 * my + (j - 1)] * hy[j];
    info[2] = masst;
    // enerdis = sum_j ep(j)/nacum hy(j) (cross.f:1152-1158)
    double enerdis = 0.0;
    if (r->st_nacum)
      for (int j = 0; j < my; ++j) enerdis += acc[(size_t)21 * my + j] / r->st_nacum * hy[j];
    info[3] = enerdis;
    // energy(i) = sqrt(|uner(i)|) of the last accumulated step (cross.f:802-816)
    for (int n = 0; n < 9; ++n) {
      double e = 0.0;
      for (int j = 0; j < my; ++j) e += last[(size_t)n * my + j] * hy[j];
      info[4 + n] = std::sqrt(std::fabs(e));
    }
  }
  return LOMA_OK;
}

int loma_cross1(loma_run* r, float* vor, float* phi, float* psi, float* scal, double* u00,
                double* w00, double* v00, int nstep) {
  int e = loma_upload(r, vor, phi, psi, scal, u00, v00, w00, 0);
  if (e) return e;
  e = loma_advance(r, nstep);
  if (e && e != LOMA_ERR_NAN) return e;
  int e2 = loma_download(r, vor, phi, psi, scal, u00, v00, w00, 0);
  return e ? e : e2;
}

// ---- per-operator entry points -----------------------------------------
int loma_deryr2_dev(loma_run* r, const double* u, double* du, int ncols) {
  if (!r || ncols < 0) return LOMA_ERR_ARG;
  LCHK(launch_dy(r->M, 1, r->T, u, du, ncols, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  return LOMA_OK;
}
int loma_deryyr2_dev(loma_run* r, const double* u, double* du, int ncols) {
  if (!r || ncols < 0) return LOMA_ERR_ARG;
  LCHK(launch_dy(r->M, 2, r->T, u, du, ncols, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  return LOMA_OK;
}

static int solve_cols(loma_run* r, const double* phi, double* v, double* dvdy,
                      const double* rK, double shift, int ncols, int mode) {
  if (!r || ncols < 0) return LOMA_ERR_ARG;
  if (ncols == 0) return LOMA_OK;
  std::vector<double> h(ncols);
  HIPCHK(hipMemcpy(h.data(), rK, sizeof(double) * ncols, hipMemcpyDeviceToHost));
  for (double& x : h) x += shift;
  double *sh = nullptr, *lu = nullptr, *bc = nullptr;
  HIPCHK(hipMalloc(&sh, sizeof(double) * ncols));
  HIPCHK(hipMalloc(&lu, sizeof(double) * (size_t)r->lu_stride * ncols));
  HIPCHK(hipMalloc(&bc, sizeof(double) * 4 * (size_t)ncols));
  int rc = LOMA_OK;
  if (hipMemcpy(sh, h.data(), sizeof(double) * ncols, hipMemcpyHostToDevice) != hipSuccess ||
      launch_helm_factor(r->dt21, r->dt22, r->my, r->M, sh, ncols, lu, r->stream) ||
      launch_helm_bc(r->M, r->T, lu, bc, nullptr, ncols, r->stream) ||
      launch_lapvdvhom(r->M, r->T, lu, bc, phi, v, dvdy, rK, ncols, mode, r->stream) ||
      hipStreamSynchronize(r->stream) != hipSuccess) {
    set_error("solve_cols: device failure");
    rc = LOMA_ERR_DEVICE;
  }
  hipFree(sh);
  hipFree(lu);
  hipFree(bc);
  return rc;
}

int loma_lapvdvhom_dev(loma_run* r, const double* phi, double* v, double* dvdy,
                       const double* rK, int ncols) {
  return solve_cols(r, phi, v, dvdy, rK, 0.0, ncols, 0);
}
int loma_solvepsi_dev(loma_run* r, const double* phi, double* v, double* dvdy,
                      const double* rK, int ncols) {
  return solve_cols(r, phi, v, dvdy, rK, 0.0, ncols, 1);
}
int loma_helmholtz_dev(loma_run* r, const double* f, double* x, const double* rK,
                       double shift, int ncols) {
  return solve_cols(r, f, x, nullptr, rK, shift, ncols, 2);
}
int loma_visc_dev(loma_run* r, double* vorwk, double* vor, double* phi, double* v,
                  double* dvdy, double* phiwk, const double* rk1, const double* rk2,
                  double dalbe, double dtri, double dtxi, int ncols) {
  if (!r || ncols < 0 || !vorwk || !vor || !phi || !v || !dvdy || !phiwk || !rk1 || !rk2)
    return LOMA_ERR_ARG;
  if (ncols == 0) return LOMA_OK;
  double *lu = nullptr, *bc = nullptr;
  const size_t L = (size_t)r->lu_stride * ncols;
  HIPCHK(hipMalloc(&lu, sizeof(double) * 2 * L));
  if (hipMalloc(&bc, sizeof(double) * 4 * (size_t)ncols) != hipSuccess) {
    hipFree(lu);
    set_error("loma_visc_dev: allocation failed");
    return LOMA_ERR_OOM;
  }
  int rc = LOMA_OK;
  if (launch_helm_factor(r->dt21, r->dt22, r->my, r->M, rk1, ncols, lu, r->stream) ||
      launch_helm_factor(r->dt21, r->dt22, r->my, r->M, rk2, ncols, lu + L, r->stream) ||
      launch_helm_bc(r->M, r->T, lu + L, bc, nullptr, ncols, r->stream) ||
      launch_visc(r->M, r->T, lu, lu + L, bc, vorwk, vor, phi, v, dvdy, phiwk, rk2, dalbe,
                  dtri, dtxi, ncols, r->stream) ||
      hipStreamSynchronize(r->stream) != hipSuccess) {
    set_error("loma_visc_dev: device failure");
    rc = LOMA_ERR_DEVICE;
  }
  hipFree(lu);
  hipFree(bc);
  return rc;
}

int loma_inty8_dev(loma_run* r, const double* du, double* u) {
  if (!r) return LOMA_ERR_ARG;
  LCHK(launch_inty8(r->M, r->T, du, u, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  return LOMA_OK;
}

// host-pointer float versions with the reference's argument shapes
static int host_colop(loma_run* r, const float* u, float* du, int m, int order) {
  if (!r || !u || !du || m < 0) return LOMA_ERR_ARG;
  const size_t n = (size_t)2 * r->my * m;
  std::vector<double> h(n);
  for (size_t q = 0; q < n; ++q) h[q] = u[q];
  double* d = nullptr;
  HIPCHK(hipMalloc(&d, n * sizeof(double)));
  int rc = LOMA_OK;
  if (hipMemcpy(d, h.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess ||
      launch_dy(r->M, order, r->T, d, d, m, r->stream) ||
      hipStreamSynchronize(r->stream) != hipSuccess ||
      hipMemcpy(h.data(), d, n * 8, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("host_colop: device failure");
    rc = LOMA_ERR_DEVICE;
  }
  hipFree(d);
  if (rc == LOMA_OK)
    for (size_t q = 0; q < n; ++q) du[q] = (float)h[q];
  return rc;
}
int loma_deryr2(loma_run* r, const float* u, float* du, int m) { return host_colop(r, u, du, m, 1); }
int loma_deryyr2(loma_run* r, const float* u, float* du, int m) { return host_colop(r, u, du, m, 2); }

static int host_solve(loma_run* r, const float* phi, float* v, float* dvdy, double rK, int mode) {
  if (!r || !phi || !v || !dvdy) return LOMA_ERR_ARG;
  const size_t n = (size_t)2 * r->my;
  std::vector<double> h(3 * n + 1);
  for (size_t q = 0; q < n; ++q) h[q] = phi[q];
  h[3 * n] = rK;
  double* d = nullptr;
  HIPCHK(hipMalloc(&d, h.size() * sizeof(double)));
  int rc = hipMemcpy(d, h.data(), h.size() * 8, hipMemcpyHostToDevice) == hipSuccess
               ? solve_cols(r, d, d + n, d + 2 * n, d + 3 * n, 0.0, 1, mode)
               : LOMA_ERR_DEVICE;
  if (rc == LOMA_OK && hipMemcpy(h.data(), d, h.size() * 8, hipMemcpyDeviceToHost) != hipSuccess)
    rc = LOMA_ERR_DEVICE;
  hipFree(d);
  if (rc == LOMA_OK)
    for (size_t q = 0; q < n; ++q) { v[q] = (float)h[n + q]; dvdy[q] = (float)h[2 * n + q]; }
  return rc;
}
int loma_lapvdvhom(loma_run* r, const float* phi, float* v, float* dvdy, double rK) {
  return host_solve(r, phi, v, dvdy, rK, 0);
}
int loma_solvepsi(loma_run* r, const float* phi, float* v, float* dvdy, double rK) {
  return host_solve(r, phi, v, dvdy, rK, 1);
}

int loma_visc(loma_run* r, float* vorwk, float* vor, float* phi, float* v, float* dvdy,
              float* phiwk, double rk1, double rk2, double dalbe, double dtri, double dtxi) {
  if (!r || !vorwk || !vor || !phi || !v || !dvdy || !phiwk) return LOMA_ERR_ARG;
  const size_t n = (size_t)2 * r->my;
  float* cols[6] = {vorwk, vor, phi, v, dvdy, phiwk};
  std::vector<double> h(6 * n + 2);
  for (int c = 0; c < 6; ++c)
    for (size_t q = 0; q < n; ++q) h[c * n + q] = cols[c][q];
  h[6 * n] = rk1;
  h[6 * n + 1] = rk2;
  double* d = nullptr;
  HIPCHK(hipMalloc(&d, h.size() * sizeof(double)));
  int rc = hipMemcpy(d, h.data(), h.size() * 8, hipMemcpyHostToDevice) == hipSuccess
               ? loma_visc_dev(r, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n,
                               d + 6 * n, d + 6 * n + 1, dalbe, dtri, dtxi, 1)
               : LOMA_ERR_DEVICE;
  if (rc == LOMA_OK && hipMemcpy(h.data(), d, h.size() * 8, hipMemcpyDeviceToHost) != hipSuccess)
    rc = LOMA_ERR_DEVICE;
  hipFree(d);
  if (rc == LOMA_OK)
    for (int c = 0; c < 6; ++c)
      for (size_t q = 0; q < n; ++q) cols[c][q] = (float)h[c * n + q];
  return rc;
}

// ---- foup2: the transform stage with the reference's argument list -------
// Arrays as foup2 declares them (cross.f:1282-1310): complex*8
// (0:my-1, 0:mgalz-1, planes), the spectral content in the first mz kz
// slots of every plane block; rhst real*8 (0:mgalx-1, lines), line
// z + mgalz*y.  The run's device buffers carry the stage; dt is the step
// size of the substep (the tem module's Deltat).
static int put_spec(loma_run* r, const float* in, int q, double* mean) {
  const int np = r->nplanes, mz = r->mz, my = r->my, mgz = r->mgalz, myp = r->rc.myp;
  const size_t n = (size_t)np * mz * myp;
  std::vector<double> h(2 * n, 0.0);
  for (int i = 0; i < np; ++i)
    for (int k = 0; k < mz; ++k)
      for (int j = 0; j < my; ++j) {
        const size_t src = 2 * (((size_t)i * mgz + k) * my + j);
        const size_t dst = 2 * (((size_t)i * mz + k) * myp + j);
        if (i == 0 && k == 0) {
          // the 00 mode travels beside the transform (mean_in), real part only
          mean[(size_t)q * my + j] = in[src];
          continue;
        }
        h[dst] = in[src];
        h[dst + 1] = in[src + 1];
      }
  if (r->fp32) {
    std::vector<float> f(h.begin(), h.end());
    HIPCHK(hipMemcpy(r->spec[q], f.data(), f.size() * 4, hipMemcpyHostToDevice));
  } else {
    HIPCHK(hipMemcpy(r->spec[q], h.data(), h.size() * 8, hipMemcpyHostToDevice));
  }
  return LOMA_OK;
}
static int get_spec(loma_run* r, int q, const double* mean, float* out) {
  const int np = r->nplanes, mz = r->mz, my = r->my, mgz = r->mgalz, myp = r->rc.myp;
  const size_t n = (size_t)np * mz * myp;
  std::vector<double> h(2 * n);
  if (r->fp32) {
    std::vector<float> f(2 * n);
    HIPCHK(hipMemcpy(f.data(), r->spec[q], f.size() * 4, hipMemcpyDeviceToHost));
    for (size_t e = 0; e < 2 * n; ++e) h[e] = f[e];
  } else {
    HIPCHK(hipMemcpy(h.data(), r->spec[q], h.size() * 8, hipMemcpyDeviceToHost));
  }
  for (int i = 0; i < np; ++i)
    for (int k = 0; k < mz; ++k)
      for (int j = 0; j < my; ++j) {
        const size_t dst = 2 * (((size_t)i * mgz + k) * my + j);
        const size_t src = 2 * (((size_t)i * mz + k) * myp + j);
        double re = h[src], im = h[src + 1];
        if (i == 0 && k == 0) { re += mean[(size_t)q * my + j]; im = 0.0; }
        out[dst] = (float)re;
        out[dst + 1] = (float)im;
      }
  return LOMA_OK;
}

int loma_foup2(loma_run* r, float* up1, float* up2, float* up3, float* ten11, float* ten22,
               float* ten33, double* rhst, float* drho, float* tnext, float* ten12,
               float* ten13, float* ten23, int rkstep, double dt, double* dtr) {
  if (!r || !up1 || !up2 || !up3 || !ten11 || !ten22 || !ten33 || !rhst || !drho ||
      !tnext || !ten12 || !ten13 || !ten23 || rkstep < 1 || rkstep > 3 || !(dt > 0.0)) {
    set_error("loma_foup2: null array, rkstep outside 1..3 or dt <= 0");
    return LOMA_ERR_ARG;
  }
  if (r->with_z) { set_error("loma_foup2: the reference's foup2 carries no scalar Z"); return LOMA_ERR_ARG; }
  if (r->inflight) { set_error("loma_foup2: a step is in flight"); return LOMA_ERR_STATE; }
  std::lock_guard<std::recursive_mutex> lk(g_setup_mu);
  int e = ensure_stream(r);
  if (e) return e;
  HIPCHK(hipStreamSynchronize(r->stream));
  const int my = r->my, mgx = r->mgalx, mgz = r->mgalz;
  // inputs in the order the transform kernels take them (cross.f:1324-1332)
  const float* in[8] = {up1, up2, up3, ten22, ten12, ten13, ten23, drho};
  std::vector<double> mean(12 * (size_t)my, 0.0);
  for (int q = 0; q < 8; ++q)
    if ((e = put_spec(r, in[q], q, mean.data()))) return e;
  HIPCHK(hipMemcpy(r->mean_in, mean.data(), mean.size() * 8, hipMemcpyHostToDevice));
  // rhst: lines z + mgalz*y (reference) -> [z][y][x] (here)
  {
    std::vector<double> h((size_t)mgx * my * mgz);
    for (int j = 0; j < my; ++j)
      for (int z = 0; z < mgz; ++z)
        std::memcpy(&h[((size_t)z * my + j) * mgx], &rhst[((size_t)j * mgz + z) * mgx],
                    sizeof(double) * mgx);
    HIPCHK(hipMemcpy(r->rhst, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  }
  // the step size of the substep
  DevScalars s{};
  HIPCHK(hipMemcpy(&s, r->sc, sizeof(s), hipMemcpyDeviceToHost));
  s.dt = dt;
  HIPCHK(hipMemcpy(r->sc, &s, sizeof(s), hipMemcpyHostToDevice));
  const int do_cfl = rkstep == 1 ? 1 : 0;
  FftArgs fa = fft_args(r, rkstep, do_cfl);
  fa.nfields = 8;
  LCHK(launch_zfft_inv(r->fp32, fa, r->stream));
  LCHK(launch_xline(r->fp32, fa, r->stream));
  fa.nfields = 11;
  LCHK(launch_zfft_fwd(r->fp32, fa, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  HIPCHK(hipMemcpy(mean.data(), r->mean_out, mean.size() * 8, hipMemcpyDeviceToHost));
  float* out[11] = {up1, up2, up3, ten11, ten22, ten33, ten12, ten13, ten23, drho, tnext};
  for (int q = 0; q < 11; ++q)
    if ((e = get_spec(r, q, mean.data(), out[q]))) return e;
  {
    std::vector<double> h((size_t)mgx * my * mgz);
    HIPCHK(hipMemcpy(h.data(), r->rhst, h.size() * 8, hipMemcpyDeviceToHost));
    for (int j = 0; j < my; ++j)
      for (int z = 0; z < mgz; ++z)
        std::memcpy(&rhst[((size_t)j * mgz + z) * mgx], &h[((size_t)z * my + j) * mgx],
                    sizeof(double) * mgx);
  }
  if (dtr && do_cfl) {
    // the CFL step of cross.f:1778-1818 from the x-line tiles' maxima, one rank
    const int ntile = xl_tiles(my, mgz, r->L);
    std::vector<double> c(4 * (size_t)ntile);
    HIPCHK(hipMemcpy(c.data(), r->cflp, c.size() * 8, hipMemcpyDeviceToHost));
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = 0; t < ntile; ++t)
      for (int q = 0; q < 4; ++q) {
        const double v = c[4 * (size_t)t + q];
        if (v > m[q] || v != v) m[q] = v;
      }
    const double conv = m[0] * r->rc.hxalp * 0.5 + m[1] + m[2] * r->rc.hzbet * 0.5;
    *dtr = r->rc.cfl / (conv + m[3]);
  }
  // the run's state no longer matches its step graphs' assumptions
  r->uploaded = false;
  return LOMA_OK;
}

// ---- host-pointer transforms and 1-D operators with the reference's shapes
// One complex FFT of `nrows` contiguous rows of length n on the device
// (FFTPACK conventions, as k_fourz: iopt < 0 forward e^{-i} scaled by 1/n,
// else inverse e^{+i} unscaled), in and out of a host array of doubles.
static int host_cfft(loma_run* r, int n, std::vector<double>& rows, int nrows, int iopt) {
  if (nrows <= 0) return LOMA_OK;
  const FftPlan pl = make_plan(n);
  if (pl.npass < 0) { set_error("transform length has no plan"); return LOMA_ERR_ARG; }
  std::vector<double> tw(2 * (size_t)n);
  for (int t = 0; t < n; ++t) {
    tw[2 * t] = std::cos(2.0 * M_PI * t / n);
    tw[2 * t + 1] = -std::sin(2.0 * M_PI * t / n);
  }
  double* d = nullptr;
  const size_t nb = rows.size() * sizeof(double), tb = tw.size() * sizeof(double);
  HIPCHK(hipMalloc(&d, nb + tb));
  double* dtw = d + rows.size();
  int rc = LOMA_OK;
  if (hipMemcpy(d, rows.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dtw, tw.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
      launch_fourz(pl, (const double2*)dtw, (double2*)d, nrows, iopt, r->stream) ||
      hipStreamSynchronize(r->stream) != hipSuccess ||
      hipMemcpy(rows.data(), d, nb, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("host transform: device failure");
    rc = LOMA_ERR_DEVICE;
  }
  hipFree(d);
  return rc;
}

// cft (cftsingle.f:2): m complex vectors of length mgalz, element j of
// vector q at c[q*isa + j*ise] (re) and the next float (im)
int loma_cft(loma_run* r, float* c, int ise, int isa, int m, int iopt) {
  if (!r || !c || m < 0 || ise < 2 || isa < 0) return LOMA_ERR_ARG;
  const int n = r->mgalz;
  std::vector<double> h(2 * (size_t)n * m);
  for (int q = 0; q < m; ++q)
    for (int j = 0; j < n; ++j) {
      const size_t o = (size_t)q * isa + (size_t)j * ise;
      h[2 * ((size_t)q * n + j)] = c[o];
      h[2 * ((size_t)q * n + j) + 1] = c[o + 1];
    }
  int e = host_cfft(r, n, h, m, iopt);
  if (e) return e;
  for (int q = 0; q < m; ++q)
    for (int j = 0; j < n; ++j) {
      const size_t o = (size_t)q * isa + (size_t)j * ise;
      c[o] = (float)h[2 * ((size_t)q * n + j)];
      c[o + 1] = (float)h[2 * ((size_t)q * n + j) + 1];
    }
  return LOMA_OK;
}

// fourz (fou3D.f:80): a real(4) plane of my rows of mgalz complex numbers
int loma_fourz(loma_run* r, float* phys, int iopt) {
  if (!r) return LOMA_ERR_ARG;
  return loma_cft(r, phys, 2, 2 * r->mgalz, r->my, iopt);
}

// rft (rftsingle.f:25): m real vectors of n = mgalx points, vector q at
// c[q*isa], isa >= n + 2.  Forward: the packed coefficients
// c[2k] + i c[2k+1] = (1/n) sum_j x_j e^{-2 pi i jk/n}, k = 0..n/2, with the
// imaginary parts of k = 0 and n/2 set to zero (rftsingle.f:64-73).
// Inverse: x_j = Re c0 + 2 sum_{0<k<n/2} Re(ck e^{2 pi i jk/n}) + Re c_{n/2}
// (-1)^j, imaginary parts of k = 0, n/2 ignored; the two entries past the n
// physical points keep their values, as the reference's shift leaves them.
int loma_rft(loma_run* r, float* c, int isa, int m, int iopt) {
  const int n = r ? r->mgalx : 0;
  if (!r || !c || m < 0 || isa < n + 2) return LOMA_ERR_ARG;
  std::vector<double> h(2 * (size_t)n * m);
  for (int q = 0; q < m; ++q) {
    const float* v = c + (size_t)q * isa;
    double* row = h.data() + 2 * (size_t)q * n;
    if (iopt < 0) {
      for (int j = 0; j < n; ++j) { row[2 * j] = v[j]; row[2 * j + 1] = 0.0; }
    } else {
      for (int k = 0; k < n; ++k) { row[2 * k] = 0.0; row[2 * k + 1] = 0.0; }
      row[0] = v[0];
      for (int k = 1; k < n / 2; ++k) {
        row[2 * k] = v[2 * k];
        row[2 * k + 1] = v[2 * k + 1];
        row[2 * (n - k)] = v[2 * k];
        row[2 * (n - k) + 1] = -v[2 * k + 1];
      }
      row[2 * (n / 2)] = v[n];
    }
  }
  int e = host_cfft(r, n, h, m, iopt);
  if (e) return e;
  for (int q = 0; q < m; ++q) {
    float* v = c + (size_t)q * isa;
    const double* row = h.data() + 2 * (size_t)q * n;
    if (iopt < 0) {
      for (int k = 0; k <= n / 2; ++k) {
        v[2 * k] = (float)row[2 * k];
        v[2 * k + 1] = (float)row[2 * k + 1];
      }
      v[1] = 0.0f;
      v[n + 1] = 0.0f;
    } else {
      for (int j = 0; j < n; ++j) v[j] = (float)row[2 * j];
    }
  }
  return LOMA_OK;
}

// fourx (fou3D.f:29): one x-line of mgalx + 2 reals.  iopt >= 0: modes past
// mx are zeroed, the line goes to physical space, and phys8 receives the
// physical values (computed in FP64, so phys8 is not the rounded phys).
// iopt < 0: forward transform of phys in place (phys8 unused).
int loma_fourx(loma_run* r, float* phys, double* phys8, int iopt) {
  if (!r || !phys || (iopt >= 0 && !phys8)) return LOMA_ERR_ARG;
  const int n = r->mgalx;
  if (iopt < 0) return loma_rft(r, phys, n + 2, 1, -1);
  for (int j = r->mx; j < n + 2; ++j) phys[j] = 0.0f;
  std::vector<double> h(2 * (size_t)n, 0.0);
  h[0] = phys[0];
  for (int k = 1; k < n / 2; ++k) {
    h[2 * k] = phys[2 * k];
    h[2 * k + 1] = phys[2 * k + 1];
    h[2 * (n - k)] = phys[2 * k];
    h[2 * (n - k) + 1] = -phys[2 * k + 1];
  }
  h[2 * (n / 2)] = phys[n];
  int e = host_cfft(r, n, h, 1, 1);
  if (e) return e;
  for (int j = 0; j < n; ++j) {
    phys8[j] = h[2 * j];
    phys[j] = (float)h[2 * j];
  }
  phys8[n] = phys[n];
  phys8[n + 1] = phys[n + 1];
  return LOMA_OK;
}

// FP64 1-D y operators on one real column of my doubles: deryr
// (cfdiff.f:615), deryyr (cfdiff.f:676), inty8 (laps.f:484)
static int host_realop(loma_run* r, const double* u, double* out, int op) {
  if (!r || !u || !out) return LOMA_ERR_ARG;
  const int my = r->my;
  std::vector<double> h(2 * (size_t)my, 0.0);
  if (op == 3) {
    for (int j = 0; j < my; ++j) h[j] = u[j];
  } else {
    for (int j = 0; j < my; ++j) h[2 * j] = u[j];
  }
  double* d = nullptr;
  HIPCHK(hipMalloc(&d, 2 * h.size() * sizeof(double)));
  double* o = d + h.size();
  int rc = LOMA_OK;
  if (hipMemcpy(d, h.data(), h.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      (op == 3 ? launch_inty8(r->M, r->T, d, o, r->stream)
               : launch_dy(r->M, op, r->T, d, o, 1, r->stream)) ||
      hipStreamSynchronize(r->stream) != hipSuccess ||
      hipMemcpy(h.data(), o, h.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("host 1-D operator: device failure");
    rc = LOMA_ERR_DEVICE;
  }
  hipFree(d);
  if (rc == LOMA_OK) {
    if (op == 3) {
      for (int j = 0; j < my; ++j) out[j] = h[j];
    } else {
      for (int j = 0; j < my; ++j) out[j] = h[2 * j];
    }
  }
  return rc;
}
int loma_deryr(loma_run* r, const double* u, double* du) { return host_realop(r, u, du, 1); }
int loma_deryyr(loma_run* r, const double* u, double* du) { return host_realop(r, u, du, 2); }
int loma_inty8(loma_run* r, const double* du, double* u) { return host_realop(r, du, u, 3); }

int loma_fourz_dev(loma_run* r, double* c, int nrows, int iopt) {
  if (!r || !c || nrows < 0) return LOMA_ERR_ARG;
  LCHK(launch_fourz(r->plz, r->twz64, (double2*)c, nrows, iopt, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  return LOMA_OK;
}

// scratch of the batched banded solve, kept per device and grown on demand.
// One area per device serves every call on it, so a call holds the lock from
// the size check until its solve has finished: concurrent calls on a device
// run one after the other instead of sharing (or freeing) each other's area.
static std::mutex g_band_mu;
static std::vector<std::pair<void*, size_t>> g_band_scr;

int loma_band5_solve_dev(const double* band, void* rhs, double* x, int n, int nsys,
                         int rhs_fp32, void* stream) {
  if (!band || !rhs || n < 3 || nsys < 0 || (rhs_fp32 && !x)) {
    set_error("loma_band5_solve_dev: bad arguments (n >= 3; x required for float rhs)");
    return LOMA_ERR_ARG;
  }
  if (nsys == 0) return LOMA_OK;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  const size_t need = (size_t)n * 5 * nsys * sizeof(double);
  std::lock_guard<std::mutex> lk(g_band_mu);
  if ((int)g_band_scr.size() <= dev) g_band_scr.resize(dev + 1, {nullptr, 0});
  auto& e = g_band_scr[dev];
  if (e.second < need) {
    if (e.first) hipFree(e.first);
    e.first = nullptr;
    e.second = 0;
    if (hipMalloc(&e.first, need) != hipSuccess) {
      set_error("loma_band5_solve_dev: scratch allocation failed");
      return LOMA_ERR_OOM;
    }
    e.second = need;
  }
  double* out = x ? x : (double*)rhs;
  hipStream_t st = (hipStream_t)stream;
  LCHK(launch_band5(band, rhs, out, (double*)e.first, n, nsys, rhs_fp32, st));
  HIPCHK(hipStreamSynchronize(st));
  return LOMA_OK;
}

}  // extern "C"
