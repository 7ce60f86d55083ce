// loma_kernels.h -- launch interface between the host orchestration
// (loma_api.hip) and the kernels (loma_yops.hip, loma_fft.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "loma_dev.h"

namespace loma {

// RK3 coefficients (trunk/lomacte3/src/modulos.f:194-209)
struct RkCoef {
  double gama[3], xi[3], alpha[3], beta[3];
};
RkCoef rk_coefficients();

// device-resident scalars of one run
struct DevScalars {
  double dt;      // current step size (Deltat)
  double time;
  double dtr;     // next step size from the CFL estimate
  int nan;
  int pad;
  double dm;      // momentum thickness of the state at the start of the last step
  double dw;      // vorticity thickness at the start of its third substep
  double trapz;   // integral of drho00 over y at the last substep (cross.f:1034-1036)
  double mass[3]; // y-integrals of the 00 modes of rho u, rho v, rho w (cross.f:1013-1019)
};

// per-run physics constants passed by value to kernels
struct RunConst {
  int mgalx, my, mgalz, mx, mz, nplanes, nz;
  int myp;  // y stride of the transform buffers (my rounded up to 16)
  double alp, bet, ire, peclet;
  double hxalp, hzbet;   // 2/dx and 2/dz as in phys2 (cross.f:1500-1501)
  double rhobot, rhotop, zbot, ztop;
  double cfl;
  int dt_fp32;
};

// FFT plan for one length
struct FftPlan {
  int n;
  int npass;
  int rad[10];
};
FftPlan make_plan(int n);

// argument blocks of the fused phase kernels (pointers are device memory)
struct PhaseArgs {
  RunConst rc;
  YTab T;
  const double* helm_lu;     // per (i,|kz|) class LU tables (lane-major)
  const double* helm_bc;     // per class Neumann constants (4 doubles)
  const double* helm_h;      // per class h1, h2 (lane-major, 2 entries per row)
  int lu_stride;             // doubles per class LU table (5 * 64 * M)
  double2 *vor, *phi, *psi, *scal, *ten12, *ten13, *zf;
  void* spec[12];            // S-sized transform buffers (float2 or double2)
  double *u00wk, *v00wk, *w00wk, *rf0u, *rf0w;
  // mean split of the transforms: y-profiles of the 00 mode of each input
  // (phase A writes them, the FP32/FP64 transform carries zero there) and
  // the per-y estimates subtracted from each output ([12][my] each)
  double *mean_in, *mean_out;
  const double* hy;          // trapezoid half-spacings of y (main.f:689-696)
  DevScalars* sc;
  int rkstep;
  int with_z;
  int mode_base;             // first mode handled by this launch
  int nmodes;                // modes handled by this launch
};

// complex entries per thread of the z transforms (J * mgalz <= 256 * E)
#ifndef LOMA_ZF_E64
#define LOMA_ZF_E64 8
#endif
#define LOMA_ZF_E32 16

// device initial condition (loma_seed): seeded mixing-layer perturbation
struct IcArgs {
  RunConst rc;
  YTab T;
  const double* helm_lu;     // per (i,|kz|) class LU tables (lane-major)
  int lu_stride;
  const double2* coef;       // [3][nplanes*mz][nymax]: coefficient of e1^(n+1)
  const double* y;           // [my]
  double* partial;           // [nplanes*mz][3]: 2|a|^2 at the centre row (pass 1)
  double2 *vor, *phi, *psi, *scal, *zf;
  double s, env_width, k0;
  double scale[3];           // target rms / centre-plane rms of u, v, w (pass 2)
  int nymax;
  int pass;                  // 1: centre-plane energies, 2: the state
  int with_z;
};

struct FftArgs {
  RunConst rc;
  FftPlan plz, plx;
  const void* twz;           // twiddles exp(-2 pi i t/mgalz), T2
  const void* twx;
  void* spec[12];
  void* wbuf[12];            // z-physical transform inputs  [i][z][j]
  void* wout[12];            // z-physical transform outputs [i][z][j]
  double* rhst;
  double* rhsz;
  double *mean_in, *mean_out;  // [12][my], see PhaseArgs
  double *rbar, *rzbar;        // [2][my] mean-profile RHS of T and Z by substep parity
  const double* hy;          // my
  DevScalars* sc;
  // CFL maxima of each x-line tile, [tile][4]: max |u|, max |v|/hy, max |w|,
  // max diffusive inverse dt (written by the substep that takes the CFL)
  double* cflp;
  // per-line sums of 1/T over x, [mgalz][my] (statistics steps only; null
  // otherwise)
  double* stat_rho;
  int nfields;
  int rkstep;
  int do_cfl;
  int with_z;
  int J;                     // y rows per z-transform block
  int L;                     // lines per x-line block
};

// statistics window (addstats1/2 and the phys2 sums of cross.f): the
// accumulated columns, in the order of a .sta record, are
//   0 um 1 vm 2 wm 3 rum 4 rvm 5 rwm 6 Tm 7 w1m 8 w2m 9 w3m 10 them
//   11 up 12 vp 13 wp 14 uvr 15 uwr 16 vwr 17 w1p 18 w2p 19 w3p 20 Tp 21 ep
//   22 ruu 23 ruv 24 ruw 25 rvv 26 rvw 27 rww 28 rhom 29 thep 30 theup
//   31 mum 32 num
constexpr int STAT_NCOL = 33;
constexpr int STAT_NPART = 13;  // per-chunk partial sums: Tp, then the 12 of addstats2
constexpr int STAT_KACH = 4;    // |kz| classes per statistics wavefront
struct StatsArgs {
  RunConst rc;
  YTab T;
  const void* spec[12];
  const double *mean_in, *mean_out;
  const int* jsp;            // spectra rows (0-based)
  int nsp;
  int kach;                  // |kz| classes per wavefront (a chunk of one kx plane)
  int nch;                   // chunks per kx plane: ceil((nz+1) / kach)
  int nw;                    // statistics wavefronts: nplanes * nch
  double* acc;               // [STAT_NCOL][my]
  double* spe;               // [nplanes][12][nsp][nz+1]
  double* part;              // [STAT_NPART][my][nw], this step's sums per chunk
  const double* rho;         // [mgalz][my] line sums of 1/T from the x lines
  double* last;              // [9][my] this step's plane-summed intensities (uner)
};

// launchers (return hipError_t as int)
int launch_dy(int M, int order, const YTab& T, const double* u, double* du, int ncols,
              hipStream_t st);
int launch_helm_factor(const double* dt21, const double* dt22, int my, int M,
                       const double* shifts, int nsys, double* lu, hipStream_t st);
// per-class Neumann constants (bc, 4 per class) and, when h is not null,
// the homogeneous solutions h1 = LU^-1 e_0, h2 = LU^-1 e_{my-1} of each
// class (lane-major, 2 entries per row, 2*64*M doubles per class)
int launch_helm_bc(int M, const YTab& T, const double* lu, double* bc, double* h, int nclass,
                   hipStream_t st);
int launch_lapvdvhom(int M, const YTab& T, const double* lu, const double* bc,
                     const double* phi, double* v, double* dv, const double* rK,
                     int ncols, int mode, hipStream_t st);
int launch_inty8(int M, const YTab& T, const double* du, double* u, hipStream_t st);
int launch_visc(int M, const YTab& T, const double* lu1, const double* lu2, const double* bc2,
                double* vorwk, double* vor, double* phi, double* v, double* dvdy,
                double* phiwk, const double* rk2, double dalbe, double dtri, double dtxi,
                int ncols, hipStream_t st);

int launch_phaseA(int M, bool fp32, const PhaseArgs& a, hipStream_t st);
int launch_phaseC(int M, bool fp32, const PhaseArgs& a, hipStream_t st);
int launch_phaseC2(int M, bool fp32, const PhaseArgs& a, hipStream_t st);
int launch_ic(int M, const IcArgs& a, hipStream_t st);
int launch_thickness(int M, bool fp32, const PhaseArgs& a, const double* hy, double ly,
                     int which, hipStream_t st);
// x-line tiles of a grid (mgalz * ceil(my / L)); at most cfl_tiles_max,
// since L >= 2
__host__ __device__ inline int xl_tiles(int my, int mgalz, int L) {
  return mgalz * ((my + L - 1) / L);
}
__host__ __device__ inline int cfl_tiles_max(int my, int mgalz) { return xl_tiles(my, mgalz, 2); }
int launch_finish_step(DevScalars* sc, RunConst rc, int icfl, const double* cflp, int ntile,
                       hipStream_t st);
// which = 1: addstats1 on the transform inputs (after phase A); 2: addstats2
// on the outputs (after phase C2); 3: the reduction of the step's partial
// sums into the accumulated columns
int launch_stats(int M, bool fp32, int which, const StatsArgs& a, hipStream_t st);

// ---- batched launches over several runs (ensembles) ----------------------
// One launch serves a class of runs that share a kernel instantiation: the
// launch's blocks are the concatenation of each run's own grid, run r
// owning blocks first[r] .. first[r+1]-1, and the kernel reads run r's
// argument block from device memory (args[r]).
struct FinishArgs {
  DevScalars* sc;
  RunConst rc;
  const double* cflp;
  int icfl;
  int ntile;
};
__device__ __forceinline__ int batch_run(const int* __restrict__ first, int nrun, int b) {
  int r = 0;
  for (int q = 1; q < nrun; ++q)
    if (b >= first[q]) r = q;
  return r;
}
// which: 0 phase A, 1 phase C, 2 phase C2 (all runs of one M and precision)
int launch_phase_batch(int M, bool fp32, int which, const PhaseArgs* args, const int* first,
                       int nrun, int nblocks, hipStream_t st);
// proto: host copy of one run's arguments (length, tiles, fields of the class)
int launch_zfft_batch(bool fp32, bool inv, const FftArgs& proto, const FftArgs* args,
                      const int* first, int nrun, int nblocks, hipStream_t st);
int launch_xline_batch(bool fp32, const FftArgs& proto, const FftArgs* args, const int* first,
                       int nrun, int nblocks, hipStream_t st);
int launch_finish_batch(const FinishArgs* args, int nrun, hipStream_t st);
// blocks of one run's launch of each kernel (the grid the single-run
// launchers use)
int phase_blocks(int which, const RunConst& rc);
int zfft_blocks(const FftArgs& a);
int xline_blocks(const FftArgs& a);

int launch_zfft_inv(bool fp32, const FftArgs& a, hipStream_t st);
int launch_xline(bool fp32, const FftArgs& a, hipStream_t st);
int launch_zfft_fwd(bool fp32, const FftArgs& a, hipStream_t st);
int launch_fourz(const FftPlan& pl, const double2* tw, double2* c, int nrows, int iopt,
                 hipStream_t st);
int launch_band5(const double* band, const void* rhs, double* x, double* scratch, int n,
                 int nsys, int rhs_fp32, hipStream_t st);

}  // namespace loma
