/*
 * loma_amd.h -- C ABI of the MI355X-native LoMa RK3 time-advance library
 * (libloma_hip.so).  Plain pointers and sizes only.
 *
 * Each entry point replaces one routine of the reference solver; the
 * replaced routine is cited as path:line relative to the reference tree.
 * Conventions shared with the reference:
 *   - complex y-columns are interleaved (re,im) pairs, like the reference's
 *     real(4) u(2,my,m) arguments (trunk/lomacte3/src/cfdiff.f:518)
 *   - spectral fields are [kx plane i][kz mode k][y j], y fastest, kz stored
 *     0..nz then -nz..-1 (trunk/lomacte3/src/cross.f:77-79)
 *   - band matrices are (nb,my) Fortran arrays, like module matrices
 *     (trunk/lomacte3/src/modulos.f:121-128)
 * Every function returns an int status (LOMA_OK on success).  Errors never
 * abort the process; the Fortran shim turns a nonzero status into `stop`.
 */
#ifndef LOMA_AMD_H
#define LOMA_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LOMA_OK 0
#define LOMA_ERR_ARG 1     /* bad size / argument */
#define LOMA_ERR_NAN 2     /* non-finite value detected in the state */
#define LOMA_ERR_OOM 3     /* device allocation failed */
#define LOMA_ERR_DEVICE 4  /* HIP runtime error */
#define LOMA_ERR_STATE 5   /* call order (e.g. advance before upload) */
#define LOMA_ERR_NODEV 6   /* no GPU visible */

/* precision modes */
#define LOMA_FP64 0   /* everything in double */
#define LOMA_MIXED 1  /* FP32 transforms, nonlinear products and the
                         y-derivatives of the vorticity/phi right-hand sides;
                         FP64 Helmholtz/Poisson solves, 00 modes and RK
                         accumulation */

/* flags */
#define LOMA_FLAG_DT_FP32 1   /* round dt to float, as the tem module stores it
                                 (trunk/lomacte3/src/modulos.f:87-90) */
#define LOMA_FLAG_SCALAR_Z 2  /* transport a passive scalar Z
                                 (branches/lomaHZB/src/cross.f:2736) */
#define LOMA_FLAG_AUTOTUNE 4  /* time the transform tile shapes for this grid
                                 at init and keep the fastest (no reference
                                 counterpart: a tuning knob of this library) */
#define LOMA_FLAG_HISTORY 8   /* compute the momentum and vorticity thickness
                                 on the device every step (the .cf history,
                                 trunk/lomacte3/src/cross.f:692-699,1414-1429) */

typedef struct loma_run loma_run;

/* ---- library ---------------------------------------------------------- */
const char *loma_version(void);
const char *loma_last_error(void);   /* message of the last failing call */
int loma_device_count(int *n);

/* Compact-FD operators for a y grid (derivadas, cfdiff.f:869-941).  Host
 * only; no GPU needed.  Outputs (nb,my) Fortran-ordered like module matrices:
 * dt11 = A1 (7-band LHS of d/deta), dt12 = B1 (7-band RHS), dt21 = A2 and
 * dt22 = B2 (5-band second derivative), fmap = 1/(dy/deta). */
int loma_fd_matrices(int my, const double *y, double *dt11, double *dt12,
                     double *dt21, double *dt22, double *fmap);

/* ---- run lifecycle ---------------------------------------------------- */
/* initcr (trunk/lomacte3/src/main.f:461-800) + cross1 set-up
 * (cross.f:100-160): grid, FD operators, wavenumbers, device buffers.
 * mgalx, mgalz: physical grid; my: y points; y[my]: coordinates; Re, alp,
 * bet: Reynolds number and fundamental wavenumbers (hre.dat:40); pr: Prandtl
 * (main.f:569 uses 0.7); cfl, ncfl: CFL number and update period
 * (hre.dat:68); dt0: first time step (main.f:559 uses 1e-4). */
int loma_init(loma_run **run, int mgalx, int my, int mgalz, const double *y,
              double Re, double alp, double bet, double pr, double cfl,
              int ncfl, double dt0, int precision, int flags);
int loma_free(loma_run *run);
/* sizes: out[0..7] = mx, mz, nplanes, nz, S (complex per field),
 * W (complex per work field), P (physical points), bytes of device memory */
int loma_dims(const loma_run *run, long long *out);

/* Restart-boundary transfers (getfil main.f:329 / escru escru.f:11).
 * Fields are complex [nplanes][mz][my]; fp64 selects double (1) or float (0)
 * host arrays (the reference's restart stores float).  u00,v00,w00 are the
 * rho*u, rho*v, rho*w mean profiles (double, my).  Upload resets the RK
 * history and computes the boundary densities (cross.f:208-213). */
int loma_upload(loma_run *run, const void *vor, const void *phi,
                const void *psi, const void *scal, const double *u00,
                const double *v00, const double *w00, int fp64);
int loma_download(loma_run *run, void *vor, void *phi, void *psi, void *scal,
                  double *u00, double *v00, double *w00, int fp64);
/* passive scalar Z field (LOMA_FLAG_SCALAR_Z runs only) */
int loma_upload_z(loma_run *run, const void *z, int fp64);
int loma_download_z(loma_run *run, void *z, int fp64);

/* Seeded initial condition built on the device, in place of an upload:
 * the temporal mixing layer of trunk/initloma3/src/main.f:95-609
 * (perturbation :123-291, rho00/U/Z profiles :360-373, psi and the
 * spectral state :452-463).  coef holds the seeded draws, complex
 * (interleaved doubles) [3][nplanes][mz][nymax]: for u, v, w the
 * coefficient amp*exp(2 pi i phs) of exp(i (n+1) (2 kh / nymax) y); s is the
 * density ratio, amp the target rms of each velocity component on the
 * centre plane, env_width the Gaussian y-window width, k0 the peak
 * wavenumber of the spectrum shape (kh/k0)^2 exp(-(kh/k0)^2).  Then as
 * loma_upload: history reset, boundary densities, 00 modes (and Z if the
 * run has it).  The draws are staged in a transform work buffer, 48*nymax
 * bytes per mode; my >= 48 always fits, smaller grids may be refused with
 * LOMA_ERR_ARG.  The reference seeds from the wall clock
 * (main.f:36-45); the caller's seeded generator replaces it. */
int loma_seed(loma_run *run, const double *coef, int nymax, double s,
              double amp, double env_width, double k0);

/* Momentum and vorticity thickness of a LOMA_FLAG_HISTORY run: out[0] =
 * delta_m of the state at the start of the last step (the reference takes it
 * at rkstep 1 from the x-z mean velocity, cross.f:1414-1429, with the
 * trapezoid weights of main.f:689-696), out[1] = delta_w = 1/max|dU/dy| at
 * the start of that step's third substep (rkstep 3, cross.f:692-699).  Both
 * are 0 until a step has run; LOMA_ERR_STATE without the flag. */
int loma_thickness(loma_run *run, double *out);
/* The columns of the reference's .cf history line that do not need the
 * statistics window (cross.f:1166-1177): out[0..7] = time, dm, dw,
 * Redw = dw*Re, trapz (the y-integral of the 00 mode of the constraint
 * source at the last substep, cross.f:1034-1036), massu, massv, massw (the
 * y-integrals of the 00 modes of rho u, rho v, rho w with the hy weights,
 * cross.f:1013-1019).  dm, dw and Redw are NaN without LOMA_FLAG_HISTORY. */
int loma_history(loma_run *run, double *out);

/* Statistics window (replaces the istati/addstats machinery of cross1:
 * addstats1 cross.f:1981-2063 at the transform inputs, the phys2 line sums
 * cross.f:1586-1607, addstats2 cross.f:2070-2287 at the transform outputs,
 * writestats escru.f:276-484 and resetstats escru.f:485-560).
 * loma_stats_config: the nspec spectra rows jspecy (1-based y indices, each
 * in 1..my; the reference's ctes3D jspecy), to which the centre row
 * (my+1)/2 is appended as main.f:698-703 does; allocates the accumulators
 * (once per call) and resets them.  loma_stats_accumulate(on): while on,
 * rkstep 1 of every step adds to the window (istati = 1, cross.f:236-237);
 * such steps run without the captured step graph, and loma_advance_many
 * refuses the run.  loma_stats_read: sta = my x 33 rows of the .sta record
 * (um vm wm rum rvm rwm Tm w1m w2m w3m them, up vp wp uvr uwr vwr w1p w2p
 * w3p Tp ep, then ruu ruv ruw rvv rvw rww rhom thep theup mum num times
 * 1/(mgalx mgalz)); spe = nplanes records of 12 x (nspec+1) x (nz+1)
 * floats, the .spe order spwk(0:nz1, 1:nspec+1, 12, plane); info[13] =
 * nacum, nacumsp, masst, enerdis, energy(9) (the .cf columns 6..16,
 * cross.f:1152-1158, 811-816; escru.f:372-376).  Any output may be NULL.
 * loma_stats_reset: resetstats (zeros, nacum = 0). */
int loma_stats_config(loma_run *run, int nspec, const int *jspecy);
int loma_stats_accumulate(loma_run *run, int on);
int loma_stats_read(loma_run *run, double *sta, float *spe, double *info);
int loma_stats_reset(loma_run *run);

/* Time advance: nsteps full RK3 steps (cross.f:225-1223 loop body),
 * device resident.  Returns LOMA_ERR_NAN if the state became non-finite. */
int loma_advance(loma_run *run, int nsteps);
/* One RK substep (rkstep 1..3) of the current step; loma_advance is three of
 * these plus the time / dt update (cross.f:1196-1204). */
int loma_substep(loma_run *run, int rkstep);
/* scalars: out[0..4] = dt, time, dtr (next dt from CFL), istep, nan flag */
int loma_scalars(loma_run *run, double *out);
/* Drop-in for cross1's time loop (cross.f:24): upload, nstep steps,
 * download, with the reference's argument meaning (float fields, double 00
 * modes ordered u00,w00,v00 as in cross1's argument list). */
int loma_cross1(loma_run *run, float *vor, float *phi, float *psi, float *scal,
                double *u00, double *w00, double *v00, int nstep);
/* Enqueue nsteps without waiting (no host synchronisation); loma_sync waits
 * for the run's stream and reports LOMA_ERR_NAN like loma_advance. */
int loma_advance_async(loma_run *run, int nsteps);
int loma_sync(loma_run *run);
/* HIP stream the run launches on (0 = default), and graph capture of one
 * full step (replayed by loma_advance when enabled). */
int loma_set_stream(loma_run *run, void *stream);
int loma_stream(loma_run *run, void **stream);
int loma_use_graph(loma_run *run, int enable);
/* Per-kernel timing with HIP events recorded on the run's stream around
 * every launch (graph replay is bypassed while enabled).  profile_read
 * returns, per kernel class (0 phase A, 1 z inverse FFT, 2 x lines,
 * 3 z forward FFT, 4 phase C, 5 end of step, 6 phase C psi solve), the
 * summed milliseconds and
 * launch counts since the last read. */
int loma_profile(loma_run *run, int enable);
int loma_profile_read(loma_run *run, double *total_ms, int *counts, int nclasses);

/* ---- per-operator entry points (device pointers, double) --------------- */
/* All act on ncols complex y-columns of length my stored back to back
 * (interleaved re/im doubles), in place allowed, on the run's stream. */
int loma_deryr2_dev(loma_run *run, const double *u, double *du, int ncols);  /* cfdiff.f:518 */
int loma_deryyr2_dev(loma_run *run, const double *u, double *du, int ncols); /* cfdiff.f:735 */
/* rK per column (double[ncols]) */
int loma_lapvdvhom_dev(loma_run *run, const double *phi, double *v,
                       double *dvdy, const double *rK, int ncols);          /* laps.f:319 */
int loma_solvepsi_dev(loma_run *run, const double *phi, double *v,
                      double *dvdy, const double *rK, int ncols);           /* laps.f:580 */
/* Helmholtz with extra shift (implicit viscous kernel of visc,
 * trunk/mixing/src/laps.f:535): (D2 - (rK+shift)) x = f, x = 0 at ends */
int loma_helmholtz_dev(loma_run *run, const double *f, double *x,
                       const double *rK, double shift, int ncols);
/* inty8 (laps.f:484) on one real column of my doubles */
int loma_inty8_dev(loma_run *run, const double *du, double *u);
/* visc (trunk/mixing/src/laps.f:535-857) with the mixing-layer boundary
 * functions (laps.f:893-1060), for ncols mode columns: implicit viscous
 * step of Omega_y and phi (Helmholtz shift rk1[col], Dirichlet) and v
 * ((D2 - rk2[col]) v = phi, sqrt(rk2) v -+ dv/dy = 0 at the bottom/top).
 * In/out roles as the reference's: vorwk in forcing, out omega; phiwk in
 * forcing, out phi; v and dvdy in the history terms, out the new v, dv/dy;
 * vor, phi out dalbe*solution + dtri*forcing + dtxi*(v, dvdy) (dalbe != 0)
 * or the solution.  rk1, rk2: device double[ncols]. */
int loma_visc_dev(loma_run *run, double *vorwk, double *vor, double *phi,
                  double *v, double *dvdy, double *phiwk, const double *rk1,
                  const double *rk2, double dalbe, double dtri, double dtxi,
                  int ncols);

/* ---- per-operator entry points (host pointers, float, reference shapes) */
int loma_deryr2(loma_run *run, const float *u, float *du, int m);   /* cfdiff.f:518 */
int loma_deryyr2(loma_run *run, const float *u, float *du, int m);  /* cfdiff.f:735 */
int loma_lapvdvhom(loma_run *run, const float *phi, float *v, float *dvdy,
                   double rK);                                      /* laps.f:319 */
int loma_solvepsi(loma_run *run, const float *phi, float *v, float *dvdy,
                  double rK);                                       /* laps.f:580 */
/* FP64 1-D operators on one real column of my doubles, host pointers:
 * the first and second derivative and the y-integral with u(1) = 0 */
int loma_deryr(loma_run *run, const double *u, double *du);         /* cfdiff.f:615 */
int loma_deryyr(loma_run *run, const double *u, double *du);        /* cfdiff.f:676 */
int loma_inty8(loma_run *run, const double *du, double *u);         /* laps.f:484 */

/* ---- transforms (device pointers) ------------------------------------- */
/* fourz (fou3D.f:80) on nrows complex rows of length mgalz, FFTPACK
 * convention: iopt<0 forward scaled by 1/n with e^{-i}, iopt>=0 inverse. */
int loma_fourz_dev(loma_run *run, double *c, int nrows, int iopt);
/* foup2 (cross.f:1282-1430, with phys2 cross.f:1440-1828): the transform
 * stage of a substep with the reference's argument list (the work array,
 * statistics array and rank are dropped).  Arrays are the reference's
 * complex*8 (my, mgalz, nplanes) blocks with the spectral content in the
 * first mz kz slots of each plane.  On entry up1..up3 = rho u, rho v,
 * rho w, ten22 = lap T, ten12, ten13, ten23 = dT/dx, dT/dy, dT/dz,
 * drho = T; on return up1..up3 = u, v, w, ten11, ten22, ten33 = rho uu,
 * rho vv, rho ww, ten12, ten13, ten23 = rho uv, rho uw, rho vw, drho = the
 * constraint source, tnext = DT.  rhst is the real*8 (mgalx, mgalz*my)
 * physical RHS history of T (line z + mgalz*y), updated.  dt is the
 * substep's step size (Deltat); at rkstep 1 *dtr (if not NULL) receives
 * CFL/(convective + diffusive maxima) of one rank.  Uses the run's device
 * buffers, so the run must be uploaded again before it advances.  Runs
 * without scalar Z only. */
int loma_foup2(loma_run *run, float *up1, float *up2, float *up3, float *ten11,
               float *ten22, float *ten33, double *rhst, float *drho, float *tnext,
               float *ten12, float *ten13, float *ten23, int rkstep, double dt,
               double *dtr);                                         /* cross.f:1282 */

/* ---- transforms with the reference's shapes (host pointers, real(4)) ---
 * FFTPACK conventions: iopt < 0 forward e^{-i} scaled by 1/n, else inverse
 * e^{+i} unscaled; computed on the device in FP64, stored back as float. */
/* cft: m complex vectors of length mgalz; element j of vector q at
 * c[q*isa + j*ise] (real part) and c[q*isa + j*ise + 1] (imaginary part) */
int loma_cft(loma_run *run, float *c, int ise, int isa, int m, int iopt); /* cftsingle.f:2 */
/* rft: m real vectors of length mgalx at stride isa >= mgalx + 2, with the
 * packed coefficients c[2k] + i c[2k+1], k = 0..mgalx/2 (rftsingle.f:64-73) */
int loma_rft(loma_run *run, float *c, int isa, int m, int iopt);  /* rftsingle.f:25 */
/* fourz: a plane of my rows of mgalz complex numbers (cft with ise = 2,
 * isa = 2*mgalz) */
int loma_fourz(loma_run *run, float *phys, int iopt);             /* fou3D.f:80 */
/* fourx: one x-line of mgalx + 2 reals; iopt >= 0 zeroes the modes past mx
 * and writes the physical line to phys and (in FP64) to phys8 */
int loma_fourx(loma_run *run, float *phys, double *phys8, int iopt); /* fou3D.f:29 */
/* visc with the reference's argument list and real(4) (2,my) columns
 * (trunk/mixing/src/laps.f:535; the trailing integer flag is unused) */
int loma_visc(loma_run *run, float *vorwk, float *vor, float *phi, float *v,
              float *dvdy, float *phiwk, double rk1, double rk2, double dalbe,
              double dtri, double dtxi);

/* ---- batched banded solves in isolation ------------------------------- */
/* nsys independent pentadiagonal systems of size n >= 3 (row-major band
 * [sys][n][5], entry d of row j is column j-2+d) with complex right-hand
 * sides rhs[sys][n] (interleaved; float when rhs_fp32, else double).  LU
 * without pivoting (bandec, cfdiff.f:977) and the two triangular sweeps
 * (banbks, cfdiff.f:954), FP64.  The solution goes to x (double,
 * interleaved); x may be NULL for a double rhs, which is then overwritten.
 * Device pointers; stream may be NULL.  Scratch of n*5*nsys doubles is kept
 * per device and grown on demand. */
int loma_band5_solve_dev(const double *band, void *rhs, double *x, int n,
                         int nsys, int rhs_fp32, void *stream);
/* loma_band5_solve_dev without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernel and results; the scratch is a fixed
 * array in the library's code object, 4,194,304 doubles (32 MiB) that every
 * GPU the library runs on holds until the process ends, so n*5*nsys must not
 * exceed 4,194,304 (LOMA_ERR_ARG otherwise; e.g. 1635 systems at n = 513).
 * Calls that overlap on one GPU share that array and race. */
int loma_band5_solve_dev_noalloc(const double *band, void *rhs, double *x, int n,
                                 int nsys, int rhs_fp32, void *stream);

/* Re-target a run at new physics without re-allocating: Reynolds number,
 * Prandtl number, CFL number and first time step (the grid, precision and
 * flags stay).  The run must be uploaded again before it advances.  For a
 * request loop that keeps finished runs of a grid for the next request of
 * the same grid.  No reference counterpart (initcr, main.f:461-800, sets
 * these once per job). */
int loma_set_physics(loma_run *run, double Re, double pr, double cfl, double dt0);

/* Device checkpoint of a run's time-advance state: the spectral fields, the
 * RK histories (RHS of Omega_y and phi, the physical RHS of T and Z), the
 * 00-mode work arrays, time, dt and the step count.  loma_snapshot copies
 * them into a device buffer the run keeps (allocated on the first call,
 * about 7 S-sized complex fields plus the physical histories, outside
 * loma_footprint); loma_restore copies them back.  For restart-on-NaN: the
 * reference's initnan (trunk/mixing/src/main.f:953) re-reads the restart
 * file and calls cross1 again; here the run returns to its last snapshot on
 * the device. */
int loma_snapshot(loma_run *run);
int loma_restore(loma_run *run);
/* Change the CFL number of a running run (the state, histories and time
 * are kept; captured step graphs are rebuilt on the next step). */
int loma_set_cfl(loma_run *run, double cfl);

/* Transform tile shapes in use: y entries per z-transform workgroup (J) and
 * lines per x-line workgroup (L).  No reference counterpart. */
int loma_tiles(const loma_run *run, int *J, int *L);

/* ---- ensembles -------------------------------------------------------- */
/* Advance several independent runs (grids, Re, density ratios may differ)
 * nsteps full RK3 steps each.  With graphs enabled on every run, all runs'
 * steps are captured into one hipGraph with one independent branch per run
 * (each on the run's own stream) and one graph launch advances the whole
 * ensemble by a step; otherwise the runs' kernels are enqueued on their own
 * streams.  Replaces a driver loop of cross1 calls over runs (no reference
 * counterpart for the batching itself). */
int loma_advance_many(loma_run **runs, int nruns, int nsteps);

/* ---- serving ---------------------------------------------------------- */
/* Device bytes that loma_init would allocate for this grid, precision and
 * flags (the sum of its requested buffer sizes, equal to the last entry of
 * loma_dims after init), without touching the device: the admission test of
 * a request loop.  No reference counterpart (the reference sizes its arrays
 * at compile time, ctes3D). */
int loma_footprint(int mgalx, int my, int mgalz, int precision, int flags,
                   long long *bytes);

/* One RK3 step in flight per run, for a request loop that interleaves many
 * runs: loma_step_begin enqueues one step on the run's stream (its graph
 * when graphs are enabled) and returns at once; loma_step_poll reports
 * without blocking whether that step has finished (*done), its device time
 * in ms and the run's non-finite flag.  Beginning a step while one is in
 * flight, or polling with none, is LOMA_ERR_STATE.  These replace one
 * iteration of the cross1 step loop (cross.f:225-1226) for a caller that
 * does not wait. */
int loma_step_begin(loma_run *run);
int loma_step_poll(loma_run *run, int *done, double *device_ms, int *nonfinite);

#ifdef __cplusplus
}
#endif
#endif
