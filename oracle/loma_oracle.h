/*
 * loma_oracle.h -- CPU FP64 restatement of the LoMa RK3 time advance.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP product path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links it.
 *
 * Every function restates one routine of the reference solver
 * (trunk/lomacte3/src unless stated otherwise) in double precision
 * ("FP64-promoted" semantics: the reference's real(4)/complex*8 state,
 * wavenumbers and time step are carried in double here).
 *
 * Array conventions follow the reference:
 *   - band matrices are Fortran (nb,my) arrays: a[j*nb + d], d = 0..nb-1
 *   - a complex y-column is interleaved (re,im) doubles, exactly like the
 *     reference's real(4) u(2,my) argument (cfdiff.f:518)
 *   - spectral fields are complex [i (kx plane)][k (kz mode)][j (y)],
 *     y fastest, kz stored 0..nz then -nz..-1 (cross.f:77-79).
 */
#ifndef LOMA_ORACLE_H
#define LOMA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid orc_grid;
typedef struct orc_state orc_state;

/* grid + compact-FD setup: initcr main.f:555-800, derivadas cfdiff.f:869 */
orc_grid *orc_grid_create(int mgalx, int my, int mgalz, const double *y,
                          double Re, double alp, double bet, double pr);
void orc_grid_destroy(orc_grid *g);
void orc_grid_dims(const orc_grid *g, int *out /* mx,mz,nplanes,nz */);
/* copies prem1,dt11,dt12 (7*my each), prem3,dt21,dt22 (5*my each),
 * fmap,hy,trp,endder (my each) */
void orc_grid_matrices(const orc_grid *g, double *prem1, double *dt11,
                       double *dt12, double *prem3, double *dt21, double *dt22,
                       double *fmap, double *hy, double *trp, double *endder);

/* raw compact-scheme coefficient generators (before LU) */
void orc_pred1m6o(int ny, double *lhs7, double *rhs7);       /* cfdiff.f:24  */
void orc_prederiv2(int ny, const double *x, double *lhs5,
                   double *rhs5);                              /* cfdiff.f:196 */
int orc_gaussj(double *a, int n, double *b, int m);           /* cfdiff.f:1144 */

/* banded LU (cfdiff.f:954-1137); a is (nb,n) Fortran band storage */
void orc_bandec(double *a, int n);            /* 5-band */
void orc_banbks(const double *a, int n, double *b);
void orc_bandec7(double *a, int n);           /* 7-band */
void orc_banbks7(const double *a, int n, double *b);

/* column operators, complex columns interleaved (2,my,m) */
void orc_deryr2(const orc_grid *g, const double *u, double *du, int m);
void orc_deryyr2(const orc_grid *g, const double *u, double *du, int m);
void orc_deryr(const orc_grid *g, const double *u, double *du);
void orc_deryyr(const orc_grid *g, const double *u, double *du);
void orc_lapvdvhom(const orc_grid *g, const double *phi, double *v,
                   double *dvdy, double rK);
void orc_solvepsi(const orc_grid *g, const double *phi, double *v,
                  double *dvdy, double rK);
void orc_inty8(const orc_grid *g, const double *du, double *u);
/* inty8 with the reference's singular dt12 factorisation (pinning only) */
void orc_inty8_ref(const orc_grid *g, const double *du, double *u);
/* implicit viscous Helmholtz (trunk/mixing/src/laps.f:535 visc kernel):
 * solves (D2 - (rK + shift)) x = f with homogeneous Dirichlet ends */
void orc_helmholtz_dirichlet(const orc_grid *g, const double *f, double *x,
                             double rK, double shift);
/* visc (trunk/mixing/src/laps.f:535-857) with the mixing-layer boundary
 * functions (laps.f:893-1060): implicit viscous step of Omega_y (Helmholtz
 * with shift rk1, Dirichlet) and of phi/v (Helmholtz rk1 for phi, then
 * (D2 - rk2) v = phi with the 4x4 boundary system).  Complex columns (2,my);
 * arguments and their in/out roles as the reference's. */
void orc_visc(const orc_grid *g, double *vorwk, double *vor, double *phi, double *v,
              double *dvdy, double *phiwk, double rk1, double rk2, double dalbe,
              double dtri, double dtxi);

/* FFTPACK-convention transforms (cftsingle.f:2, rftsingle.f:25).
 * cft: m complex vectors of length n, stride between vectors = n complex
 * rft: m real vectors of n+2 doubles, packing of rftsingle.f:64-96 */
void orc_cft(double *c, int n, int m, int iopt);
void orc_rft(double *c, int n, int m, int iopt);

/* full time advance (cross.f:24-1259) */
orc_state *orc_state_create(const orc_grid *g, double dt0, double cfl,
                            int ncfl, int flags);
void orc_state_destroy(orc_state *st);
/* complex fields interleaved, reference layout [i][k][j]; 00 modes my each */
void orc_state_set(orc_state *st, const double *vor, const double *phi,
                   const double *psi, const double *scal, const double *u00,
                   const double *v00, const double *w00);
void orc_state_get(const orc_state *st, double *vor, double *phi, double *psi,
                   double *scal, double *u00, double *v00, double *w00);
/* history: ten12/ten13 (S complex), rhst (mgalx*my*mgalz), rf0u/rf0w (my) */
void orc_state_history(const orc_state *st, double *ten12, double *ten13,
                       double *rhst, double *rf0u, double *rf0w);
/* Reduce the CFL maxima as the reference does over `ranks` compute ranks
 * (cross.f:1778-1799, lines split as in change.f:188-201): the largest
 * per-rank sum of the convective maxima.  1 (the default) is one rank. */
void orc_state_set_cfl_ranks(orc_state *st, int ranks);
/* .cf history columns of the last substep: trapz, massu, massv, massw
 * (cross.f:1013-1019, 1034-1036) */
void orc_state_cf(const orc_state *st, double *out);
/* foup2 (cross.f:1282-1430) alone: 8 transform inputs -> 11 outputs with the
 * state's rhst, substep rkstep of step size dt; *dtr at rkstep 1 */
void orc_foup2(orc_state *st, const double *in, double *out, int rkstep, double dt,
               double *dtr);
/* set the physical RHS history of T ([y][z][x], cross.f:92) */
void orc_state_set_rhst(orc_state *st, const double *rhst);
int orc_step(orc_state *st, int nsteps);           /* returns 1 on NaN */
int orc_substep(orc_state *st, int rkstep);        /* one RK substep */
void orc_state_scalars(const orc_state *st, double *out /* dt,time,dtr,istep */);

/* statistics window (addstats1/2, phys2 sums, writestats): spectra rows
 * (0-based), accumulate on/off, read, reset */
int orc_stats_config(orc_state *st, int nsp, const int *rows);
void orc_stats_accumulate(orc_state *st, int on);
void orc_stats_get(const orc_state *st, double *sta /* [my][33] */,
                   double *spe /* [nplanes][12][nsp][nz+1] */,
                   double *info /* nacum, nacumsp, masst, enerdis, energy[9] */);
void orc_stats_reset(orc_state *st);

/* flags for orc_state_create */
#define ORC_DT_FP32 1   /* round dt to float like tem module (modulos.f:89) */
#define ORC_SCALAR_Z 2  /* transport a passive scalar Z (lomaHZB cross.f:2736) */

/* passive scalar Z (when ORC_SCALAR_Z): spectral field + physical history */
void orc_state_set_z(orc_state *st, const double *z);
void orc_state_get_z(const orc_state *st, double *z);

/* number of OpenMP threads the oracle uses (0 = runtime default) */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
