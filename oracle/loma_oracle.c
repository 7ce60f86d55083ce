/*
 * loma_oracle.c -- CPU FP64 restatement of the LoMa RK3 time advance.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py load this library; the HIP product path
 * never links or calls it.
 *
 * What it restates (paths relative to the reference tree):
 *   compact FD coefficients   trunk/lomacte3/src/cfdiff.f:24-501 (pred1m6o,
 *                             prederiv2), derivadas cfdiff.f:869-941
 *   banded LU                 cfdiff.f:954-1137 (banbks, bandec, banbks7,
 *                             bandec7), Gauss-Jordan cfdiff.f:1144
 *   y operators               deryr2 cfdiff.f:518, deryr :615, deryyr :676,
 *                             deryyr2 :735; laps.f Lapvdvhom :319,
 *                             inty8 :484, solvepsi :580, calcrhslap :17,
 *                             calcrhsd1 :87
 *   transforms                fou3D.f:29-189 (fourx, fourz, local
 *                             transposes), cftsingle.f:2-66, rftsingle.f:25-96
 *   physical kernel           cross.f:1440-1828 (phys2)
 *   transform pipeline        cross.f:1282-1430 (foup2)
 *   substep                   cross.f:307-1141 (cross1 RK loop body)
 *   RK coefficients           modulos.f:194-209
 *   grid / wavenumbers        main.f:555-570, 689-696, 769-800
 *   passive scalar Z          branches/lomaHZB/src/cross.f:2736-2777
 *                             (calcrhshz with sigma = 0, dZb/dZ = 1, Le = 1)
 *
 * Everything is carried in double precision ("FP64-promoted" semantics).
 */
#include "loma_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* grid                                                                      */
/* ------------------------------------------------------------------------ */
struct orc_grid {
  int mgalx, my, mgalz;
  int mx, mz, nplanes, nz, nz2; /* ctes3D:27-36 */
  double Re, alp, bet, peclet;
  double *y, *hy, *trp, *fmap, *endder;
  double *prem1, *dt11, *dt12; /* (7,my) */
  double *prem3, *dt21, *dt22; /* (5,my) */
  double *alp2, *bet2;         /* kx^2 (nplanes), kz^2 (mz) */
  double *kx, *kz;             /* signed wavenumbers */
};

/* RK3 coefficients, modulos.f:194-209 */
static const double RK_P0 = 5.0 / 66.0, RK_P1 = -1.0;
static const double RK_XI[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
static const double RK_GAMA[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static void rk_coefs(double *alpha, double *beta) {
  const double *g = RK_GAMA, *xi = RK_XI;
  double p1 = RK_P1;
  double b3 = (p1 * (g[0] * g[2] + g[1] * g[2] + xi[1] * g[2] + g[0] * xi[2]) -
               g[0] * g[1] * g[2] - g[0] * g[1] * xi[2]) /
              (p1 - g[0] * g[1]);
  beta[0] = g[0] - RK_P0;
  beta[1] = p1;
  beta[2] = b3;
  alpha[0] = RK_P0;
  alpha[1] = g[1] + xi[1] - p1;
  alpha[2] = g[2] + xi[2] - beta[2];
}

/* Gauss-Jordan elimination with full pivoting (the method of gaussj,
 * cfdiff.f:1144).  a is n x n column-major, b is n x m column-major. */
int orc_gaussj(double *a, int n, double *b, int m) {
  int *ipiv = calloc(n, sizeof(int)), *ir = malloc(n * sizeof(int)),
      *ic = malloc(n * sizeof(int));
#define A(r, c) a[(r) + (size_t)(c) * n]
#define B(r, c) b[(r) + (size_t)(c) * n]
  int bad = 0;
  for (int it = 0; it < n; it++) {
    double big = 0.0;
    int irow = -1, icol = -1;
    for (int j = 0; j < n; j++) {
      if (ipiv[j] == 1) continue;
      for (int k = 0; k < n; k++) {
        if (ipiv[k] == 0 && fabs(A(j, k)) >= big) {
          big = fabs(A(j, k));
          irow = j;
          icol = k;
        }
      }
    }
    if (icol < 0) { bad = 1; break; }
    ipiv[icol]++;
    if (irow != icol) {
      for (int l = 0; l < n; l++) { double t = A(irow, l); A(irow, l) = A(icol, l); A(icol, l) = t; }
      for (int l = 0; l < m; l++) { double t = B(irow, l); B(irow, l) = B(icol, l); B(icol, l) = t; }
    }
    ir[it] = irow;
    ic[it] = icol;
    if (A(icol, icol) == 0.0) { bad = 1; break; }
    double pinv = 1.0 / A(icol, icol);
    A(icol, icol) = 1.0;
    for (int l = 0; l < n; l++) A(icol, l) *= pinv;
    for (int l = 0; l < m; l++) B(icol, l) *= pinv;
    for (int ll = 0; ll < n; ll++) {
      if (ll == icol) continue;
      double d = A(ll, icol);
      A(ll, icol) = 0.0;
      for (int l = 0; l < n; l++) A(ll, l) -= A(icol, l) * d;
      for (int l = 0; l < m; l++) B(ll, l) -= B(icol, l) * d;
    }
  }
  if (!bad) {
    for (int l = n - 1; l >= 0; l--) {
      if (ir[l] != ic[l]) {
        for (int k = 0; k < n; k++) {
          double t = A(k, ir[l]); A(k, ir[l]) = A(k, ic[l]); A(k, ic[l]) = t;
        }
      }
    }
  }
#undef A
#undef B
  free(ipiv); free(ir); free(ic);
  return bad;
}

/* pred1m6o, cfdiff.f:24-182.  lhs7/rhs7 are (ny,7) column arrays, i.e.
 * lhs7[d*ny + j]: d = 0..6 <-> (den,ben,aln,one,alp,bep,dep) and
 * rhs7 <-> (a3n,a2n,a1n,a0,a1p,a2p,a3p). */
void orc_pred1m6o(int ny, double *L, double *R) {
  const double aaa = 0.56926818603810, bbb = 0.31997939759919,
               ccc = 0.03751544855390;
  const double alpha = -3.0 / 4.0 + (33.0 / 20.0) * ccc +
                       (239.0 / 180.0) * aaa + (82.0 / 45.0) * bbb;
  const double beta = 3.0 / 10.0 + (24.0 / 25.0) * ccc -
                      (88.0 / 225.0) * aaa + (34.0 / 225.0) * bbb;
  const double delta = -1.0 / 2e1 + (39.0 / 100.0) * ccc +
                       (19.0 / 300.0) * aaa + (2.0 / 75.0) * bbb;
  const double h = 2.0 / (double)(ny - 1);
#define LL(d, j) L[(d) * ny + (j)]
#define RR(d, j) R[(d) * ny + (j)]
  for (int j = 0; j < ny; j++) LL(3, j) = 1.0;
  for (int j = 3; j < ny - 3; j++) {
    LL(0, j) = delta; LL(1, j) = beta; LL(2, j) = alpha;
    LL(4, j) = alpha; LL(5, j) = beta; LL(6, j) = delta;
    RR(0, j) = -ccc / h; RR(1, j) = -bbb / h; RR(2, j) = -aaa / h;
    RR(3, j) = 0.0;
    RR(4, j) = aaa / h; RR(5, j) = bbb / h; RR(6, j) = ccc / h;
  }
  /* first three rows (cfdiff.f:83-130) */
  double l0[7] = {0, 0, 0, 1, 1, 0, 0};
  double r0[7] = {0, 0, 0, -2.0 / h, 2.0 / h, 0, 0};
  double l1[7] = {0, 0, 0.25, 1, 0.25, 0, 0};
  double r1[7] = {0, 0, -3.0 / 4.0 / h, 0, 3.0 / 4.0 / h, 0, 0};
  double l2[7] = {0, 1.0 / 36.0, 4.0 / 9.0, 1, 4.0 / 9.0, 1.0 / 36.0, 0};
  double r2[7] = {0, -25.0 / 216.0 / h, -20.0 / 27.0 / h, 0,
                  20.0 / 27.0 / h, 25.0 / 216.0 / h, 0};
  double *ls[3] = {l0, l1, l2}, *rs[3] = {r0, r1, r2};
  for (int b = 0; b < 3; b++)
    for (int d = 0; d < 7; d++) { LL(d, b) = ls[b][d]; RR(d, b) = rs[b][d]; }
  /* mirrored last three rows (cfdiff.f:134-179): LHS reversed, RHS
   * reversed and negated */
  for (int b = 0; b < 3; b++) {
    int j = ny - 1 - b;
    for (int d = 0; d < 7; d++) {
      LL(d, j) = ls[b][6 - d];
      RR(d, j) = -rs[b][6 - d];
    }
  }
#undef LL
#undef RR
}

/* prederiv2, cfdiff.f:196-501.  lhs5/rhs5 are (ny,5) column arrays:
 * lhs d = 0..4 <-> (ben,aln,1,alp,bep), rhs <-> (a2n,a1n,a0,a1p,a2p). */
void orc_prederiv2(int ny, const double *x, double *L, double *R) {
#define LL(d, j) L[(d) * ny + (j)]
#define RR(d, j) R[(d) * ny + (j)]
  double c[81], b[9];
  for (int j = 0; j < ny; j++) LL(2, j) = 1.0;
  /* interior: unknowns (ben,aln,alp,bep,a2n,a1n,a0,a1p,a2p); Taylor
   * conditions of orders 0..8 (cfdiff.f:216-270) */
  for (int j = 2; j < ny - 2; j++) {
    double d[5] = {x[j - 2] - x[j], x[j - 1] - x[j], 0.0, x[j + 1] - x[j],
                   x[j + 2] - x[j]};
    double e[4] = {d[0], d[1], d[3], d[4]}; /* LHS offsets */
    memset(c, 0, sizeof c);
#define C(r, k) c[(r) + 9 * (k)]
    for (int k = 4; k < 9; k++) C(0, k) = 1.0;
    b[0] = 0.0;
    for (int k = 0; k < 5; k++) C(1, 4 + k) = d[k];
    b[1] = 0.0;
    for (int k = 0; k < 4; k++) C(2, k) = 2.0;
    for (int k = 0; k < 5; k++) C(2, 4 + k) = -d[k] * d[k];
    b[2] = -2.0;
    for (int n = 3; n <= 8; n++) {
      for (int k = 0; k < 4; k++) C(n, k) = n * (n - 1) * pow(e[k], n - 2);
      for (int k = 0; k < 5; k++) C(n, 4 + k) = (k == 2) ? 0.0 : -pow(d[k], n);
      b[n] = 0.0;
    }
    orc_gaussj(c, 9, b, 1);
    LL(0, j) = b[0]; LL(1, j) = b[1]; LL(3, j) = b[2]; LL(4, j) = b[3];
    RR(0, j) = b[4]; RR(1, j) = b[5]; RR(2, j) = b[6]; RR(3, j) = b[7];
    RR(4, j) = b[8];
#undef C
  }
  /* boundary rows use 5x5 / 7x7 systems embedded in a 9-leading-dim array */
  double cc[81];
#define C(r, k) cc[(r) + n_ * (k)]
  {
    /* j = 1 (cfdiff.f:278-325): unknowns (alp,bep,a0,a1p,a2p) */
    int j = 0, n_ = 5;
    double d1 = x[j + 1] - x[j], d2 = x[j + 2] - x[j];
    memset(cc, 0, sizeof cc);
    C(0, 2) = 1; C(0, 3) = 1; C(0, 4) = 1; b[0] = 0;
    C(1, 3) = d1; C(1, 4) = d2; b[1] = 0;
    C(2, 0) = 2; C(2, 1) = 2; C(2, 3) = -d1 * d1; C(2, 4) = -d2 * d2; b[2] = -2;
    for (int n = 3; n <= 4; n++) {
      C(n, 0) = n * (n - 1) * pow(d1, n - 2);
      C(n, 1) = n * (n - 1) * pow(d2, n - 2);
      C(n, 3) = -pow(d1, n); C(n, 4) = -pow(d2, n); b[n] = 0;
    }
    orc_gaussj(cc, 5, b, 1);
    LL(0, j) = 0; LL(1, j) = 0; LL(3, j) = b[0]; LL(4, j) = b[1];
    RR(0, j) = 0; RR(1, j) = 0; RR(2, j) = b[2]; RR(3, j) = b[3]; RR(4, j) = b[4];
  }
  {
    /* j = 2 (cfdiff.f:328-376): unknowns (aln,alp,bep,a1n,a0,a1p,a2p) */
    int j = 1, n_ = 7;
    double dm = x[j - 1] - x[j], d1 = x[j + 1] - x[j], d2 = x[j + 2] - x[j];
    memset(cc, 0, sizeof cc);
    C(0, 3) = 1; C(0, 4) = 1; C(0, 5) = 1; C(0, 6) = 1; b[0] = 0;
    C(1, 3) = dm; C(1, 5) = d1; C(1, 6) = d2; b[1] = 0;
    C(2, 0) = 2; C(2, 1) = 2; C(2, 2) = 2;
    C(2, 3) = -dm * dm; C(2, 5) = -d1 * d1; C(2, 6) = -d2 * d2; b[2] = -2;
    for (int n = 3; n <= 6; n++) {
      C(n, 0) = n * (n - 1) * pow(dm, n - 2);
      C(n, 1) = n * (n - 1) * pow(d1, n - 2);
      C(n, 2) = n * (n - 1) * pow(d2, n - 2);
      C(n, 3) = -pow(dm, n); C(n, 5) = -pow(d1, n); C(n, 6) = -pow(d2, n);
      b[n] = 0;
    }
    orc_gaussj(cc, 7, b, 1);
    LL(0, j) = 0; LL(1, j) = b[0]; LL(3, j) = b[1]; LL(4, j) = b[2];
    RR(0, j) = 0; RR(1, j) = b[3]; RR(2, j) = b[4]; RR(3, j) = b[5]; RR(4, j) = b[6];
  }
  {
    /* j = ny-1 (cfdiff.f:380-428): unknowns (ben,aln,alp,a2n,a1n,a0,a1p) */
    int j = ny - 2, n_ = 7;
    double dmm = x[j - 2] - x[j], dm = x[j - 1] - x[j], d1 = x[j + 1] - x[j];
    memset(cc, 0, sizeof cc);
    C(0, 3) = 1; C(0, 4) = 1; C(0, 5) = 1; C(0, 6) = 1; b[0] = 0;
    C(1, 3) = dmm; C(1, 4) = dm; C(1, 6) = d1; b[1] = 0;
    C(2, 0) = 2; C(2, 1) = 2; C(2, 2) = 2;
    C(2, 3) = -dmm * dmm; C(2, 4) = -dm * dm; C(2, 6) = -d1 * d1; b[2] = -2;
    for (int n = 3; n <= 6; n++) {
      C(n, 0) = n * (n - 1) * pow(dmm, n - 2);
      C(n, 1) = n * (n - 1) * pow(dm, n - 2);
      C(n, 2) = n * (n - 1) * pow(d1, n - 2);
      C(n, 3) = -pow(dmm, n); C(n, 4) = -pow(dm, n); C(n, 6) = -pow(d1, n);
      b[n] = 0;
    }
    orc_gaussj(cc, 7, b, 1);
    LL(0, j) = b[0]; LL(1, j) = b[1]; LL(3, j) = b[2]; LL(4, j) = 0;
    RR(0, j) = b[3]; RR(1, j) = b[4]; RR(2, j) = b[5]; RR(3, j) = b[6]; RR(4, j) = 0;
  }
  {
    /* j = ny (cfdiff.f:432-478): unknowns (aln,ben,a0,a1n,a2n) */
    int j = ny - 1, n_ = 5;
    double dm = x[j - 1] - x[j], dmm = x[j - 2] - x[j];
    memset(cc, 0, sizeof cc);
    C(0, 2) = 1; C(0, 3) = 1; C(0, 4) = 1; b[0] = 0;
    C(1, 3) = dm; C(1, 4) = dmm; b[1] = 0;
    C(2, 0) = 2; C(2, 1) = 2; C(2, 3) = -dm * dm; C(2, 4) = -dmm * dmm; b[2] = -2;
    for (int n = 3; n <= 4; n++) {
      C(n, 0) = n * (n - 1) * pow(dm, n - 2);
      C(n, 1) = n * (n - 1) * pow(dmm, n - 2);
      C(n, 3) = -pow(dm, n); C(n, 4) = -pow(dmm, n); b[n] = 0;
    }
    orc_gaussj(cc, 5, b, 1);
    LL(0, j) = b[1]; LL(1, j) = b[0]; LL(3, j) = 0; LL(4, j) = 0;
    RR(0, j) = b[4]; RR(1, j) = b[3]; RR(2, j) = b[2]; RR(3, j) = 0; RR(4, j) = 0;
  }
#undef C
#undef LL
#undef RR
}

/* ------------------------------------------------------------------------ */
/* banded LU without pivoting in the reference's packed storage              */
/* (cfdiff.f:954-1137): after factorisation a(1,k) = 1/pivot, a(2..p+1,k) = */
/* U superdiagonals, a(p+2..2p+1,k) = L multipliers for rows k+1..k+p.      */
/* ------------------------------------------------------------------------ */
static void band_factor(double *a, int n, int p) {
  const int nb = 2 * p + 1;
  /* left-align the first p rows (cfdiff.f:984-989, 1063-1073) */
  for (int r = 0; r < p && r < n; r++) {
    int sh = p - r;
    for (int d = 0; d + sh < nb; d++) a[r * nb + d] = a[r * nb + d + sh];
  }
  for (int k = 0; k < n; k++) {
    double inv = 1.0 / a[k * nb + 0];
    a[k * nb + 0] = inv;
    if (k == n - 1) break;
    double mult[4];
    for (int i = 1; i <= p && k + i < n; i++) {
      double *row = a + (size_t)(k + i) * nb;
      double l = row[0] * inv;
      mult[i] = l;
      /* shift row left by one while eliminating column k */
      for (int d = 0; d < nb - 1; d++) {
        double ukd = (d + 1 <= p) ? a[k * nb + d + 1] : 0.0;
        if (d + 1 <= p)
          row[d] = row[d + 1] - l * ukd;
        else
          row[d] = row[d + 1];
      }
    }
    for (int i = 1; i <= p && k + i < n; i++) a[k * nb + p + i] = mult[i];
  }
}

static void band_solve(const double *a, int n, int p, double *b) {
  const int nb = 2 * p + 1;
  for (int k = 0; k < n - 1; k++)
    for (int i = 1; i <= p && k + i < n; i++) b[k + i] -= a[k * nb + p + i] * b[k];
  for (int i = n - 1; i >= 0; i--) {
    double s = b[i];
    for (int d = 1; d <= p && i + d < n; d++) s -= a[i * nb + d] * b[i + d];
    b[i] = s * a[i * nb + 0];
  }
}

void orc_bandec(double *a, int n) { band_factor(a, n, 2); }
void orc_banbks(const double *a, int n, double *b) { band_solve(a, n, 2, b); }
void orc_bandec7(double *a, int n) { band_factor(a, n, 3); }
void orc_banbks7(const double *a, int n, double *b) { band_solve(a, n, 3, b); }

/* banded matvec out = B*u for a (nb,n) band B (row j uses u[j-p..j+p]) */
static void band_matvec(const double *B, int n, int p, const double *u,
                        int ustride, double *out) {
  const int nb = 2 * p + 1;
  for (int j = 0; j < n; j++) {
    double s = 0.0;
    int first = 1;
    for (int d = 0; d < nb; d++) {
      int c = j - p + d;
      if (c < 0 || c >= n) continue;
      double t = B[j * nb + d] * u[(size_t)c * ustride];
      if (first) { s = t; first = 0; } else s += t;
    }
    out[j] = s;
  }
}

/* ------------------------------------------------------------------------ */
orc_grid *orc_grid_create(int mgalx, int my, int mgalz, const double *y,
                          double Re, double alp, double bet, double pr) {
  orc_grid *g = calloc(1, sizeof *g);
  g->mgalx = mgalx; g->my = my; g->mgalz = mgalz;
  g->mx = 2 * (mgalx / 3);
  g->nplanes = g->mx / 2;
  g->mz = 2 * (mgalz / 3) - 1;
  g->nz = (g->mz - 1) / 2;
  g->nz2 = mgalz - g->nz;
  g->Re = Re; g->alp = alp; g->bet = bet;
  g->peclet = Re * pr; /* main.f:569 */
  g->y = malloc(my * sizeof(double));
  memcpy(g->y, y, my * sizeof(double));
  g->hy = malloc(my * sizeof(double));
  g->trp = malloc(my * sizeof(double));
  g->fmap = malloc(my * sizeof(double));
  g->endder = malloc(my * sizeof(double));
  g->prem1 = malloc(7 * my * sizeof(double));
  g->dt11 = malloc(7 * my * sizeof(double));
  g->dt12 = malloc(7 * my * sizeof(double));
  g->prem3 = malloc(5 * my * sizeof(double));
  g->dt21 = malloc(5 * my * sizeof(double));
  g->dt22 = malloc(5 * my * sizeof(double));
  /* derivadas, cfdiff.f:869-941 */
  double *d11 = calloc(7 * my, sizeof(double)), *d12 = calloc(7 * my, sizeof(double));
  double *d21 = calloc(5 * my, sizeof(double)), *d22 = calloc(5 * my, sizeof(double));
  orc_pred1m6o(my, d11, d12);
  orc_prederiv2(my, y, d21, d22);
  for (int j = 0; j < my; j++) {
    for (int i = 0; i < 7; i++) {
      g->prem1[j * 7 + i] = d11[i * my + j];
      g->dt11[j * 7 + i] = d11[i * my + j];
      g->dt12[j * 7 + i] = d12[i * my + j];
    }
    for (int i = 0; i < 5; i++) {
      g->prem3[j * 5 + i] = d21[i * my + j];
      g->dt21[j * 5 + i] = d21[i * my + j];
      g->dt22[j * 5 + i] = d22[i * my + j];
    }
  }
  free(d11); free(d12); free(d21); free(d22);
  orc_bandec7(g->prem1, my);
  orc_bandec(g->prem3, my);
  /* fmap = 1 / (d y / d eta) (cfdiff.f:912-937) */
  band_matvec(g->dt12, my, 3, y, 1, g->fmap);
  orc_banbks7(g->prem1, my, g->fmap);
  for (int j = 0; j < my; j++) g->fmap[j] = 1.0 / g->fmap[j];
  /* endderiv, cfdiff.f:819-863 (transpose solve; diagnostics only) */
  {
    double *f = calloc(my, sizeof(double));
    const double *p = g->prem1;
#define P(d, j) p[(j) * 7 + (d)]
    f[0] = P(0, 0);
    f[1] = -P(0, 1) * P(1, 0) * f[0];
    f[2] = -P(0, 2) * (P(1, 1) * f[1] + P(2, 0) * f[0]);
    for (int j = 3; j < my; j++)
      f[j] = -P(0, j) * (P(1, j - 1) * f[j - 1] + P(2, j - 2) * f[j - 2] +
                         P(3, j - 3) * f[j - 3]);
    f[my - 2] = f[my - 2] - P(4, my - 2) * f[my - 1];
    f[my - 3] = f[my - 3] - P(4, my - 3) * f[my - 2] - P(5, my - 3) * f[my - 1];
    for (int j = my - 4; j >= 0; j--)
      f[j] = f[j] - P(4, j) * f[j + 1] - P(5, j) * f[j + 2] - P(6, j) * f[j + 3];
#undef P
    for (int j = 0; j < my; j++) {
      double s = 0.0;
      int lo = j - 3 < 0 ? 0 : j - 3, hi = j + 3 > my - 1 ? my - 1 : j + 3;
      for (int i = lo; i <= hi; i++) s += f[i] * g->dt12[i * 7 + (j - i + 3)];
      g->endder[j] = s * g->fmap[0];
    }
    free(f);
  }
  /* trapezoid weights, main.f:689-696 */
  for (int j = 1; j < my - 1; j++) {
    g->trp[j] = 0.25 * (y[j + 1] - y[j - 1]);
    g->hy[j] = (y[j + 1] - y[j - 1]) / 2.0;
  }
  g->trp[0] = 0.25 * (y[1] - y[0]);
  g->trp[my - 1] = 0.25 * (y[my - 1] - y[my - 2]);
  g->hy[0] = (y[1] - y[0]) / 2.0;
  g->hy[my - 1] = (y[my - 1] - y[my - 2]) / 2.0;
  /* wavenumbers, main.f:769-800 */
  g->alp2 = malloc(g->nplanes * sizeof(double));
  g->kx = malloc(g->nplanes * sizeof(double));
  g->bet2 = malloc(g->mz * sizeof(double));
  g->kz = malloc(g->mz * sizeof(double));
  for (int i = 0; i < g->nplanes; i++) {
    g->kx[i] = alp * i;
    g->alp2[i] = g->kx[i] * g->kx[i];
  }
  for (int k = 0; k < g->mz; k++) {
    g->kz[k] = (k <= g->nz) ? bet * k : -bet * (g->mz - k);
    g->bet2[k] = g->kz[k] * g->kz[k];
  }
  return g;
}

void orc_grid_destroy(orc_grid *g) {
  if (!g) return;
  free(g->y); free(g->hy); free(g->trp); free(g->fmap); free(g->endder);
  free(g->prem1); free(g->dt11); free(g->dt12);
  free(g->prem3); free(g->dt21); free(g->dt22);
  free(g->alp2); free(g->bet2); free(g->kx); free(g->kz);
  free(g);
}

void orc_grid_dims(const orc_grid *g, int *out) {
  out[0] = g->mx; out[1] = g->mz; out[2] = g->nplanes; out[3] = g->nz;
}

void orc_grid_matrices(const orc_grid *g, double *prem1, double *dt11,
                       double *dt12, double *prem3, double *dt21, double *dt22,
                       double *fmap, double *hy, double *trp, double *endder) {
  int my = g->my;
  if (prem1) memcpy(prem1, g->prem1, 7 * my * sizeof(double));
  if (dt11) memcpy(dt11, g->dt11, 7 * my * sizeof(double));
  if (dt12) memcpy(dt12, g->dt12, 7 * my * sizeof(double));
  if (prem3) memcpy(prem3, g->prem3, 5 * my * sizeof(double));
  if (dt21) memcpy(dt21, g->dt21, 5 * my * sizeof(double));
  if (dt22) memcpy(dt22, g->dt22, 5 * my * sizeof(double));
  if (fmap) memcpy(fmap, g->fmap, my * sizeof(double));
  if (hy) memcpy(hy, g->hy, my * sizeof(double));
  if (trp) memcpy(trp, g->trp, my * sizeof(double));
  if (endder) memcpy(endder, g->endder, my * sizeof(double));
}

/* ------------------------------------------------------------------------ */
/* y operators                                                               */
/* ------------------------------------------------------------------------ */
/* first derivative of one real column with stride (deryr, cfdiff.f:615) */
static void dy1_col(const orc_grid *g, const double *u, int stride,
                    double *du, int ostride) {
  int my = g->my;
  double *w = malloc(my * sizeof(double));
  band_matvec(g->dt12, my, 3, u, stride, w);
  orc_banbks7(g->prem1, my, w);
  for (int j = 0; j < my; j++) du[(size_t)j * ostride] = w[j] * g->fmap[j];
  free(w);
}
/* second derivative (deryyr, cfdiff.f:676) */
static void dy2_col(const orc_grid *g, const double *u, int stride,
                    double *du, int ostride) {
  int my = g->my;
  double *w = malloc(my * sizeof(double));
  band_matvec(g->dt22, my, 2, u, stride, w);
  orc_banbks(g->prem3, my, w);
  for (int j = 0; j < my; j++) du[(size_t)j * ostride] = w[j];
  free(w);
}

void orc_deryr(const orc_grid *g, const double *u, double *du) { dy1_col(g, u, 1, du, 1); }
void orc_deryyr(const orc_grid *g, const double *u, double *du) { dy2_col(g, u, 1, du, 1); }

/* deryr2 cfdiff.f:518: complex columns (2,my,m); in place allowed */
void orc_deryr2(const orc_grid *g, const double *u, double *du, int m) {
  int my = g->my;
  double *t = malloc(2 * my * sizeof(double));
  for (int k = 0; k < m; k++) {
    const double *uk = u + (size_t)k * 2 * my;
    dy1_col(g, uk, 2, t, 2);
    dy1_col(g, uk + 1, 2, t + 1, 2);
    memcpy(du + (size_t)k * 2 * my, t, 2 * my * sizeof(double));
  }
  free(t);
}
void orc_deryyr2(const orc_grid *g, const double *u, double *du, int m) {
  int my = g->my;
  double *t = malloc(2 * my * sizeof(double));
  for (int k = 0; k < m; k++) {
    const double *uk = u + (size_t)k * 2 * my;
    dy2_col(g, uk, 2, t, 2);
    dy2_col(g, uk + 1, 2, t + 1, 2);
    memcpy(du + (size_t)k * 2 * my, t, 2 * my * sizeof(double));
  }
  free(t);
}

/* LHS of the Helmholtz/Poisson solves, laps.f:349-366: (dt22 - rK dt21)
 * with identity rows at both ends, LU factored */
static double *helm_lhs(const orc_grid *g, double rK) {
  int my = g->my;
  double *a = malloc(5 * my * sizeof(double));
  for (int j = 1; j < my - 1; j++)
    for (int i = 0; i < 5; i++) a[j * 5 + i] = g->dt22[j * 5 + i] - rK * g->dt21[j * 5 + i];
  for (int i = 0; i < 5; i++) { a[i] = 0.0; a[(my - 1) * 5 + i] = 0.0; }
  a[2] = 1.0;
  a[(my - 1) * 5 + 2] = 1.0;
  orc_bandec(a, my);
  return a;
}
/* calcrhslap, laps.f:17-46: dt21*f with zero end rows */
static void rhslap(const orc_grid *g, const double *f, int stride, double *out) {
  band_matvec(g->dt21, g->my, 2, f, stride, out);
  out[0] = 0.0;
  out[g->my - 1] = 0.0;
}

/* Lapvdvhom, laps.f:319-477 */
void orc_lapvdvhom(const orc_grid *g, const double *phi, double *v,
                   double *dvdy, double rK) {
  int my = g->my;
  if (rK == 0.0) {
    memset(v, 0, 2 * my * sizeof(double));
    memset(dvdy, 0, 2 * my * sizeof(double));
    return;
  }
  double *a = helm_lhs(g, rK);
  double *f = malloc(2 * my * sizeof(double));
  for (int c = 0; c < 2; c++) {
    double *fc = f + c * my;
    rhslap(g, phi + c, 2, fc);
    orc_banbks(a, my, fc);
  }
  for (int c = 0; c < 2; c++) {
    for (int j = 0; j < my; j++) v[2 * j + c] = f[c * my + j];
    dy1_col(g, f + c * my, 1, dvdy + c, 2);
  }
  free(f);
  free(a);
}

/* solvepsi, laps.f:580-730: Poisson with d/dy = 0 at both ends via two
 * homogeneous solutions and a 2x2 Gauss-Jordan */
void orc_solvepsi(const orc_grid *g, const double *phi, double *v,
                  double *dvdy, double rK) {
  int my = g->my;
  if (rK == 0.0) {
    memset(v, 0, 2 * my * sizeof(double));
    memset(dvdy, 0, 2 * my * sizeof(double));
    return;
  }
  double *a = helm_lhs(g, rK);
  double *buf = malloc(8 * my * sizeof(double));
  double *vpr = buf, *vpi = buf + my, *v1 = buf + 2 * my, *v2 = buf + 3 * my;
  double *dvpr = buf + 4 * my, *dvpi = buf + 5 * my, *dv1 = buf + 6 * my,
         *dv2 = buf + 7 * my;
  rhslap(g, phi, 2, vpr);
  rhslap(g, phi + 1, 2, vpi);
  memset(v1, 0, my * sizeof(double));
  memset(v2, 0, my * sizeof(double));
  v1[0] = 1.0;
  v2[my - 1] = 1.0;
  orc_banbks(a, my, vpr);
  orc_banbks(a, my, vpi);
  orc_banbks(a, my, v1);
  orc_banbks(a, my, v2);
  /* calcrhsd1 + banbks7 (laps.f:696-705): derivative without fmap */
  double *src[4] = {vpr, vpi, v1, v2}, *dst[4] = {dvpr, dvpi, dv1, dv2};
  for (int q = 0; q < 4; q++) {
    band_matvec(g->dt12, my, 3, src[q], 1, dst[q]);
    orc_banbks7(g->prem1, my, dst[q]);
  }
  /* BCs dpsi/dy = 0 at both ends (laps.f:631-637, 707-721) */
  const double Cvbot = 0.0, Cdyvbot = 1.0, BCbot = 0.0;
  const double Cvtop = 0.0, Cdyvtop = 1.0, BCtop = 0.0;
  double AA[4], coef[4];
  AA[0] = Cvbot + Cdyvbot * dv1[0] * g->fmap[0];          /* (1,1) */
  AA[2] = Cdyvbot * dv2[0] * g->fmap[0];                  /* (1,2) */
  AA[1] = Cdyvtop * dv1[my - 1] * g->fmap[my - 1];        /* (2,1) */
  AA[3] = Cvtop + Cdyvtop * dv2[my - 1] * g->fmap[my - 1]; /* (2,2) */
  coef[0] = BCbot - Cdyvbot * dvpr[0] * g->fmap[0];
  coef[1] = BCtop - Cdyvtop * dvpr[my - 1] * g->fmap[my - 1];
  coef[2] = BCbot - Cdyvbot * dvpi[0] * g->fmap[0];
  coef[3] = BCtop - Cdyvtop * dvpi[my - 1] * g->fmap[my - 1];
  orc_gaussj(AA, 2, coef, 2);
  for (int j = 0; j < my; j++) {
    v[2 * j] = vpr[j] + coef[0] * v1[j] + coef[1] * v2[j];
    v[2 * j + 1] = vpi[j] + coef[2] * v1[j] + coef[3] * v2[j];
    dvdy[2 * j] = g->fmap[j] * (dvpr[j] + coef[0] * dv1[j] + coef[1] * dv2[j]);
    dvdy[2 * j + 1] = g->fmap[j] * (dvpi[j] + coef[2] * dv1[j] + coef[3] * dv2[j]);
  }
  free(buf);
  free(a);
}

/* inty8, laps.f:484-575, exactly as the reference computes it: LU of dt12
 * (bandec7) and u = dt12^{-1} dt11 (du/fmap) - (value at row 1).  dt12 has
 * the constant vector in its null space (each of its rows sums to zero), so
 * the last pivot of this LU is rounding noise (|1/pivot| ~ 1e11-1e12) and the
 * result carries cancellation noise of relative size ~eps/|pivot|. */
void orc_inty8_ref(const orc_grid *g, const double *du, double *u) {
  int my = g->my;
  double *w = malloc(7 * my * sizeof(double));
  memcpy(w, g->dt12, 7 * my * sizeof(double));
  orc_bandec7(w, my);
  double *f = malloc(my * sizeof(double)), *v = malloc(my * sizeof(double));
  for (int j = 0; j < my; j++) f[j] = du[j] / g->fmap[j];
  band_matvec(g->dt11, my, 3, f, 1, v);
  orc_banbks7(w, my, v);
  double v0 = v[0];
  for (int j = 0; j < my; j++) u[j] = v[j] - v0;
  free(f); free(v); free(w);
}

/* inty8 in the well-conditioned form used by the time advance: the same
 * function in exact arithmetic.  The reference's solve (see orc_inty8_ref)
 * returns x~ + C*1 where x~ solves rows 1..my-1 of dt12 x = dt11 f with
 * x(my) = 0, and C (huge) is removed by the final subtraction.  Here the
 * last row of dt12 is replaced by the gauge x(my) = 0, so no huge C arises. */
void orc_inty8(const orc_grid *g, const double *du, double *u) {
  int my = g->my;
  double *w = malloc(7 * my * sizeof(double));
  memcpy(w, g->dt12, 7 * my * sizeof(double));
  for (int d = 0; d < 7; d++) w[(my - 1) * 7 + d] = 0.0;
  w[(my - 1) * 7 + 3] = 1.0;
  orc_bandec7(w, my);
  double *f = malloc(my * sizeof(double)), *v = malloc(my * sizeof(double));
  for (int j = 0; j < my; j++) f[j] = du[j] / g->fmap[j];
  band_matvec(g->dt11, my, 3, f, 1, v);
  v[my - 1] = 0.0;
  orc_banbks7(w, my, v);
  double v0 = v[0];
  for (int j = 0; j < my; j++) u[j] = v[j] - v0;
  free(f); free(v); free(w);
}

/* Implicit viscous Helmholtz kernel of the incompressible solver's visc
 * (trunk/mixing/src/laps.f:535-600): (dt22 - (rK+shift) dt21) x =
 * dt21 f with homogeneous Dirichlet ends.  Complex column (2,my). */
void orc_helmholtz_dirichlet(const orc_grid *g, const double *f, double *x,
                             double rK, double shift) {
  int my = g->my;
  double *a = helm_lhs(g, rK + shift);
  double *t = malloc(my * sizeof(double));
  for (int c = 0; c < 2; c++) {
    rhslap(g, f + c, 2, t);
    orc_banbks(a, my, t);
    for (int j = 0; j < my; j++) x[2 * j + c] = t[j];
  }
  free(t);
  free(a);
}

/* Boundary functions of visc for the mixing layer (trunk/mixing/src/
 * laps.f:893-1060): phi/v rows  C_phi*phi + C_vvv*v + C_dyv*dv/dy = BC  (bot1,
 * bot2 at y(1); top1, top2 at y(my)); omega rows C_ome*ome + C_dyo*dome/dy. */
typedef struct { double phi, vvv, dyv, bc; } visc_row;
static visc_row visc_bot1(double rk2) { visc_row r = {0.0, sqrt(rk2), -1.0, 0.0}; return r; }
static visc_row visc_bot2(double rk2) { visc_row r = {1.0, 0.0, 0.0, 0.0}; (void)rk2; return r; }
static visc_row visc_top1(double rk2) { visc_row r = {0.0, sqrt(rk2), 1.0, 0.0}; return r; }
static visc_row visc_top2(double rk2) { visc_row r = {1.0, 0.0, 0.0, 0.0}; (void)rk2; return r; }
static const double VISC_OME = 1.0, VISC_DYO = 0.0, VISC_BCO = 0.0;  /* bot and top */

/* calcrhsd1 + banbks7(prem1): d/deta without fmap (laps.f:61-111) */
static void d1raw(const orc_grid *g, const double *u, double *du) {
  band_matvec(g->dt12, g->my, 3, u, 1, du);
  orc_banbks7(g->prem1, g->my, du);
}

/* visc, trunk/mixing/src/laps.f:535-857 */
void orc_visc(const orc_grid *g, double *vorwk, double *vor, double *phi, double *v,
              double *dvdy, double *phiwk, double rk1, double rk2, double dalbe,
              double dtri, double dtxi) {
  const int my = g->my, L = my - 1;
  const double f0 = g->fmap[0], fL = g->fmap[L];
  double *a1 = helm_lhs(g, rk1);  /* wk1 with rk1, laps.f:596-610 */
  double *buf = calloc((size_t)16 * my, sizeof(double));
  double *phipr = buf, *phipi = buf + my, *phi1 = buf + 2 * my, *phi2 = buf + 3 * my;
  double *vpr = buf + 4 * my, *vpi = buf + 5 * my, *v1 = buf + 6 * my, *v2 = buf + 7 * my;
  double *v3 = buf + 8 * my, *v4 = buf + 9 * my, *dvpr = buf + 10 * my, *dvpi = buf + 11 * my;
  double *dv1 = buf + 12 * my, *dv2 = buf + 13 * my, *dv3 = buf + 14 * my, *dv4 = buf + 15 * my;
  /* omega (laps.f:615-695) */
  for (int j = 0; j < 2 * my; j++) vor[j] = vorwk[j];
  rhslap(g, vorwk, 2, phipr);
  rhslap(g, vorwk + 1, 2, phipi);
  phi1[0] = 1.0;
  phi2[L] = 1.0;
  orc_banbks(a1, my, phipr);
  orc_banbks(a1, my, phipi);
  orc_banbks(a1, my, phi1);
  orc_banbks(a1, my, phi2);
  d1raw(g, phipr, dvpr);
  d1raw(g, phipi, dvpi);
  d1raw(g, phi1, dv1);
  d1raw(g, phi2, dv2);
  {
    double AA[4], co[4];
    AA[0] = VISC_OME + VISC_DYO * dv1[0] * f0;   /* AA(1,1) */
    AA[1] = VISC_DYO * dv1[L] * fL;              /* AA(2,1) */
    AA[2] = VISC_DYO * dv2[0] * f0;              /* AA(1,2) */
    AA[3] = VISC_OME + VISC_DYO * dv2[L] * fL;   /* AA(2,2) */
    co[0] = VISC_BCO - VISC_DYO * dvpr[0] * f0;
    co[1] = VISC_BCO - VISC_DYO * dvpr[L] * fL;
    co[2] = VISC_BCO - VISC_DYO * dvpi[0] * f0;
    co[3] = VISC_BCO - VISC_DYO * dvpi[L] * fL;
    orc_gaussj(AA, 2, co, 2);
    for (int j = 0; j < my; j++) {
      vorwk[2 * j] = phipr[j] + co[0] * phi1[j] + co[1] * phi2[j];
      vorwk[2 * j + 1] = phipi[j] + co[2] * phi1[j] + co[3] * phi2[j];
    }
  }
  for (int j = 0; j < 2 * my; j++)
    vor[j] = dalbe != 0.0 ? dalbe * vorwk[j] + dtri * vor[j] + dtxi * v[j] : vorwk[j];
  /* phi and v (laps.f:700-856) */
  for (int j = 0; j < 2 * my; j++) phi[j] = phiwk[j];
  rhslap(g, phiwk, 2, phipr);
  rhslap(g, phiwk + 1, 2, phipi);
  orc_banbks(a1, my, phipr);
  orc_banbks(a1, my, phipi);
  double *a2 = helm_lhs(g, rk2);  /* wk1 with rk2, laps.f:730-742 */
  rhslap(g, phipr, 1, vpr);
  rhslap(g, phipi, 1, vpi);
  rhslap(g, phi1, 1, v1);
  rhslap(g, phi2, 1, v2);
  memset(v3, 0, my * sizeof(double));
  memset(v4, 0, my * sizeof(double));
  v3[0] = 1.0;
  v4[L] = 1.0;
  double *vs[6] = {vpr, vpi, v1, v2, v3, v4}, *ds[6] = {dvpr, dvpi, dv1, dv2, dv3, dv4};
  for (int q = 0; q < 6; q++) orc_banbks(a2, my, vs[q]);
  for (int q = 0; q < 6; q++) d1raw(g, vs[q], ds[q]);
  {
    /* rows bot1, bot2, top1, top2; columns phi1, phi2, v3, v4 (laps.f:793-816) */
    const visc_row rb1 = visc_bot1(rk2), rb2 = visc_bot2(rk2);
    const visc_row rt1 = visc_top1(rk2), rt2 = visc_top2(rk2);
    double AA[16], co[8];
#define AAM(r, c) AA[(r) + 4 * (c)]
    AAM(0, 0) = rb1.phi + rb1.dyv * dv1[0] * f0;
    AAM(0, 1) = rb1.dyv * dv2[0] * f0;
    AAM(0, 2) = rb1.vvv + rb1.dyv * dv3[0] * f0;
    AAM(0, 3) = rb1.dyv * dv4[0] * f0;
    AAM(1, 0) = rb2.phi + rb2.dyv * dv1[0] * f0;
    AAM(1, 1) = rb2.dyv * dv2[0] * f0;
    AAM(1, 2) = rb2.vvv + rb2.dyv * dv3[0] * f0;
    AAM(1, 3) = rb2.dyv * dv4[0] * f0;
    AAM(2, 0) = rt1.dyv * dv1[L] * fL;
    AAM(2, 1) = rt1.phi + rt1.dyv * dv2[L] * fL;
    AAM(2, 2) = rt1.dyv * dv3[L] * fL;
    AAM(2, 3) = rt1.vvv + rt1.dyv * dv4[L] * fL;
    AAM(3, 0) = rt2.dyv * dv1[L] * fL;
    AAM(3, 1) = rt2.phi + rt2.dyv * dv2[L] * fL;
    AAM(3, 2) = rt2.dyv * dv3[L] * fL;
    AAM(3, 3) = rt2.vvv + rt2.dyv * dv4[L] * fL;
#undef AAM
    co[0] = rb1.bc - rb1.dyv * dvpr[0] * f0;
    co[1] = rb2.bc - rb2.dyv * dvpr[0] * f0;
    co[2] = rt1.bc - rt1.dyv * dvpr[L] * fL;
    co[3] = rt2.bc - rt2.dyv * dvpr[L] * fL;
    co[4] = rb1.bc - rb1.dyv * dvpi[0] * f0;
    co[5] = rb2.bc - rb2.dyv * dvpi[0] * f0;
    co[6] = rt1.bc - rt1.dyv * dvpi[L] * fL;
    co[7] = rt2.bc - rt2.dyv * dvpi[L] * fL;
    orc_gaussj(AA, 4, co, 2);
    for (int j = 0; j < my; j++) {
      phiwk[2 * j] = phipr[j] + co[0] * phi1[j] + co[1] * phi2[j];
      phiwk[2 * j + 1] = phipi[j] + co[4] * phi1[j] + co[5] * phi2[j];
    }
    for (int j = 0; j < 2 * my; j++)
      phi[j] = dalbe != 0.0 ? dalbe * phiwk[j] + dtri * phi[j] + dtxi * dvdy[j] : phiwk[j];
    for (int j = 0; j < my; j++) {
      if (rk2 != 0.0) {
        dvdy[2 * j] = (dvpr[j] + co[0] * dv1[j] + co[1] * dv2[j] + co[2] * dv3[j] +
                       co[3] * dv4[j]) * g->fmap[j];
        dvdy[2 * j + 1] = (dvpi[j] + co[4] * dv1[j] + co[5] * dv2[j] + co[6] * dv3[j] +
                           co[7] * dv4[j]) * g->fmap[j];
        v[2 * j] = vpr[j] + co[0] * v1[j] + co[1] * v2[j] + co[2] * v3[j] + co[3] * v4[j];
        v[2 * j + 1] = vpi[j] + co[4] * v1[j] + co[5] * v2[j] + co[6] * v3[j] + co[7] * v4[j];
      } else {
        dvdy[2 * j] = dvdy[2 * j + 1] = v[2 * j] = v[2 * j + 1] = 0.0;
      }
    }
  }
  free(a2);
  free(buf);
  free(a1);
}

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time with an explicit DFT for   */
/* each prime factor.  sign = -1 forward (e^{-i}), +1 backward.  Unscaled.  */
/* ------------------------------------------------------------------------ */
/* twiddle tables exp(-2 pi i t / N), cached per N */
typedef struct { int n; cplx *w; } twtab;
static twtab tw_cache[16];
static int tw_count = 0;
static const cplx *twiddles(int n) {
  const cplx *res = NULL;
#pragma omp critical(orc_twiddle)
  {
    for (int c = 0; c < tw_count; c++)
      if (tw_cache[c].n == n) res = tw_cache[c].w;
    if (!res) {
      cplx *w = malloc(n * sizeof(cplx));
      for (int t = 0; t < n; t++) w[t] = cexp(-2.0 * M_PI * I * (double)t / (double)n);
      if (tw_count < 16) {
        tw_cache[tw_count].n = n;
        tw_cache[tw_count].w = w;
        tw_count++;
      }
      res = w;
    }
  }
  return res;
}

/* length-n sub-transform of the sequence in[0], in[istride], ...;
 * w is the table of the top-level length N = n * s */
static void fft_rec(const cplx *in, int istride, cplx *out, int n, int sign,
                    cplx *scratch, const cplx *w, int N, int s) {
  if (n == 1) { out[0] = in[0]; return; }
  int p = 2;
  while (p * p <= n && n % p) p++;
  if (n % p) p = n;
  int m = n / p;
  for (int r = 0; r < p; r++)
    fft_rec(in + (size_t)r * istride, istride * p, scratch + (size_t)r * m, m, sign,
            out, w, N, s * p);
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      int kk = k + q * m;
      cplx acc = 0.0;
      for (int r = 0; r < p; r++) {
        long t = ((long)r * kk % n) * s % N;
        cplx tw = sign < 0 ? w[t] : conj(w[t]);
        acc += scratch[(size_t)r * m + k] * tw;
      }
      out[kk] = acc;
    }
  }
}
/* in-place complex transform of length n, contiguous */
static void fft_inplace(cplx *a, int n, int sign) {
  const cplx *w = twiddles(n);
  cplx *tmp = malloc(2 * (size_t)n * sizeof(cplx));
  memcpy(tmp, a, n * sizeof(cplx));
  fft_rec(tmp, 1, a, n, sign, tmp + n, w, n, 1);
  free(tmp);
}

/* cft, cftsingle.f:2-66: forward scaled by 1/n with e^{-i}, inverse e^{+i} */
void orc_cft(double *c, int n, int m, int iopt) {
  for (int v = 0; v < m; v++) {
    cplx *a = (cplx *)(c + (size_t)v * 2 * n);
    if (iopt < 0) {
      fft_inplace(a, n, -1);
      for (int k = 0; k < n; k++) a[k] /= (double)n;
    } else {
      fft_inplace(a, n, +1);
    }
  }
}

/* rft, rftsingle.f:25-96 packing: c(1)=mean, c(2)=0 (imag mode 0),
 * c(2k+1),c(2k+2) = Re,Im mode k, c(n+1) = Nyquist, c(n+2) = 0.  Each
 * vector has n+2 doubles. */
void orc_rft(double *c, int n, int m, int iopt) {
  cplx *a = malloc(n * sizeof(cplx));
  for (int v = 0; v < m; v++) {
    double *x = c + (size_t)v * (n + 2);
    if (iopt < 0) {
      for (int j = 0; j < n; j++) a[j] = x[j];
      fft_inplace(a, n, -1);
      for (int k = 0; k <= n / 2; k++) {
        x[2 * k] = creal(a[k]) / n;
        x[2 * k + 1] = cimag(a[k]) / n;
      }
      x[1] = 0.0;
      x[n + 1] = 0.0;
    } else {
      /* imag parts of mode 0 and of the Nyquist mode are ignored */
      a[0] = x[0];
      for (int k = 1; k < n / 2; k++) {
        a[k] = x[2 * k] + I * x[2 * k + 1];
        a[n - k] = conj(a[k]);
      }
      a[n / 2] = x[n];
      fft_inplace(a, n, +1);
      for (int j = 0; j < n; j++) x[j] = creal(a[j]);
      x[n] = x[n + 1] = 0.0;
    }
  }
  free(a);
}

/* ------------------------------------------------------------------------ */
/* state and time advance                                                    */
/* ------------------------------------------------------------------------ */
struct orc_state {
  const orc_grid *g;
  int flags, ncfl, istep, icfl, nS;
  double dt, time, dtr, cfl;
  double rhobot, rhotop, zbot, ztop;
  cplx *vor, *phi, *psi, *scal, *ten12, *ten13, *zf;
  double *rhst, *rhsz; /* physical [j][z][x] (= reference line order) */
  double *u00wk, *v00wk, *w00wk, *rf0u, *rf0w;
  int have_bounds;
  int cfl_ranks; /* compute ranks of the reference's CFL reduction (1 = one rank) */
  double trapz, mass[3]; /* .cf history: cross.f:1013-1019, 1034-1036 */
  struct orc_stats *stats; /* statistics window, NULL until configured */
};

static size_t S_of(const orc_grid *g) { return (size_t)g->nplanes * g->mz * g->my; }
#define SIDX(g, i, k, j) (((size_t)(i) * (g)->mz + (k)) * (g)->my + (j))

/* ------------------------------------------------------------------------ */
/* statistics window: addstats1 (cross.f:1981-2063), the phys2 sums          */
/* (cross.f:1586-1607), addstats2 (cross.f:2070-2287), writestats            */
/* (escru.f:276-484) and resetstats (escru.f:485-560)                        */
/* ------------------------------------------------------------------------ */
/* Accumulated columns, in the order writestats puts them in a .sta record:
 *  0 um  1 vm  2 wm  3 rum  4 rvm  5 rwm  6 Tm  7 w1m  8 w2m  9 w3m 10 them
 * 11 up 12 vp 13 wp 14 uvr 15 uwr 16 vwr 17 w1p 18 w2p 19 w3p 20 Tp 21 ep
 * 22 ruu 23 ruv 24 ruw 25 rvv 26 rvw 27 rww 28 rhom 29 thep 30 theup
 * 31 mum 32 num  (22..32 are written times 1/(mgalx mgalz)). */
#define ORC_NSTA 33
typedef struct orc_stats orc_stats;
static void col_dy1(const orc_grid *g, const cplx *u, cplx *du);
struct orc_stats {
  int nsp, *jsp;        /* spectra rows (0-based), jsptot of main.f:698-703 */
  int on, nacum, nacumsp;
  double *acc;          /* [ORC_NSTA][my] */
  double *spe;          /* [nplanes][12][nsp][nz+1], the .spe record order */
  double ener[9];       /* uner of the last accumulated step (cross.f:802-816) */
};

static void orc_stats_free(orc_state *st) {
  if (!st->stats) return;
  free(st->stats->jsp);
  free(st->stats->acc);
  free(st->stats->spe);
  free(st->stats);
  st->stats = NULL;
}

/* resetstats: every accumulator and both counters to zero */
void orc_stats_reset(orc_state *st) {
  orc_stats *q = st->stats;
  if (!q) return;
  const orc_grid *g = st->g;
  memset(q->acc, 0, sizeof(double) * ORC_NSTA * g->my);
  memset(q->spe, 0, sizeof(double) * (size_t)g->nplanes * 12 * q->nsp * (g->nz + 1));
  memset(q->ener, 0, sizeof q->ener);
  q->nacum = q->nacumsp = 0;
}

/* Spectra rows (0-based, each < my) and a reset; returns 0, or -1 for a row
 * outside the grid. */
int orc_stats_config(orc_state *st, int nsp, const int *rows) {
  const orc_grid *g = st->g;
  for (int n = 0; n < nsp; n++)
    if (rows[n] < 0 || rows[n] >= g->my) return -1;
  orc_stats_free(st);
  orc_stats *q = calloc(1, sizeof *q);
  q->nsp = nsp;
  q->jsp = malloc(sizeof(int) * (nsp > 0 ? nsp : 1));
  memcpy(q->jsp, rows, sizeof(int) * nsp);
  q->acc = calloc((size_t)ORC_NSTA * g->my, sizeof(double));
  q->spe = calloc((size_t)g->nplanes * 12 * (nsp > 0 ? nsp : 1) * (g->nz + 1), sizeof(double));
  st->stats = q;
  return 0;
}

/* istati: accumulate at rkstep 1 of the following steps while on */
void orc_stats_accumulate(orc_state *st, int on) {
  if (st->stats) st->stats->on = on != 0;
}

static int stats_active(const orc_state *st, int rkstep) {
  return rkstep == 1 && st->stats && st->stats->on;
}
static int kz_abs(const orc_grid *g, int k) { return k <= g->nz ? k : g->mz - k; }
static double cabs2(cplx a) { return creal(a) * creal(a) + cimag(a) * cimag(a); }
static double *spe_at(const orc_state *st, int i, int q, int iy, int ka) {
  const orc_stats *s = st->stats;
  return s->spe + (((size_t)i * 12 + q) * s->nsp + iy) * (st->g->nz + 1) + ka;
}

/* addstats1 on the transform inputs: rho u_i (mxwk, mywk, mzwk) and T
 * (drho), after the 00 modes are set (cross.f:578-586). */
static void addstats1(orc_state *st, const cplx *ru, const cplx *rv, const cplx *rw,
                      const cplx *T) {
  const orc_grid *g = st->g;
  orc_stats *q = st->stats;
  const int my = g->my;
  double *acc = q->acc;
  for (int i = 0; i < g->nplanes; i++)
    for (int k = 0; k < g->mz; k++) {
      const size_t o = SIDX(g, i, k, 0);
      const double f = i == 0 ? 1.0 : 2.0;
      const int ka = kz_abs(g, k);
      for (int n = 0; n < q->nsp; n++) {
        const int j = q->jsp[n];
        *spe_at(st, i, 8, n, ka) += f * cabs2(ru[o + j]);
        *spe_at(st, i, 9, n, ka) += f * cabs2(rv[o + j]);
        *spe_at(st, i, 10, n, ka) += f * cabs2(rw[o + j]);
        *spe_at(st, i, 11, n, ka) += f * cabs2(T[o + j]);
      }
      for (int j = 0; j < my; j++) acc[20 * my + j] += f * cabs2(T[o + j]);
      if (i == 0 && k == 0)
        for (int j = 0; j < my; j++) {
          acc[3 * my + j] += creal(ru[o + j]);
          acc[4 * my + j] += creal(rv[o + j]);
          acc[5 * my + j] += creal(rw[o + j]);
          acc[6 * my + j] += creal(T[o + j]);
        }
    }
}

/* addstats2 on the velocities u, v, w that foup2 returns, with the
 * vorticities of cross.f:748-767 (o1 = dw/dy - ikz v, o2 = ikz u - ikx w,
 * o3 = ikx v - du/dy) and dv/dy (cross.f:794-817, 2070-2287). */
static void addstats2(orc_state *st, const cplx *u, const cplx *v, const cplx *w) {
  const orc_grid *g = st->g;
  orc_stats *q = st->stats;
  const int my = g->my;
  double *acc = q->acc;
  cplx *du = malloc(my * sizeof(cplx)), *dv = malloc(my * sizeof(cplx)),
       *dw = malloc(my * sizeof(cplx));
  double ener[9] = {0};
  for (int i = 0; i < g->nplanes; i++)
    for (int k = 0; k < g->mz; k++) {
      const size_t o = SIDX(g, i, k, 0);
      const double f = i == 0 ? 1.0 : 2.0;
      const int ka = kz_abs(g, k);
      const cplx ikx = I * g->kx[i], ikz = I * g->kz[k];
      const double rK = g->alp2[i] + g->bet2[k];
      col_dy1(g, u + o, du);
      col_dy1(g, v + o, dv);
      col_dy1(g, w + o, dw);
      for (int j = 0; j < my; j++) {
        const cplx uu = u[o + j], vv = v[o + j], ww = w[o + j];
        const cplx o1 = dw[j] - ikz * vv, o2 = ikz * uu - ikx * ww, o3 = ikx * vv - du[j];
        const cplx cc = ikx * uu + dv[j] + ikz * ww;
        double e[9];
        e[0] = cabs2(uu); e[1] = cabs2(vv); e[2] = cabs2(ww);
        e[3] = creal(uu * conj(vv)); e[4] = creal(uu * conj(ww)); e[5] = creal(ww * conj(vv));
        e[6] = cabs2(o1); e[7] = cabs2(o2); e[8] = cabs2(o3);
        for (int n = 0; n < 9; n++) {
          acc[(11 + n) * my + j] += f * e[n];
          ener[n] += f * e[n] * g->hy[j];
        }
        acc[29 * my + j] += f * cabs2(cc);
        acc[30 * my + j] += f * creal(cc * conj(vv));
        /* pseudo-dissipation: k^2 |u_i|^2 + |dv/dy|^2 + |o1 + ikz v|^2 +
         * |o3 - ikx v|^2 */
        const double aa = rK * (e[0] + e[1] + e[2]) + cabs2(dv[j]) + cabs2(o1 + ikz * vv) +
                          cabs2(o3 - ikx * vv);
        acc[21 * my + j] += f * aa;
        if (i == 0 && k == 0) {
          acc[0 * my + j] += creal(uu);
          acc[1 * my + j] += creal(vv);
          acc[2 * my + j] += creal(ww);
          acc[7 * my + j] += creal(o1);
          acc[8 * my + j] += creal(o2);
          acc[9 * my + j] += creal(o3);
          acc[10 * my + j] += creal(dv[j]);
        }
      }
      for (int n = 0; n < q->nsp; n++) {
        const int j = q->jsp[n];
        const cplx uu = u[o + j], vv = v[o + j], ww = w[o + j];
        const cplx uv = uu * conj(vv);
        const double sp[8] = {cabs2(uu), cabs2(vv), cabs2(ww), creal(uv), cimag(uv),
                              cabs2(dw[j] - ikz * vv), cabs2(ikz * uu - ikx * ww),
                              cabs2(ikx * vv - du[j])};
        for (int c = 0; c < 8; c++) *spe_at(st, i, c, n, ka) += f * sp[c];
      }
    }
  memcpy(q->ener, ener, sizeof ener);
  q->nacum++;
  q->nacumsp++;
  free(du); free(dv); free(dw);
}

/* writestats' outputs: the .sta columns (22..32 scaled by 1/(mgalx mgalz),
 * [my][33] as one record row per y), the spectra ([nplanes][12][nsp][nz+1])
 * and info = {nacum, nacumsp, masst, enerdis, energy(9)}. */
void orc_stats_get(const orc_state *st, double *sta, double *spe, double *info) {
  const orc_stats *q = st->stats;
  const orc_grid *g = st->g;
  const int my = g->my;
  if (!q) return;
  const double fac = 1.0 / ((double)g->mgalx * g->mgalz);
  if (sta)
    for (int j = 0; j < my; j++)
      for (int c = 0; c < ORC_NSTA; c++)
        sta[(size_t)j * ORC_NSTA + c] = q->acc[(size_t)c * my + j] * (c >= 22 ? fac : 1.0);
  if (spe)
    memcpy(spe, q->spe, sizeof(double) * (size_t)g->nplanes * 12 * q->nsp * (g->nz + 1));
  if (info) {
    info[0] = q->nacum;
    info[1] = q->nacumsp;
    /* masst (escru.f:372-376) runs j = 0..my-1 over wkstats(j,18) with
     * hy(j+1); wkstats is declared (my,22), so its j = 0 entry is the
     * element before wkstats(1,18) in memory, wkstats(my,17) (rww of the
     * top row).  The reference's sum is restated as it runs. */
    double masst = q->acc[27 * my + (my - 1)] * fac * g->hy[0];
    for (int j = 1; j < my; j++) masst += fac * q->acc[28 * my + (j - 1)] * g->hy[j];
    info[2] = masst;
    /* enerdis: sum_j ep(j) / nacum * hy(j)  (cross.f:1152-1158) */
    double enerdis = 0.0;
    for (int j = 0; j < my; j++)
      enerdis += q->nacum ? q->acc[21 * my + j] / q->nacum * g->hy[j] : 0.0;
    info[3] = enerdis;
    /* energy(i) = sqrt(|uner(i)|) (cross.f:811-816) */
    for (int n = 0; n < 9; n++) info[4 + n] = sqrt(fabs(q->ener[n]));
  }
}

orc_state *orc_state_create(const orc_grid *g, double dt0, double cfl,
                            int ncfl, int flags) {
  orc_state *st = calloc(1, sizeof *st);
  st->g = g;
  st->flags = flags;
  st->ncfl = ncfl > 0 ? ncfl : 1;
  st->dt = (flags & ORC_DT_FP32) ? (double)(float)dt0 : dt0;
  st->cfl = cfl;
  st->cfl_ranks = 1;
  size_t S = S_of(g), P = (size_t)g->mgalx * g->my * g->mgalz;
  st->nS = (int)S;
  st->vor = calloc(S, sizeof(cplx));
  st->phi = calloc(S, sizeof(cplx));
  st->psi = calloc(S, sizeof(cplx));
  st->scal = calloc(S, sizeof(cplx));
  st->ten12 = calloc(S, sizeof(cplx));
  st->ten13 = calloc(S, sizeof(cplx));
  st->rhst = calloc(P, sizeof(double));
  if (flags & ORC_SCALAR_Z) {
    st->zf = calloc(S, sizeof(cplx));
    st->rhsz = calloc(P, sizeof(double));
  }
  int my = g->my;
  st->u00wk = calloc(my, sizeof(double));
  st->v00wk = calloc(my, sizeof(double));
  st->w00wk = calloc(my, sizeof(double));
  st->rf0u = calloc(my, sizeof(double));
  st->rf0w = calloc(my, sizeof(double));
  return st;
}

void orc_state_destroy(orc_state *st) {
  if (!st) return;
  free(st->vor); free(st->phi); free(st->psi); free(st->scal);
  free(st->ten12); free(st->ten13); free(st->rhst);
  free(st->zf); free(st->rhsz);
  free(st->u00wk); free(st->v00wk); free(st->w00wk); free(st->rf0u); free(st->rf0w);
  orc_stats_free(st);
  free(st);
}

void orc_state_set(orc_state *st, const double *vor, const double *phi,
                   const double *psi, const double *scal, const double *u00,
                   const double *v00, const double *w00) {
  size_t S = S_of(st->g);
  memcpy(st->vor, vor, S * sizeof(cplx));
  memcpy(st->phi, phi, S * sizeof(cplx));
  memcpy(st->psi, psi, S * sizeof(cplx));
  memcpy(st->scal, scal, S * sizeof(cplx));
  /* first-step special, cross.f:264-291: u00wk = u00 etc. */
  memcpy(st->u00wk, u00, st->g->my * sizeof(double));
  memcpy(st->v00wk, v00, st->g->my * sizeof(double));
  memcpy(st->w00wk, w00, st->g->my * sizeof(double));
  /* densities at the boundaries, cross.f:208-213 */
  st->rhobot = 1.0 / creal(st->scal[SIDX(st->g, 0, 0, 0)]);
  st->rhotop = 1.0 / creal(st->scal[SIDX(st->g, 0, 0, st->g->my - 1)]);
  st->istep = 0;
  st->icfl = 1;
  st->have_bounds = 1;
}

void orc_state_set_z(orc_state *st, const double *z) {
  if (!st->zf) return;
  memcpy(st->zf, z, S_of(st->g) * sizeof(cplx));
  /* lomaHZB cross.f:226-227 */
  st->zbot = creal(st->zf[SIDX(st->g, 0, 0, 0)]);
  st->ztop = creal(st->zf[SIDX(st->g, 0, 0, st->g->my - 1)]);
}
void orc_state_get_z(const orc_state *st, double *z) {
  if (st->zf) memcpy(z, st->zf, S_of(st->g) * sizeof(cplx));
}

void orc_state_get(const orc_state *st, double *vor, double *phi, double *psi,
                   double *scal, double *u00, double *v00, double *w00) {
  size_t S = S_of(st->g);
  if (vor) memcpy(vor, st->vor, S * sizeof(cplx));
  if (phi) memcpy(phi, st->phi, S * sizeof(cplx));
  if (psi) memcpy(psi, st->psi, S * sizeof(cplx));
  if (scal) memcpy(scal, st->scal, S * sizeof(cplx));
  if (u00) memcpy(u00, st->u00wk, st->g->my * sizeof(double));
  if (v00) memcpy(v00, st->v00wk, st->g->my * sizeof(double));
  if (w00) memcpy(w00, st->w00wk, st->g->my * sizeof(double));
}

void orc_state_history(const orc_state *st, double *ten12, double *ten13,
                       double *rhst, double *rf0u, double *rf0w) {
  size_t S = S_of(st->g), P = (size_t)st->g->mgalx * st->g->my * st->g->mgalz;
  if (ten12) memcpy(ten12, st->ten12, S * sizeof(cplx));
  if (ten13) memcpy(ten13, st->ten13, S * sizeof(cplx));
  if (rhst) memcpy(rhst, st->rhst, P * sizeof(double));
  if (rf0u) memcpy(rf0u, st->rf0u, st->g->my * sizeof(double));
  if (rf0w) memcpy(rf0w, st->rf0w, st->g->my * sizeof(double));
}

void orc_state_scalars(const orc_state *st, double *out) {
  out[0] = st->dt; out[1] = st->time; out[2] = st->dtr; out[3] = st->istep;
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* complex column helpers on the reference layout */
static void col_dy1(const orc_grid *g, const cplx *u, cplx *du) {
  orc_deryr2(g, (const double *)u, (double *)du, 1);
}
static void col_dy2(const orc_grid *g, const cplx *u, cplx *du) {
  orc_deryyr2(g, (const double *)u, (double *)du, 1);
}

/*
 * foup2 + phys2 (cross.f:1282-1828), processed one y-plane at a time.
 * in[8]  = rhou, rhov, rhow, lapT, dT/dx, dT/dy, dT/dz, T   (spectral)
 * out[11]= u, v, w, rhouu, rhovv, rhoww, rhouv, rhouw, rhovw, drho, DT
 * zin[4] = dZ/dx, dZ/dy, dZ/dz, lapZ ; zout = DZ  (scalar option)
 * cfl[2] returns (convective, diffusive) inverse time-step maxima.
 */
static void transform_pipeline(orc_state *st, cplx *const *in, cplx *const *out,
                               cplx *const *zin, cplx *zout, int rkstep,
                               double dtxi, double dtgamma, double dtbeta,
                               int do_cfl, double *cflres) {
  const orc_grid *g = st->g;
  const int mgalx = g->mgalx, mgalz = g->mgalz, my = g->my;
  const int np = g->nplanes, mz = g->mz, nz = g->nz;
  const int nin = zin ? 12 : 8, nout = zout ? 12 : 11;
  double alpha[3], beta[3];
  rk_coefs(alpha, beta);
  const double ab = alpha[rkstep - 1] / beta[rkstep - 1];
  const double hxalp = g->alp * (mgalx - 1) / M_PI; /* cross.f:1500 */
  const double hzbet = g->bet * (mgalz - 1) / M_PI;
  double cflc = 0.0, cfld = 0.0;
  double cflx_all = 0.0, cfly_all = 0.0, cflz_all = 0.0;
  /* per-line maxima, kept only when the reduction is over several ranks */
  const int nlines = mgalz * my, ranks = st->cfl_ranks;
  double *lmax = (do_cfl && ranks > 1) ? calloc((size_t)4 * nlines, sizeof(double)) : NULL;

#pragma omp parallel
  {
    cplx *zrow = malloc(mgalz * sizeof(cplx));
    cplx *plane = malloc((size_t)nin * np * mgalz * sizeof(cplx)); /* [f][i][z] */
    double *ph = malloc((size_t)(nin + nout) * mgalz * mgalx * sizeof(double));
    cplx *xl = malloc(mgalx * sizeof(cplx));
    cplx *oplane = malloc((size_t)nout * np * mgalz * sizeof(cplx));
    double lcx = 0, lcy = 0, lcz = 0, lct = 0;
#pragma omp for schedule(dynamic)
    for (int j = 0; j < my; j++) {
      /* localyz2zy + fourz(+1): pad kz, inverse complex transform (fou3D.f:103-146) */
      for (int f = 0; f < nin; f++) {
        const cplx *src = f < 8 ? in[f] : zin[f - 8];
        for (int i = 0; i < np; i++) {
          for (int z = 0; z < mgalz; z++) zrow[z] = 0.0;
          for (int k = 0; k <= nz; k++) zrow[k] = src[SIDX(g, i, k, j)];
          for (int k = nz + 1; k < mz; k++) zrow[k + mgalz - mz] = src[SIDX(g, i, k, j)];
          fft_inplace(zrow, mgalz, +1);
          memcpy(plane + ((size_t)f * np + i) * mgalz, zrow, mgalz * sizeof(cplx));
        }
      }
      /* fourx(+1): inverse real transform of each z line (fou3D.f:47-55) */
      for (int f = 0; f < nin; f++) {
        for (int z = 0; z < mgalz; z++) {
          for (int x = 0; x < mgalx; x++) xl[x] = 0.0;
          xl[0] = creal(plane[((size_t)f * np + 0) * mgalz + z]);
          for (int i = 1; i < np; i++) {
            xl[i] = plane[((size_t)f * np + i) * mgalz + z];
            xl[mgalx - i] = conj(xl[i]);
          }
          fft_inplace(xl, mgalx, +1);
          double *dst = ph + ((size_t)f * mgalz + z) * mgalx;
          for (int x = 0; x < mgalx; x++) dst[x] = creal(xl[x]);
        }
      }
      /* phys2 pointwise algebra (cross.f:1613-1673) */
      for (int z = 0; z < mgalz; z++) {
        const double *ru = ph + ((size_t)0 * mgalz + z) * mgalx;
        const double *rv = ph + ((size_t)1 * mgalz + z) * mgalx;
        const double *rw = ph + ((size_t)2 * mgalz + z) * mgalx;
        const double *lapT = ph + ((size_t)3 * mgalz + z) * mgalx;
        const double *Tx = ph + ((size_t)4 * mgalz + z) * mgalx;
        const double *Ty = ph + ((size_t)5 * mgalz + z) * mgalx;
        const double *Tz = ph + ((size_t)6 * mgalz + z) * mgalx;
        const double *T = ph + ((size_t)7 * mgalz + z) * mgalx;
        double *o[12];
        for (int q = 0; q < nout; q++) o[q] = ph + ((size_t)(nin + q) * mgalz + z) * mgalx;
        double *rh = st->rhst + ((size_t)j * mgalz + z) * mgalx;
        double lx = 0, l0 = 0, lz = 0, Tmax = 0;
        if (stats_active(st, rkstep)) {
          /* statistics sums over the line (cross.f:1586-1607, sigma = 0:
           * mu/mu0 = T^0, nu/nu0 = T^1); line j is this iteration's own */
          double *acc = st->stats->acc;
          for (int x = 0; x < mgalx; x++) {
            double a = ru[x] * T[x];
            acc[22 * my + j] += a * ru[x];
            acc[23 * my + j] += a * rv[x];
            acc[24 * my + j] += a * rw[x];
            a = rv[x] * T[x];
            acc[25 * my + j] += a * rv[x];
            acc[26 * my + j] += a * rw[x];
            a = rw[x] * T[x];
            acc[27 * my + j] += a * rw[x];
          }
          for (int x = 0; x < mgalx; x++) {
            acc[28 * my + j] += 1.0 / T[x];
            acc[31 * my + j] += pow(T[x], 0.0);
            acc[32 * my + j] += pow(T[x], 1.0);
          }
        }
        for (int x = 0; x < mgalx; x++) {
          double u = ru[x] * T[x], v = rv[x] * T[x], w = rw[x] * T[x];
          double tnext = T[x] + dtxi * rh[x];
          double dT = dtxi * rh[x];
          double r = -u * Tx[x] - v * Ty[x] - w * Tz[x] + T[x] * lapT[x] / g->peclet;
          rh[x] = r;
          tnext = tnext + dtgamma * r;
          dT = dT + dtgamma * r;
          o[0][x] = u; o[1][x] = v; o[2][x] = w;
          o[3][x] = ru[x] * u; o[4][x] = rv[x] * v; o[5][x] = rw[x] * w;
          o[6][x] = ru[x] * v; o[7][x] = ru[x] * w; o[8][x] = rv[x] * w;
          o[9][x] = (dT / (T[x] * tnext)) / dtbeta - ab * r / (T[x] * T[x]);
          o[10][x] = dT;
          if (zout) {
            /* passive scalar, lomaHZB cross.f:2766-2770 with sigma=0 */
            const double *Zx = ph + ((size_t)8 * mgalz + z) * mgalx;
            const double *Zy = ph + ((size_t)9 * mgalz + z) * mgalx;
            const double *Zz = ph + ((size_t)10 * mgalz + z) * mgalx;
            const double *lapZ = ph + ((size_t)11 * mgalz + z) * mgalx;
            double *rz = st->rhsz + ((size_t)j * mgalz + z) * mgalx;
            double dZ = dtxi * rz[x];
            double rzn = -u * Zx[x] - v * Zy[x] - w * Zz[x] + T[x] * lapZ[x] / g->peclet;
            rz[x] = rzn;
            o[11][x] = dZ + dtgamma * rzn;
          }
          if (do_cfl) {
            lx = fmax(lx, fabs(u));
            l0 = fmax(l0, fabs(v));
            lz = fmax(lz, fabs(w));
            Tmax = fmax(Tmax, T[x]);
          }
        }
        if (do_cfl) {
          /* cross.f:1690-1708 */
          double DD = 1.0 / g->peclet * Tmax;
          lcx = fmax(lcx, lx);
          lcz = fmax(lcz, lz);
          lcy = fmax(lcy, l0 / g->hy[j]);
          double ldiff = 2 * DD * ((hxalp / 2) * (hxalp / 2) + (hzbet / 2) * (hzbet / 2) +
                                   (1 / g->hy[j]) * (1 / g->hy[j]));
          lct = fmax(lct, ldiff);
          if (lmax) { /* line index z + mgalz*j, as jj = (j-1)/mgalz+1 at cross.f:1583 */
            double *lm = lmax + (size_t)4 * (z + (size_t)mgalz * j);
            lm[0] = lx; lm[1] = l0 / g->hy[j]; lm[2] = lz; lm[3] = ldiff;
          }
        }
      }
      /* fourx(-1) with truncation to nplanes modes (rftsingle.f:64-73) */
      for (int q = 0; q < nout; q++) {
        for (int z = 0; z < mgalz; z++) {
          const double *srcx = ph + ((size_t)(nin + q) * mgalz + z) * mgalx;
          for (int x = 0; x < mgalx; x++) xl[x] = srcx[x];
          fft_inplace(xl, mgalx, -1);
          for (int i = 0; i < np; i++)
            oplane[((size_t)q * np + i) * mgalz + z] = xl[i] / (double)mgalx;
          oplane[((size_t)q * np + 0) * mgalz + z] = creal(xl[0]) / (double)mgalx;
        }
      }
      /* fourz(-1) + localzy2yz: forward complex transform, keep mz modes */
      for (int q = 0; q < nout; q++) {
        cplx *dst = q < 11 ? out[q] : zout;
        for (int i = 0; i < np; i++) {
          memcpy(zrow, oplane + ((size_t)q * np + i) * mgalz, mgalz * sizeof(cplx));
          fft_inplace(zrow, mgalz, -1);
          for (int k = 0; k <= nz; k++) dst[SIDX(g, i, k, j)] = zrow[k] / (double)mgalz;
          for (int k = nz + 1; k < mz; k++)
            dst[SIDX(g, i, k, j)] = zrow[k + mgalz - mz] / (double)mgalz;
        }
      }
    }
#pragma omp critical
    {
      cflx_all = fmax(cflx_all, lcx);
      cfly_all = fmax(cfly_all, lcy);
      cflz_all = fmax(cflz_all, lcz);
      cfld = fmax(cfld, lct);
    }
    free(zrow); free(plane); free(ph); free(xl); free(oplane);
  }
  if (do_cfl && lmax) {
    /* cross.f:1778-1799 over several ranks: each rank sums ITS OWN maxima,
     * cfl0 = cflx + cfly + cflz, and MPI_ALLREDUCE(MAX) takes the largest
     * sum; cflT is reduced on its own.  Lines are split as pointers_calc
     * does (change.f:188-201): the first nlines % ranks ranks get one more. */
    int n1 = nlines / ranks, n2 = nlines - ranks * n1, lb = 0;
    cflc = 0.0; cfld = 0.0;
    for (int r = 0; r < ranks; r++) {
      int ln = r < n2 ? n1 + 1 : n1;
      double cx = 0, cy = 0, cz = 0, ct = 0;
      for (int L = lb; L < lb + ln; L++) {
        const double *lm = lmax + (size_t)4 * L;
        cx = fmax(cx, lm[0]); cy = fmax(cy, lm[1]); cz = fmax(cz, lm[2]); ct = fmax(ct, lm[3]);
      }
      cflc = fmax(cflc, cx * hxalp * 0.5 + cy + cz * hzbet * 0.5);
      cfld = fmax(cfld, ct);
      lb += ln;
    }
    cflres[0] = cflc;
    cflres[1] = cfld;
  } else if (do_cfl) {
    /* cross.f:1769-1817 */
    cflc = cflx_all * hxalp * 0.5 + cfly_all + cflz_all * hzbet * 0.5;
    cflres[0] = cflc;
    cflres[1] = cfld;
  }
  free(lmax);
}

/* foup2 (cross.f:1282-1430) on its own: the 8 transform inputs (rho u,
 * rho v, rho w, lap T, T_x, T_y, T_z, T; S complex each, mode-major
 * [i][k][j]) to the 11 outputs (u, v, w, rho uu, rho vv, rho ww, rho uv,
 * rho uw, rho vw, drho, DT) with the state's rhst history, for substep
 * rkstep of step size dt; at rkstep 1 *dtr receives CFL/(conv + diff) of
 * one rank (cross.f:1769-1818). */
void orc_foup2(orc_state *st, const double *in, double *out, int rkstep, double dt,
               double *dtr) {
  const orc_grid *g = st->g;
  const size_t S = S_of(g);
  cplx *pin[8], *pout[11];
  for (int q = 0; q < 8; q++) pin[q] = (cplx *)in + q * S;
  for (int q = 0; q < 11; q++) pout[q] = (cplx *)out + q * S;
  double alpha[3], beta[3];
  rk_coefs(alpha, beta);
  const double dtxi = dt * RK_XI[rkstep - 1];
  const double dtgamma = dt * RK_GAMA[rkstep - 1];
  const double dtbeta = dt * beta[rkstep - 1];
  double cfl[2] = {0, 0};
  const int do_cfl = rkstep == 1;
  transform_pipeline(st, pin, pout, NULL, NULL, rkstep, dtxi, dtgamma, dtbeta, do_cfl, cfl);
  if (dtr && do_cfl) *dtr = st->cfl / (cfl[0] + cfl[1]);
}

void orc_state_set_rhst(orc_state *st, const double *rhst) {
  memcpy(st->rhst, rhst, (size_t)st->g->mgalx * st->g->my * st->g->mgalz * sizeof(double));
}

void orc_state_cf(const orc_state *st, double *out) {
  out[0] = st->trapz;
  out[1] = st->mass[0];
  out[2] = st->mass[1];
  out[3] = st->mass[2];
}

void orc_state_set_cfl_ranks(orc_state *st, int ranks) {
  st->cfl_ranks = ranks > 1 ? ranks : 1;
}

int orc_substep(orc_state *st, int rkstep) {
  const orc_grid *g = st->g;
  const int my = g->my, np = g->nplanes, mz = g->mz;
  const size_t S = S_of(g);
  const int zon = (st->flags & ORC_SCALAR_Z) != 0;
  double alpha[3], beta[3];
  rk_coefs(alpha, beta);
  const double dtxi = st->dt * RK_XI[rkstep - 1];
  const double dtgamma = st->dt * RK_GAMA[rkstep - 1];
  const double dtbeta = st->dt * beta[rkstep - 1];
  const double ire = 1.0 / g->Re;

  cplx *mxw = malloc(S * sizeof(cplx)), *myw = malloc(S * sizeof(cplx)),
       *mzw = malloc(S * sizeof(cplx)), *nxw = malloc(S * sizeof(cplx)),
       *nyw = malloc(S * sizeof(cplx)), *nzw = malloc(S * sizeof(cplx)),
       *t23 = malloc(S * sizeof(cplx)), *drho = malloc(S * sizeof(cplx)),
       *chwk2 = malloc(S * sizeof(cplx)), *chwk = malloc(S * sizeof(cplx)),
       *t12 = st->ten12, *t13 = st->ten13;
  cplx *zx = NULL, *zy = NULL, *zz = NULL, *zl = NULL, *dz = NULL;
  if (zon) {
    zx = malloc(S * sizeof(cplx)); zy = malloc(S * sizeof(cplx));
    zz = malloc(S * sizeof(cplx)); zl = malloc(S * sizeof(cplx));
    dz = malloc(S * sizeof(cplx));
  }

  /* BCs, cross.f:315-349 */
  for (int i = 0; i < np; i++)
    for (int k = 0; k < mz; k++) {
      st->vor[SIDX(g, i, k, 0)] = st->vor[SIDX(g, i, k, 1)];
      st->vor[SIDX(g, i, k, my - 1)] = st->vor[SIDX(g, i, k, my - 2)];
      st->phi[SIDX(g, i, k, 0)] = 0.0;
      st->phi[SIDX(g, i, k, my - 1)] = 0.0;
      st->scal[SIDX(g, i, k, 0)] = 0.0;
      st->scal[SIDX(g, i, k, my - 1)] = 0.0;
      if (zon) {
        st->zf[SIDX(g, i, k, 0)] = 0.0;
        st->zf[SIDX(g, i, k, my - 1)] = 0.0;
      }
    }
  st->scal[SIDX(g, 0, 0, 0)] = 1.0 / st->rhobot;
  st->scal[SIDX(g, 0, 0, my - 1)] = 1.0 / st->rhotop;
  if (zon) {
    st->zf[SIDX(g, 0, 0, 0)] = st->zbot;
    st->zf[SIDX(g, 0, 0, my - 1)] = st->ztop;
  }

  /* phase A (cross.f:352-554), per mode */
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int i = 0; i < np; i++)
    for (int k = 0; k < mz; k++) {
      size_t o = SIDX(g, i, k, 0);
      double rK = g->bet2[k] + g->alp2[i];
      cplx ikx = I * g->kx[i], ikz = I * g->kz[k];
      /* nxwk = vor, nywk = psi (cross.f:352-360) */
      memcpy(nxw + o, st->vor + o, my * sizeof(cplx));
      memcpy(nyw + o, st->psi + o, my * sizeof(cplx));
      /* (1) Lapvdvhom: my, dmy/dy */
      orc_lapvdvhom(g, (const double *)(st->phi + o), (double *)(myw + o),
                    (double *)(nzw + o), rK);
      /* (2) mx, mz with tempa = ikx/k^2, tempb = ikz/k^2 (cross.f:125-154) */
      cplx ta = 0.0, tb = 0.0;
      if (rK != 0.0) { ta = ikx / rK; tb = ikz / rK; }
      for (int j = 0; j < my; j++) {
        mxw[o + j] = nzw[o + j] * ta - nxw[o + j] * tb;
        mzw[o + j] = nzw[o + j] * tb + nxw[o + j] * ta;
        drho[o + j] = st->scal[o + j];
        st->vor[o + j] = st->vor[o + j] + t12[o + j] * dtxi;
        st->phi[o + j] = st->phi[o + j] + t13[o + j] * dtxi;
      }
      /* d(psi)/dy (cross.f:434-437) */
      col_dy1(g, nyw + o, t23 + o);
      /* (3) rho u_i (cross.f:463-475) */
      for (int j = 0; j < my; j++) {
        mxw[o + j] = mxw[o + j] + ikx * nyw[o + j];
        mzw[o + j] = mzw[o + j] + ikz * nyw[o + j];
        myw[o + j] = myw[o + j] + t23[o + j];
      }
      /* (4) T derivatives (cross.f:521-554) */
      col_dy1(g, drho + o, t13 + o);
      col_dy2(g, drho + o, nyw + o);
      for (int j = 0; j < my; j++) {
        nyw[o + j] = nyw[o + j] - rK * drho[o + j];
        t23[o + j] = ikz * drho[o + j];
        t12[o + j] = ikx * drho[o + j];
      }
      if (zon) {
        /* Z derivatives (lomaHZB cross.f:632-666) */
        const cplx *Z = st->zf + o;
        col_dy1(g, Z, zy + o);
        col_dy2(g, Z, zl + o);
        for (int j = 0; j < my; j++) {
          zl[o + j] = zl[o + j] - rK * Z[j];
          zx[o + j] = ikx * Z[j];
          zz[o + j] = ikz * Z[j];
        }
      }
    }
  /* 00 modes (cross.f:486-499) */
  {
    size_t o = SIDX(g, 0, 0, 0);
    for (int j = 0; j < my; j++) {
      mxw[o + j] = st->u00wk[j];
      myw[o + j] = st->v00wk[j];
      mzw[o + j] = st->w00wk[j];
    }
    for (int j = 1; j < my - 1; j++) {
      st->u00wk[j] = st->u00wk[j] + dtxi * st->rf0u[j];
      st->w00wk[j] = st->w00wk[j] + dtxi * st->rf0w[j];
    }
  }

  if (stats_active(st, rkstep)) addstats1(st, mxw, myw, mzw, drho);

  /* (5) foup2 */
  int do_cfl = (rkstep == 1 && st->icfl);
  double cflres[2] = {0, 0};
  {
    cplx *in[8] = {mxw, myw, mzw, nyw, t12, t13, t23, drho};
    cplx *out[11] = {mxw, myw, mzw, nxw, nyw, nzw, t12, t13, t23, drho, chwk2};
    cplx *zin[4] = {zx, zy, zz, zl};
    /* inputs are read per y plane before any output of that plane is
     * written, and outputs are written per plane: alias-safe because every
     * plane j is read and written by the same iteration */
    transform_pipeline(st, in, out, zon ? zin : NULL, zon ? dz : NULL, rkstep, dtxi,
                       dtgamma, dtbeta, do_cfl, cflres);
  }
  if (do_cfl) st->dtr = st->cfl / (cflres[0] + cflres[1]); /* cross.f:1810-1818 */
  if (stats_active(st, rkstep)) addstats2(st, mxw, myw, mzw);

  /* (6a) T += DT (cross.f:619-627); Z += DZ */
  for (size_t s = 0; s < S; s++) st->scal[s] += chwk2[s];
  if (zon)
    for (size_t s = 0; s < S; s++) st->zf[s] += dz[s];

  /* 00-mode RHS (cross.f:686-710) needs d(rhouv)/dy, d(rhovw)/dy of 00 */
  double *rf0u = st->rf0u, *rf0w = st->rf0w;
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int i = 0; i < np; i++)
    for (int k = 0; k < mz; k++) {
      size_t o = SIDX(g, i, k, 0);
      double rK = g->bet2[k] + g->alp2[i];
      cplx ikx = I * g->kx[i], ikz = I * g->kz[k];
      /* (6b) N_i (cross.f:651-682) */
      col_dy1(g, nyw + o, nyw + o);
      for (int j = 0; j < my; j++)
        nyw[o + j] = -ikx * t12[o + j] - nyw[o + j] - ikz * t23[o + j];
      col_dy1(g, t12 + o, t12 + o);
      col_dy1(g, t23 + o, t23 + o);
      for (int j = 0; j < my; j++) {
        cplx nx = -ikx * nxw[o + j] - t12[o + j] - ikz * t13[o + j];
        cplx nz = -ikx * t13[o + j] - t23[o + j] - ikz * nzw[o + j];
        nxw[o + j] = nx;
        nzw[o + j] = nz;
      }
      if (i == 0 && k == 0) {
        /* 00 RHS: -d(rhouv)/dy + u00''/Re (cross.f:686-710) */
        double *a = malloc(my * sizeof(double)), *b = malloc(my * sizeof(double));
        for (int j = 0; j < my; j++) { a[j] = creal(mxw[o + j]); b[j] = creal(mzw[o + j]); }
        orc_deryyr(g, a, a);
        orc_deryyr(g, b, b);
        for (int j = 0; j < my; j++) {
          rf0u[j] = -creal(t12[o + j]) + ire * a[j];
          rf0w[j] = -creal(t23[o + j]) + ire * b[j];
        }
        free(a); free(b);
      }
      /* (6c) vorticities (cross.f:748-767) */
      col_dy1(g, mxw + o, t12 + o);
      col_dy1(g, mzw + o, t23 + o);
      for (int j = 0; j < my; j++) {
        cplx vx = t23[o + j] - ikz * myw[o + j];
        cplx vy = ikz * mxw[o + j] - ikx * mzw[o + j];
        cplx vz = ikx * myw[o + j] - t12[o + j];
        t23[o + j] = vx; t13[o + j] = vy; t12[o + j] = vz;
      }
      /* (6d) Laplacians (cross.f:821-857) */
      col_dy2(g, t13 + o, chwk + o);
      for (int j = 0; j < my; j++) t13[o + j] = chwk[o + j] - rK * t13[o + j];
      for (int j = 0; j < my; j++) t23[o + j] = ikz * t23[o + j] - ikx * t12[o + j];
      col_dy2(g, t23 + o, chwk + o);
      for (int j = 0; j < my; j++) t23[o + j] = chwk[o + j] - rK * t23[o + j];
      /* (9) RHS (cross.f:885-921) */
      for (int j = 0; j < my; j++) {
        t12[o + j] = ikz * nxw[o + j] - ikx * nzw[o + j] + ire * t13[o + j];
        nxw[o + j] = ikx * nxw[o + j] + ikz * nzw[o + j];
      }
      col_dy1(g, nxw + o, nxw + o);
      for (int j = 0; j < my; j++) t13[o + j] = -rK * nyw[o + j] - nxw[o + j] - ire * t23[o + j];
      /* gamma update (cross.f:966-981) */
      for (int j = 0; j < my; j++) {
        st->vor[o + j] = st->vor[o + j] + dtgamma * t12[o + j];
        st->phi[o + j] = st->phi[o + j] + dtgamma * t13[o + j];
      }
    }
  /* 00 gamma update (cross.f:1006-1011) */
  for (int j = 1; j < my - 1; j++) {
    st->u00wk[j] = st->u00wk[j] + dtgamma * rf0u[j];
    st->w00wk[j] = st->w00wk[j] + dtgamma * rf0w[j];
  }
  /* momentum of the 00 modes for the .cf history (cross.f:1013-1019),
   * before rho v 00 is updated */
  st->mass[0] = st->mass[1] = st->mass[2] = 0.0;
  for (int j = 0; j < my; j++) {
    st->mass[0] += st->u00wk[j] * g->hy[j];
    st->mass[1] += st->v00wk[j] * g->hy[j];
    st->mass[2] += st->w00wk[j] * g->hy[j];
  }
  /* (8) 00 mode of rho v from the constraint (cross.f:1027-1072) */
  {
    double *d00 = malloc(my * sizeof(double)), *rf0v = malloc(my * sizeof(double));
    for (int j = 0; j < my; j++) d00[j] = creal(drho[SIDX(g, 0, 0, j)]);
    orc_inty8(g, d00, rf0v);
    double trapz = -rf0v[my - 1];
    st->trapz = trapz;
    for (int j = 0; j < my; j++)
      st->v00wk[j] = rf0v[j] + trapz / (1.0 + sqrt(st->rhotop / st->rhobot));
    free(d00); free(rf0v);
  }
  /* solvepsi per mode (cross.f:1092-1104) */
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int i = 0; i < np; i++)
    for (int k = 0; k < mz; k++) {
      size_t o = SIDX(g, i, k, 0);
      double rK = g->bet2[k] + g->alp2[i];
      orc_solvepsi(g, (const double *)(drho + o), (double *)(st->psi + o),
                   (double *)(t23 + o), rK);
    }

  int nan = 0;
  for (size_t s = 0; s < S && !nan; s++)
    if (isnan(creal(st->vor[s])) || isnan(cimag(st->vor[s]))) nan = 1;
  free(mxw); free(myw); free(mzw); free(nxw); free(nyw); free(nzw);
  free(t23); free(drho); free(chwk2); free(chwk);
  free(zx); free(zy); free(zz); free(zl); free(dz);
  return nan;
}

int orc_step(orc_state *st, int nsteps) {
  for (int n = 0; n < nsteps; n++) {
    st->istep++;
    /* cross.f:238 */
    if ((st->istep - 1) % st->ncfl == 0) st->icfl = 1;
    for (int rk = 1; rk <= 3; rk++)
      if (orc_substep(st, rk)) return 1;
    /* cross.f:1196-1204 */
    st->time += st->dt;
    if (st->icfl) {
      st->icfl = 0;
      st->dt = (st->flags & ORC_DT_FP32) ? (double)(float)st->dtr : st->dtr;
    }
  }
  return 0;
}
