/*
 * CPU oracle for the compact finite-difference core (TEST INFRASTRUCTURE ONLY).
 * Plain-C FP64 restatement of the reference algorithm; every function cites
 * the reference file:line it follows. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, as the checker.
 */
#include "compact_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;
void ora_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

/* ------------------------------------------------------------------------- */
/* Symmetry folding, stencils.f90:2390-2453 (lower/upper_symm_weights_int).   */
/* alb/arb rows are [row][col]; ncl/ncr are the row lengths.                  */
/* ------------------------------------------------------------------------- */
static void lower_symm(double alb[4][5], double arb[4][9], const double* ali,
                       const double* ari, int ncl, int ncr, int nol, int nor,
                       int syml, int symr) {
  for (int r = 0; r < 4; ++r) {
    for (int c = 0; c < ncl; ++c) alb[r][c] = ali[c];
    for (int c = 0; c < ncr; ++c) arb[r][c] = ari[c];
  }
  for (int j = 1; j <= nol; ++j) {
    int k = nol - j + 1;
    for (int i = 1; i <= k; ++i) {
      int i1 = k - i + 1, i2 = k + i;
      alb[j - 1][i2 - 1] += (double)syml * alb[j - 1][i1 - 1];
      alb[j - 1][i1 - 1] = 0.0;
    }
  }
  for (int j = 1; j <= nor; ++j) {
    int k = nor - j + 1;
    for (int i = 1; i <= k; ++i) {
      int i1 = k - i + 1, i2 = k + i;
      arb[j - 1][i2 - 1] += (double)symr * arb[j - 1][i1 - 1];
      arb[j - 1][i1 - 1] = 0.0;
    }
  }
}

static void upper_symm(double alb[4][5], double arb[4][9], const double* ali,
                       const double* ari, int ncl, int ncr, int nol, int nor,
                       int syml, int symr) {
  const int n = 4;
  for (int r = 0; r < 4; ++r) {
    for (int c = 0; c < ncl; ++c) alb[r][c] = ali[c];
    for (int c = 0; c < ncr; ++c) arb[r][c] = ari[c];
  }
  for (int j = 1; j <= nol; ++j) {
    int jj = n + 1 - j, k = nol - j + 1;
    for (int i = 1; i <= k; ++i) {
      int noff = ncl - k;
      alb[jj - 1][noff - i] += (double)syml * alb[jj - 1][noff + i - 1];
      alb[jj - 1][noff + i - 1] = 0.0;
    }
  }
  for (int j = n - nor + 1; j <= n; ++j) {
    int k = j - n + nor;
    for (int i = 1; i <= k; ++i) {
      int noff = ncr - k;
      arb[j - 1][noff - i] += (double)symr * arb[j - 1][noff + i - 1];
      arb[j - 1][noff + i - 1] = 0.0;
    }
  }
}

/* upper rows by reversal of the lower ones: alb2(:,j) = alb1(ncl:1:-1,5-j) */
static void reverse_rows(ora_weight* w, int var, double sign_r) {
  for (int j = 0; j < 4; ++j) {
    for (int c = 0; c < w->ncl; ++c)
      w->alb2[var][j][c] = w->alb1[var][3 - j][w->ncl - 1 - c];
    for (int c = 0; c < w->ncr; ++c)
      w->arb2[var][j][c] = sign_r * w->arb1[var][3 - j][w->ncr - 1 - c];
  }
}

static void set_row(double* dst, const double* src, int n) { memcpy(dst, src, sizeof(double) * n); }

/* c10d1, stencils.f90:207-277 */
static void w_c10d1(ora_weight* w) {
  static const double ali[5] = {0.45, 4.5, 9.0, 4.5, 0.45};
  static const double ari[7] = {-0.015, -1.515, -6.375, 0.0, 6.375, 1.515, 0.015};
  static const double alb[4][5] = {{0.0, 0.0, 4.725, 9.45, 0.0},
                                   {0.0, 1.94578125, 7.783125, 1.94578125, 0.0},
                                   {0.2964375, 4.743, 10.67175, 4.743, 0.2964375},
                                   {0.451390625, 4.63271875, 9.38146875, 4.63271875, 0.451390625}};
  static const double arb[4][7] = {{0.0, 0.0, 0.0, -11.8125, 9.45, 2.3625, 0.0},
                                   {0.0, 0.0, -5.83734375, 0.0, 5.83734375, 0.0, 0.0},
                                   {0.0, -1.23515625, -7.905, 0.0, 7.905, 1.23515625, 0.0},
                                   {-0.015, -1.53, -6.66984375, 0.0, 6.66984375, 1.53, 0.015}};
  w->nol = 2; w->nor = 3; w->ncl = 5; w->ncr = 7; w->nbc1 = 4; w->nbc2 = 4;
  w->null_option = 0; w->implicit_op = 1;
  set_row(w->ali, ali, 5); set_row(w->ari, ari, 7);
  for (int j = 0; j < 4; ++j) { set_row(w->alb1[1][j], alb[j], 5); set_row(w->arb1[1][j], arb[j], 7); }
  reverse_rows(w, 1, -1.0);
  lower_symm(w->alb1[2], w->arb1[2], ali, ari, 5, 7, 2, 3, -1, +1);
  lower_symm(w->alb1[0], w->arb1[0], ali, ari, 5, 7, 2, 3, +1, -1);
  upper_symm(w->alb2[2], w->arb2[2], ali, ari, 5, 7, 2, 3, -1, +1);
  upper_symm(w->alb2[0], w->arb2[0], ali, ari, 5, 7, 2, 3, +1, -1);
}

/* c10d2, stencils.f90:358-428 */
static void w_c10d2(ora_weight* w) {
  static const double ali[5] = {387.0, 6012.0, 16182.0, 6012.0, 387.0};
  static const double ari[7] = {79.0, 4671.0, 9585.0, -28670.0, 9585.0, 4671.0, 79.0};
  static const double alb[4][5] = {{0.0, 0.0, 1.0, 11.0, 0.0},
                                   {0.0, 1.0, 10.0, 1.0, 0.0},
                                   {23.0, 688.0, 2358.0, 688.0, 23.0},
                                   {387.0, 6012.0, 16182.0, 6012.0, 387.0}};
  static const double arb[4][7] = {{0.0, 0.0, 0.0, 13.0, -27.0, 15.0, -1.0},
                                   {0.0, 0.0, 12.0, -24.0, 12.0, 0.0, 0.0},
                                   {0.0, 465.0, 1920.0, -4770.0, 1920.0, 465.0, 0.0},
                                   {79.0, 4671.0, 9585.0, -28670.0, 9585.0, 4671.0, 79.0}};
  w->nol = 2; w->nor = 3; w->ncl = 5; w->ncr = 7; w->nbc1 = 4; w->nbc2 = 4;
  w->null_option = 0; w->implicit_op = 1;
  set_row(w->ali, ali, 5); set_row(w->ari, ari, 7);
  for (int j = 0; j < 4; ++j) { set_row(w->alb1[1][j], alb[j], 5); set_row(w->arb1[1][j], arb[j], 7); }
  reverse_rows(w, 1, 1.0);
  lower_symm(w->alb1[2], w->arb1[2], ali, ari, 5, 7, 2, 3, +1, +1);
  lower_symm(w->alb1[0], w->arb1[0], ali, ari, 5, 7, 2, 3, -1, -1);
  upper_symm(w->alb2[2], w->arb2[2], ali, ari, 5, 7, 2, 3, +1, +1);
  upper_symm(w->alb2[0], w->arb2[0], ali, ari, 5, 7, 2, 3, -1, -1);
}

/* c10d8, stencils.f90:515-621 */
static void w_c10d8(ora_weight* w) {
  const double zeta = 29.0, alpha = 14.0, beta = 1.5;
  const double aa = 4200.0, bb = -3360.0, cc = 1680.0, dd = -480.0, ee = 60.0;
  const double alpha2 = beta + alpha, zeta1 = alpha + zeta, alpha1 = beta + alpha;
  const double dd4 = ee + dd, cc3 = dd + cc, bb3 = ee + bb, bb2 = cc + bb, aa2 = dd + aa,
               b22 = ee + bb, aa1 = bb + aa, bb1 = cc + bb, cc1 = dd + cc, dd1 = ee + dd;
  const double ali[5] = {beta, alpha, zeta, alpha, beta};
  const double ari[9] = {ee, dd, cc, bb, aa, bb, cc, dd, ee};
  const double alb[4][5] = {{0.0, 0.0, zeta1, alpha1, beta},
                            {0.0, alpha2, zeta, alpha, beta},
                            {beta, alpha, zeta, alpha, beta},
                            {beta, alpha, zeta, alpha, beta}};
  const double arb[4][9] = {{0.0, 0.0, 0.0, 0.0, aa1, bb1, cc1, dd1, ee},
                            {0.0, 0.0, 0.0, bb2, aa2, b22, cc, dd, ee},
                            {0.0, 0.0, cc3, bb3, aa, bb, cc, dd, ee},
                            {0.0, dd4, cc, bb, aa, bb, cc, dd, ee}};
  w->nol = 2; w->nor = 4; w->ncl = 5; w->ncr = 9; w->nbc1 = 4; w->nbc2 = 4;
  w->null_option = 0; w->implicit_op = 1;
  set_row(w->ali, ali, 5); set_row(w->ari, ari, 9);
  for (int j = 0; j < 4; ++j) { set_row(w->alb1[1][j], alb[j], 5); set_row(w->arb1[1][j], arb[j], 9); }
  reverse_rows(w, 1, 1.0);
  lower_symm(w->alb1[2], w->arb1[2], ali, ari, 5, 9, 2, 4, +1, +1);
  lower_symm(w->alb1[0], w->arb1[0], ali, ari, 5, 9, 2, 4, -1, -1);
  upper_symm(w->alb2[2], w->arb2[2], ali, ari, 5, 9, 2, 4, +1, +1);
  upper_symm(w->alb2[0], w->arb2[0], ali, ari, 5, 9, 2, 4, -1, -1);
}

/* c8ff8 (sfspec=2, fbar), stencils.f90:713-835, difference form :832-834 */
static void w_c8ff8(ora_weight* w) {
  const double beta = 1.6688e-1, alpha = 6.6624e-1, zeta = 1.0;
  const double aa = 9.9965e-1, bb = 6.6652e-1, cc = 1.6674e-1, dd = 4.0e-5, ee = -5.0e-6;
  const double ali[5] = {beta, alpha, zeta, alpha, beta};
  const double ari[9] = {ee, dd, cc, bb, aa, bb, cc, dd, ee};
  const double alb[4][5] = {{0.0, 0.0, 1.0, 0.0, 0.0},
                            {0.0, 4.997e-1, 1.0, 4.997e-1, 0.0},
                            {1.6688e-1, 6.6624e-1, 1.0, 6.6624e-1, 1.6688e-1},
                            {1.6688e-1, 6.6624e-1, 1.0, 6.6624e-1, 1.6688e-1}};
  const double arb[4][9] = {
      {0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0},
      {0.0, 0.0, 0.0, 4.9985e-1, 9.997e-1, 4.9985e-1, 0.0, 0.0, 0.0},
      {0.0, 0.0, 1.668e-1, 6.6656e-1, 9.9952e-1, 6.6656e-1, 1.668e-1, 0.0, 0.0},
      {0.0, 4.0e-5, 1.6672e-1, 6.6652e-1, 9.9968e-1, 6.6652e-1, 1.6672e-1, 4.0e-5, 0.0}};
  w->nol = 2; w->nor = 4; w->ncl = 5; w->ncr = 9; w->nbc1 = 4; w->nbc2 = 4;
  w->null_option = 1; w->implicit_op = 1;
  set_row(w->ali, ali, 5); set_row(w->ari, ari, 9);
  for (int j = 0; j < 4; ++j) { set_row(w->alb1[1][j], alb[j], 5); set_row(w->arb1[1][j], arb[j], 9); }
  reverse_rows(w, 1, 1.0);
  lower_symm(w->alb1[2], w->arb1[2], ali, ari, 5, 9, 2, 4, +1, +1);
  lower_symm(w->alb1[0], w->arb1[0], ali, ari, 5, 9, 2, 4, -1, -1);
  upper_symm(w->alb2[2], w->arb2[2], ali, ari, 5, 9, 2, 4, +1, +1);
  upper_symm(w->alb2[0], w->arb2[0], ali, ari, 5, 9, 2, 4, -1, -1);
  /* difference: ari(3:7) -= ali(1:5), same for every boundary table */
  for (int c = 0; c < 5; ++c) w->ari[2 + c] -= w->ali[c];
  for (int v = 0; v < 3; ++v)
    for (int j = 0; j < 4; ++j)
      for (int c = 0; c < 5; ++c) {
        w->arb1[v][j][2 + c] -= w->alb1[v][j][c];
        w->arb2[v][j][2 + c] -= w->alb2[v][j][c];
      }
}

/* cgfs4 (gfspec=6, gbar), stencils.f90:1387-1476, difference form :1473-1475 */
static void w_cgfs4(ora_weight* w) {
  const double agau = 3565.0 / 10368.0, bgau = 3091.0 / 12960.0, cgau = 1997.0 / 25920.0,
               dgau = 149.0 / 12960.0, egau = 107.0 / 103680.0;
  const double dg14 = dgau + egau, cg13 = cgau + dgau, bg23 = bgau + egau, bg12 = bgau + cgau,
               ag22 = agau + dgau, bg32 = bgau + egau, ag11 = agau + bgau, bg21 = bgau + cgau,
               cg31 = cgau + dgau, dg41 = dgau + egau;
  const double ali[5] = {1.0, 0, 0, 0, 0};
  const double ari[9] = {egau, dgau, cgau, bgau, agau, bgau, cgau, dgau, egau};
  const double arb[4][9] = {{0.0, 0.0, 0.0, 0.0, ag11, bg21, cg31, dg41, egau},
                            {0.0, 0.0, 0.0, bg12, ag22, bg32, cgau, dgau, egau},
                            {0.0, 0.0, cg13, bg23, agau, bgau, cgau, dgau, egau},
                            {0.0, dg14, cgau, bgau, agau, bgau, cgau, dgau, egau}};
  w->nol = 0; w->nor = 4; w->ncl = 1; w->ncr = 9; w->nbc1 = 4; w->nbc2 = 4;
  w->null_option = 1; w->implicit_op = 0;
  set_row(w->ali, ali, 5); set_row(w->ari, ari, 9);
  for (int j = 0; j < 4; ++j) { w->alb1[1][j][0] = 1.0; set_row(w->arb1[1][j], arb[j], 9); }
  reverse_rows(w, 1, 1.0);
  lower_symm(w->alb1[2], w->arb1[2], ali, ari, 1, 9, 0, 4, +1, +1);
  lower_symm(w->alb1[0], w->arb1[0], ali, ari, 1, 9, 0, 4, -1, -1);
  upper_symm(w->alb2[2], w->arb2[2], ali, ari, 1, 9, 0, 4, +1, +1);
  upper_symm(w->alb2[0], w->arb2[0], ali, ari, 1, 9, 0, 4, -1, -1);
  w->ari[4] -= w->ali[0];
  for (int v = 0; v < 3; ++v)
    for (int j = 0; j < 4; ++j) {
      w->arb1[v][j][4] -= w->alb1[v][j][0];
      w->arb2[v][j][4] -= w->alb2[v][j][0];
    }
}

void ora_get_weight(int op, ora_weight* w) {
  memset(w, 0, sizeof(*w));
  switch (op) {
    case ORA_D1: w_c10d1(w); break;
    case ORA_D2: w_c10d2(w); break;
    case ORA_D8: w_c10d8(w); break;
    case ORA_SF: w_c8ff8(w); break;
    case ORA_GF: w_cgfs4(w); break;
    default: break;
  }
}

/* ------------------------------------------------------------------------- */
/* Banded LU, pentadiagonal.f90. Column-major, 1-based accessor.              */
/* ------------------------------------------------------------------------- */
#define CC(i, j) c[((j)-1) * (size_t)n + ((i)-1)]

/* bpentLUD1, pentadiagonal.f90:18-39 */
void ora_bpentLUD1(double* c, int n) {
  for (int i = 1; i <= n - 2; ++i) {
    CC(i + 1, 2) = CC(i + 1, 2) / CC(i, 3);
    CC(i + 1, 3) = CC(i + 1, 3) - CC(i, 4) * CC(i + 1, 2);
    CC(i + 1, 4) = CC(i + 1, 4) - CC(i, 5) * CC(i + 1, 2);
    CC(i + 2, 1) = CC(i + 2, 1) / CC(i, 3);
    CC(i + 2, 2) = CC(i + 2, 2) - CC(i, 4) * CC(i + 2, 1);
    CC(i + 2, 3) = CC(i + 2, 3) - CC(i, 5) * CC(i + 2, 1);
  }
  CC(n, 2) = CC(n, 2) / CC(n - 1, 3);
  CC(n, 3) = CC(n, 3) - CC(n - 1, 4) * CC(n, 2);
  for (int i = 1; i <= n; ++i) CC(i, 3) = 1.0 / CC(i, 3);
}

/* bpentLUS1, pentadiagonal.f90:42-61 */
void ora_bpentLUS1(const double* c, double* r, int n) {
  for (int i = 1; i <= n - 2; ++i) {
    r[i] = r[i] - r[i - 1] * CC(i + 1, 2);
    r[i + 1] = r[i + 1] - r[i - 1] * CC(i + 2, 1);
  }
  r[n - 1] = r[n - 1] - r[n - 2] * CC(n, 2);
  r[n - 1] = r[n - 1] * CC(n, 3);
  r[n - 2] = (r[n - 2] - CC(n - 1, 4) * r[n - 1]) * CC(n - 1, 3);
  for (int i = n - 2; i >= 1; --i)
    r[i - 1] = (r[i - 1] - CC(i, 4) * r[i] - CC(i, 5) * r[i + 1]) * CC(i, 3);
}

/* ppentLUD1 (bordered cyclic LU), pentadiagonal.f90:64-131 */
void ora_ppentLUD1(double* c, int N) {
  const int n = N;
  CC(1, 8) = CC(1, 1);
  CC(1, 9) = CC(1, 2);
  CC(1, 3) = 1.0 / CC(1, 3);
  CC(1, 6) = CC(N - 1, 5) * CC(1, 3);
  CC(1, 7) = CC(N, 4) * CC(1, 3);

  CC(2, 2) = CC(2, 2) * CC(1, 3);
  CC(2, 3) = CC(2, 3) - CC(2, 2) * CC(1, 4);
  CC(2, 4) = CC(2, 4) - CC(2, 2) * CC(1, 5);
  CC(2, 8) = -CC(2, 2) * CC(1, 8);
  CC(2, 9) = CC(2, 1) - CC(2, 2) * CC(1, 9);
  CC(2, 3) = 1.0 / CC(2, 3);
  CC(2, 6) = -CC(1, 6) * CC(1, 4) * CC(2, 3);
  CC(2, 7) = (CC(N, 5) - CC(1, 7) * CC(1, 4)) * CC(2, 3);

  for (int K = 3; K <= N - 4; ++K) {
    CC(K, 1) = CC(K, 1) * CC(K - 2, 3);
    CC(K, 2) = (CC(K, 2) - CC(K, 1) * CC(K - 2, 4)) * CC(K - 1, 3);
    CC(K, 3) = CC(K, 3) - (CC(K, 2) * CC(K - 1, 4) + CC(K, 1) * CC(K - 2, 5));
    CC(K, 4) = CC(K, 4) - CC(K, 2) * CC(K - 1, 5);
    CC(K, 8) = -(CC(K, 2) * CC(K - 1, 8) + CC(K, 1) * CC(K - 2, 8));
    CC(K, 9) = -(CC(K, 2) * CC(K - 1, 9) + CC(K, 1) * CC(K - 2, 9));
    CC(K, 3) = 1.0 / CC(K, 3);
    CC(K, 6) = -(CC(K - 1, 6) * CC(K - 1, 4) + CC(K - 2, 6) * CC(K - 2, 5)) * CC(K, 3);
    CC(K, 7) = -(CC(K - 1, 7) * CC(K - 1, 4) + CC(K - 2, 7) * CC(K - 2, 5)) * CC(K, 3);
  }

  CC(N - 3, 1) = CC(N - 3, 1) * CC(N - 5, 3);
  CC(N - 3, 2) = (CC(N - 3, 2) - CC(N - 3, 1) * CC(N - 5, 4)) * CC(N - 4, 3);
  CC(N - 3, 3) = CC(N - 3, 3) - (CC(N - 3, 2) * CC(N - 4, 4) + CC(N - 3, 1) * CC(N - 5, 5));
  CC(N - 3, 4) = CC(N - 3, 4) - CC(N - 3, 2) * CC(N - 4, 5);
  CC(N - 3, 8) = CC(N - 3, 5) - (CC(N - 3, 2) * CC(N - 4, 8) + CC(N - 3, 1) * CC(N - 5, 8));
  CC(N - 3, 9) = -(CC(N - 3, 2) * CC(N - 4, 9) + CC(N - 3, 1) * CC(N - 5, 9));
  CC(N - 3, 3) = 1.0 / CC(N - 3, 3);
  CC(N - 3, 6) = (CC(N - 1, 1) - (CC(N - 4, 6) * CC(N - 4, 4) + CC(N - 5, 6) * CC(N - 5, 5))) * CC(N - 3, 3);
  CC(N - 3, 7) = -(CC(N - 4, 7) * CC(N - 4, 4) + CC(N - 5, 7) * CC(N - 5, 5)) * CC(N - 3, 3);

  CC(N - 2, 1) = CC(N - 2, 1) * CC(N - 4, 3);
  CC(N - 2, 2) = (CC(N - 2, 2) - CC(N - 2, 1) * CC(N - 4, 4)) * CC(N - 3, 3);
  CC(N - 2, 3) = CC(N - 2, 3) - (CC(N - 2, 2) * CC(N - 3, 4) + CC(N - 2, 1) * CC(N - 4, 5));
  CC(N - 2, 8) = CC(N - 2, 4) - (CC(N - 2, 2) * CC(N - 3, 8) + CC(N - 2, 1) * CC(N - 4, 8));
  CC(N - 2, 9) = CC(N - 2, 5) - (CC(N - 2, 2) * CC(N - 3, 9) + CC(N - 2, 1) * CC(N - 4, 9));
  CC(N - 2, 3) = 1.0 / CC(N - 2, 3);
  CC(N - 2, 6) = (CC(N - 1, 2) - CC(N - 3, 6) * CC(N - 3, 4) - CC(N - 4, 6) * CC(N - 4, 5)) * CC(N - 2, 3);
  CC(N - 2, 7) = (CC(N, 1) - CC(N - 3, 7) * CC(N - 3, 4) - CC(N - 4, 7) * CC(N - 4, 5)) * CC(N - 2, 3);

  for (int K = 1; K <= N - 2; ++K) {
    CC(N - 1, 3) = CC(N - 1, 3) - CC(K, 6) * CC(K, 8);
    CC(N - 1, 4) = CC(N - 1, 4) - CC(K, 6) * CC(K, 9);
  }
  CC(N - 1, 9) = CC(N - 1, 4);
  for (int K = 1; K <= N - 2; ++K) CC(N, 2) = CC(N, 2) - CC(K, 7) * CC(K, 8);
  CC(N - 1, 3) = 1.0 / CC(N - 1, 3);
  CC(N - 1, 7) = CC(N, 2) * CC(N - 1, 3);
  for (int K = 1; K <= N - 1; ++K) CC(N, 3) = CC(N, 3) - CC(K, 7) * CC(K, 9);
  CC(N, 3) = 1.0 / CC(N, 3);
}

/* ppentLUS1, pentadiagonal.f90:134-157 */
void ora_ppentLUS1(const double* c, double* r, int N) {
  const int n = N;
#define RR(i) r[(i)-1]
  RR(2) = RR(2) - CC(2, 2) * RR(1);
  for (int i = 3; i <= N - 2; ++i) RR(i) = RR(i) - (CC(i, 2) * RR(i - 1) + CC(i, 1) * RR(i - 2));
  for (int i = 1; i <= N - 2; ++i) RR(N - 1) = RR(N - 1) - CC(i, 6) * RR(i);
  for (int i = 1; i <= N - 1; ++i) RR(N) = RR(N) - CC(i, 7) * RR(i);
  RR(N) = RR(N) * CC(N, 3);
  RR(N - 1) = (RR(N - 1) - CC(N - 1, 9) * RR(N)) * CC(N - 1, 3);
  RR(N - 2) = (RR(N - 2) - CC(N - 2, 8) * RR(N - 1) - CC(N - 2, 9) * RR(N)) * CC(N - 2, 3);
  RR(N - 3) = (RR(N - 3) - (CC(N - 3, 4) * RR(N - 2) + CC(N - 3, 8) * RR(N - 1) + CC(N - 3, 9) * RR(N))) * CC(N - 3, 3);
  for (int i = N - 4; i >= 1; --i)
    RR(i) = (RR(i) - (CC(i, 4) * RR(i + 1) + CC(i, 5) * RR(i + 2) + CC(i, 8) * RR(n - 1) + CC(i, 9) * RR(n))) * CC(i, 3);
#undef RR
}

/* Per-line solves as the 3-D sweeps do them (pentadiagonal.f90:629-684). */
static void line_bpentLUS3(const double* c, double* r, int n) {
#define RR(i) r[(i)-1]
  for (int i = 1; i <= n - 2; ++i) {
    RR(i + 1) = RR(i + 1) - RR(i) * CC(i + 1, 2);
    RR(i + 2) = RR(i + 2) - RR(i) * CC(i + 2, 1);
  }
  RR(n) = (RR(n) - RR(n - 1) * CC(n, 2)) * CC(n, 3);
  RR(n - 1) = (RR(n - 1) - CC(n - 1, 4) * RR(n)) * CC(n - 1, 3);
  for (int i = n - 2; i >= 1; --i) RR(i) = (RR(i) - CC(i, 4) * RR(i + 1) - CC(i, 5) * RR(i + 2)) * CC(i, 3);
#undef RR
}

static void line_ppentLUS3(const double* c, double* r, int N) {
  const int n = N;
#define RR(i) r[(i)-1]
  RR(2) = RR(2) - CC(2, 2) * RR(1);
  double tmp1 = CC(1, 6) * RR(1) + CC(2, 6) * RR(2);
  double tmp2 = CC(1, 7) * RR(1) + CC(2, 7) * RR(2);
  for (int i = 3; i <= N - 2; ++i) {
    RR(i) = RR(i) - (CC(i, 2) * RR(i - 1) + CC(i, 1) * RR(i - 2));
    tmp1 = tmp1 + CC(i, 6) * RR(i);
    tmp2 = tmp2 + CC(i, 7) * RR(i);
  }
  RR(N - 1) = RR(N - 1) - tmp1;
  RR(N) = (RR(N) - tmp2 - CC(N - 1, 7) * RR(N - 1)) * CC(N, 3);
  RR(N - 1) = (RR(N - 1) - CC(N - 1, 9) * RR(N)) * CC(N - 1, 3);
  RR(N - 2) = (RR(N - 2) - CC(N - 2, 8) * RR(N - 1) - CC(N - 2, 9) * RR(N)) * CC(N - 2, 3);
  RR(N - 3) = (RR(N - 3) - (CC(N - 3, 4) * RR(N - 2) + CC(N - 3, 8) * RR(N - 1) + CC(N - 3, 9) * RR(N))) * CC(N - 3, 3);
  for (int i = N - 4; i >= 1; --i)
    RR(i) = (RR(i) - (CC(i, 4) * RR(i + 1) + CC(i, 5) * RR(i + 2) + CC(i, 8) * RR(n - 1) + CC(i, 9) * RR(n))) * CC(i, 3);
#undef RR
}
#undef CC

/* The per-line sweep of the 3-D solves, exported for pinning against _ref. */
void ora_line_solve(int periodic, const double* c, double* r, int n) {
  if (periodic) line_ppentLUS3(c, r, n);
  else line_bpentLUS3(c, r, n);
}

/* ------------------------------------------------------------------------- */
/* Band assembly, compact_basetype.f90:124-148 and :165-170 (np=1).           */
/* ------------------------------------------------------------------------- */
int ora_assemble(int op, int n, int periodic, int bc1, int bc2, double* al, double* ar) {
  ora_weight w;
  ora_get_weight(op, &w);
  const int nl = w.ncl, nr = w.ncr;
  const int nal = periodic ? nl + 4 : nl;
  /* al(m,nal) column-major = transpose(gal(nal,n)) */
  for (int c = 0; c < nal; ++c)
    for (int i = 0; i < n; ++i) al[(size_t)c * n + i] = (c < nl) ? w.ali[c] : 0.0;
  for (int i = 0; i < n; ++i)
    for (int l = 0; l < nr; ++l) ar[(size_t)i * nr + l] = w.ari[l];
  if (!periodic) {
    for (int j = 0; j < 4; ++j) {
      for (int c = 0; c < nl; ++c) {
        al[(size_t)c * n + j] = w.alb1[bc1 + 1][j][c];
        al[(size_t)c * n + (n - 4 + j)] = w.alb2[bc2 + 1][j][c];
      }
      for (int l = 0; l < nr; ++l) {
        ar[(size_t)j * nr + l] = w.arb1[bc1 + 1][j][l];
        ar[(size_t)(n - 4 + j) * nr + l] = w.arb2[bc2 + 1][j][l];
      }
    }
  }
  if (w.implicit_op) {
    if (periodic) ora_ppentLUD1(al, n);
    else ora_bpentLUD1(al, n);
  }
  return nal;
}

/* ------------------------------------------------------------------------- */
/* One operator on one line: RHS (reference forms) + solve + (+v).           */
/* ------------------------------------------------------------------------- */
typedef struct {
  int op, n, periodic, bc1, bc2, null_op, null_option, implicit_op, nor, nr;
  double* al; /* (n, nal) */
  double* ar; /* (nr, n) */
} ora_op1;

static void op1_setup(ora_op1* o, int op, int n, int periodic, int bc1, int bc2) {
  ora_weight w;
  ora_get_weight(op, &w);
  memset(o, 0, sizeof(*o));
  o->op = op; o->n = n; o->periodic = periodic; o->bc1 = bc1; o->bc2 = bc2;
  o->null_option = w.null_option; o->implicit_op = w.implicit_op;
  o->nor = w.nor; o->nr = w.ncr;
  o->null_op = (n < 4); /* compact_basetype.f90:86-90 */
  if (o->null_op) return;
  o->al = (double*)malloc(sizeof(double) * (size_t)n * 9);
  o->ar = (double*)malloc(sizeof(double) * (size_t)n * 9);
  ora_assemble(op, n, periodic, bc1, bc2, o->al, o->ar);
}

static void op1_free(ora_op1* o) { free(o->al); free(o->ar); o->al = o->ar = NULL; }

#define V(i) v[(i)-1]
#define DV(i) dv[(i)-1]

/* compact_d1.f90:100-227 (x version; y/z identical per line) */
static void rhs_d1(const ora_op1* o, const double* v, double* dv) {
  const int n = o->n, nr = o->nr;
  const double* ar = o->ar;
#define AR(l, i) ar[((size_t)(i)-1) * nr + ((l)-1)]
  double vbr1[3] = {0, 0, 0}, vbr2[3] = {0, 0, 0};
  if (o->periodic) {
    vbr1[0] = V(n - 2); vbr1[1] = V(n - 1); vbr1[2] = V(n);
    vbr2[0] = V(1); vbr2[1] = V(2); vbr2[2] = V(3);
  }
  if (!o->periodic) {
    if (o->bc1 == -1) {
      double s = 0; for (int l = 4; l <= 7; ++l) s += AR(l, 1) * V(l - 3); DV(1) = s;
      s = 0; for (int l = 3; l <= 7; ++l) s += AR(l, 2) * V(l - 2); DV(2) = s;
      s = 0; for (int l = 2; l <= 7; ++l) s += AR(l, 3) * V(l - 1); DV(3) = s;
    } else {
      double s = 0; for (int l = 5; l <= 7; ++l) s += AR(l, 1) * (V(l - 3) - V(1)); DV(1) = s;
      s = 0; for (int l = 4; l <= 7; ++l) s += AR(l, 2) * (V(l - 2) - V(1)); DV(2) = s;
      DV(3) = AR(5, 3) * (V(4) - V(2)) + AR(6, 3) * (V(5) - V(1)) + AR(7, 3) * (V(6) - V(1));
    }
  } else {
    DV(1) = AR(5, 1) * (V(2) - vbr1[2]) + AR(6, 1) * (V(3) - vbr1[1]) + AR(7, 1) * (V(4) - vbr1[0]);
    DV(2) = AR(5, 2) * (V(3) - V(1)) + AR(6, 2) * (V(4) - vbr1[2]) + AR(7, 2) * (V(5) - vbr1[1]);
    DV(3) = AR(5, 3) * (V(4) - V(2)) + AR(6, 3) * (V(5) - V(1)) + AR(7, 3) * (V(6) - vbr1[2]);
  }
  for (int i = 4; i <= n - 3; ++i)
    DV(i) = AR(5, i) * (V(i + 1) - V(i - 1)) + AR(6, i) * (V(i + 2) - V(i - 2)) + AR(7, i) * (V(i + 3) - V(i - 3));
  if (!o->periodic) {
    if (o->bc2 == -1) {
      double s = 0; for (int l = 1; l <= 6; ++l) s += AR(l, n - 2) * V(n - 6 + l); DV(n - 2) = s;
      s = 0; for (int l = 1; l <= 5; ++l) s += AR(l, n - 1) * V(n - 5 + l); DV(n - 1) = s;
      s = 0; for (int l = 1; l <= 4; ++l) s += AR(l, n) * V(n - 4 + l); DV(n) = s;
    } else {
      DV(n - 2) = AR(1, n - 2) * (V(n - 5) - V(n)) + AR(2, n - 2) * (V(n - 4) - V(n)) + AR(3, n - 2) * (V(n - 3) - V(n - 1));
      double s = 0; for (int l = 1; l <= 4; ++l) s += AR(l, n - 1) * (V(n - 5 + l) - V(n)); DV(n - 1) = s;
      s = 0; for (int l = 1; l <= 3; ++l) s += AR(l, n) * (V(n - 4 + l) - V(n)); DV(n) = s;
    }
  } else {
    DV(n - 2) = AR(5, n - 2) * (V(n - 1) - V(n - 3)) + AR(6, n - 2) * (V(n) - V(n - 4)) + AR(7, n - 2) * (vbr2[0] - V(n - 5));
    DV(n - 1) = AR(5, n - 1) * (V(n) - V(n - 2)) + AR(6, n - 1) * (vbr2[0] - V(n - 3)) + AR(7, n - 1) * (vbr2[1] - V(n - 4));
    DV(n) = AR(5, n) * (vbr2[0] - V(n - 1)) + AR(6, n) * (vbr2[1] - V(n - 2)) + AR(7, n) * (vbr2[2] - V(n - 3));
  }
#undef AR
}

/* compact_r3.f90:73-157 */
static void rhs_r3(const ora_op1* o, const double* v, double* dv) {
  const int n = o->n, nr = o->nr;
  const double* ar = o->ar;
#define AR(l, i) ar[((size_t)(i)-1) * nr + ((l)-1)]
  double vbr1[3] = {0, 0, 0}, vbr2[3] = {0, 0, 0};
  if (o->periodic) {
    vbr1[0] = V(n - 2); vbr1[1] = V(n - 1); vbr1[2] = V(n);
    vbr2[0] = V(1); vbr2[1] = V(2); vbr2[2] = V(3);
  }
  const int anti1 = (!o->periodic && o->bc1 == -1), anti2 = (!o->periodic && o->bc2 == -1);
  {
    double s1 = 0, s2 = 0;
    for (int l = 1; l <= 3; ++l) s1 += AR(l, 1) * (vbr1[l - 1] - V(1));
    for (int l = 4; l <= 7; ++l) s2 += AR(l, 1) * (V(l - 3) - V(1));
    DV(1) = s1 + s2;
    if (anti1) { double sr = 0; for (int l = 1; l <= 7; ++l) sr += AR(l, 1); DV(1) = DV(1) + sr * V(1); }
    s1 = 0; s2 = 0;
    for (int l = 1; l <= 2; ++l) s1 += AR(l, 2) * (vbr1[l] - V(2));
    for (int l = 3; l <= 7; ++l) s2 += AR(l, 2) * (V(l - 2) - V(2));
    DV(2) = s1 + s2;
    if (anti1) { double sr = 0; for (int l = 1; l <= 7; ++l) sr += AR(l, 2); DV(2) = DV(2) + sr * V(2); }
    s1 = AR(1, 3) * (vbr1[2] - V(3)); s2 = 0;
    for (int l = 2; l <= 7; ++l) s2 += AR(l, 3) * (V(l - 1) - V(3));
    DV(3) = s1 + s2;
    if (anti1) { double sr = 0; for (int l = 1; l <= 7; ++l) sr += AR(l, 3); DV(3) = DV(3) + sr * V(3); }
  }
  for (int i = 4; i <= n - 3; ++i) {
    double s = 0;
    for (int l = 1; l <= 7; ++l) s += AR(l, i) * (V(i - 4 + l) - V(i));
    DV(i) = s;
  }
  {
    double s1 = 0, s2 = 0;
    for (int l = 1; l <= 6; ++l) s1 += AR(l, n - 2) * (V(n - 6 + l) - V(n - 2));
    s2 = AR(7, n - 2) * (vbr2[0] - V(n - 2));
    DV(n - 2) = s1 + s2;
    if (anti2) { double sr = 0; for (int l = 1; l <= 7; ++l) sr += AR(l, n - 2); DV(n - 2) = DV(n - 2) + sr * V(n - 2); }
    s1 = 0; s2 = 0;
    for (int l = 1; l <= 5; ++l) s1 += AR(l, n - 1) * (V(n - 5 + l) - V(n - 1));
    for (int l = 6; l <= 7; ++l) s2 += AR(l, n - 1) * (vbr2[l - 6] - V(n - 1));
    DV(n - 1) = s1 + s2;
    if (anti2) { double sr = 0; for (int l = 1; l <= 7; ++l) sr += AR(l, n - 1); DV(n - 1) = DV(n - 1) + sr * V(n - 1); }
    s1 = 0; s2 = 0;
    for (int l = 1; l <= 4; ++l) s1 += AR(l, n) * (V(n - 4 + l) - V(n));
    for (int l = 5; l <= 7; ++l) s2 += AR(l, n) * (vbr2[l - 5] - V(n));
    DV(n) = s1 + s2;
    if (anti2) { double sr = 0; for (int l = 1; l <= 7; ++l) sr += AR(l, n); DV(n) = DV(n) + sr * V(n); }
  }
#undef AR
}

/* compact_r4.f90:79-207 */
static void rhs_r4(const ora_op1* o, const double* v, double* dv) {
  const int n = o->n, nr = o->nr;
  const double* ar = o->ar;
#define AR(l, i) ar[((size_t)(i)-1) * nr + ((l)-1)]
  double vbr1[4] = {0, 0, 0, 0}, vbr2[4] = {0, 0, 0, 0};
  if (o->periodic) {
    for (int q = 0; q < 4; ++q) { vbr1[q] = V(n - 3 + q); vbr2[q] = V(1 + q); }
  }
  if (!o->periodic) {
    if (o->bc1 == -1) {
      double s = 0; for (int l = 5; l <= 9; ++l) s += AR(l, 1) * V(l - 4); DV(1) = s;
      s = 0; for (int l = 4; l <= 9; ++l) s += AR(l, 2) * V(l - 3); DV(2) = s;
      s = 0; for (int l = 3; l <= 9; ++l) s += AR(l, 3) * V(l - 2); DV(3) = s;
      s = 0; for (int l = 2; l <= 9; ++l) s += AR(l, 4) * V(l - 1); DV(4) = s;
    } else {
      double s = 0; for (int l = 6; l <= 9; ++l) s += AR(l, 1) * (V(l - 4) - V(1)); DV(1) = s;
      s = 0; for (int l = 5; l <= 9; ++l) s += AR(l, 2) * (V(l - 3) - V(1)); DV(2) = s;
      s = 0; for (int l = 4; l <= 9; ++l) s += AR(l, 3) * (V(l - 2) - V(1)); DV(3) = s;
      s = 0; for (int l = 3; l <= 9; ++l) s += AR(l, 4) * (V(l - 1) - V(1)); DV(4) = s;
    }
  } else {
    double s1 = 0, s2 = 0;
    for (int l = 1; l <= 4; ++l) s1 += AR(l, 1) * (vbr1[l - 1] - V(1));
    for (int l = 5; l <= 9; ++l) s2 += AR(l, 1) * (V(l - 4) - V(1));
    DV(1) = s1 + s2;
    s1 = 0; s2 = 0;
    for (int l = 1; l <= 3; ++l) s1 += AR(l, 2) * (vbr1[l] - V(2));
    for (int l = 4; l <= 9; ++l) s2 += AR(l, 2) * (V(l - 3) - V(2));
    DV(2) = s1 + s2;
    s1 = 0; s2 = 0;
    for (int l = 1; l <= 2; ++l) s1 += AR(l, 3) * (vbr1[l + 1] - V(3));
    for (int l = 3; l <= 9; ++l) s2 += AR(l, 3) * (V(l - 2) - V(3));
    DV(3) = s1 + s2;
    s1 = AR(1, 4) * (vbr1[3] - V(4)); s2 = 0;
    for (int l = 2; l <= 9; ++l) s2 += AR(l, 4) * (V(l - 1) - V(4));
    DV(4) = s1 + s2;
  }
  for (int i = 5; i <= n - 4; ++i) {
    double s = 0.0;
    const double vc = V(i);
    for (int l = 1; l <= 9; ++l) s = s + (V(l + i - 5) - vc) * AR(l, i);
    DV(i) = s;
  }
  if (!o->periodic) {
    if (o->bc2 == -1) {
      double s = 0; for (int l = 1; l <= 8; ++l) s += AR(l, n - 3) * V(n - 8 + l); DV(n - 3) = s;
      s = 0; for (int l = 1; l <= 7; ++l) s += AR(l, n - 2) * V(n - 7 + l); DV(n - 2) = s;
      s = 0; for (int l = 1; l <= 6; ++l) s += AR(l, n - 1) * V(n - 6 + l); DV(n - 1) = s;
      s = 0; for (int l = 1; l <= 5; ++l) s += AR(l, n) * V(n - 5 + l); DV(n) = s;
    } else {
      double s = 0; for (int l = 1; l <= 7; ++l) s += AR(l, n - 3) * (V(n - 8 + l) - V(n)); DV(n - 3) = s;
      s = 0; for (int l = 1; l <= 6; ++l) s += AR(l, n - 2) * (V(n - 7 + l) - V(n)); DV(n - 2) = s;
      s = 0; for (int l = 1; l <= 5; ++l) s += AR(l, n - 1) * (V(n - 6 + l) - V(n)); DV(n - 1) = s;
      s = 0; for (int l = 1; l <= 4; ++l) s += AR(l, n) * (V(n - 5 + l) - V(n)); DV(n) = s;
    }
  } else {
    double s1 = 0, s2 = 0;
    for (int l = 1; l <= 8; ++l) s1 += AR(l, n - 3) * (V(n - 8 + l) - V(n - 3));
    s2 = AR(9, n - 3) * (vbr2[0] - V(n - 3));
    DV(n - 3) = s1 + s2;
    s1 = 0; s2 = 0;
    for (int l = 1; l <= 7; ++l) s1 += AR(l, n - 2) * (V(n - 7 + l) - V(n - 2));
    for (int l = 8; l <= 9; ++l) s2 += AR(l, n - 2) * (vbr2[l - 8] - V(n - 2));
    DV(n - 2) = s1 + s2;
    s1 = 0; s2 = 0;
    for (int l = 1; l <= 6; ++l) s1 += AR(l, n - 1) * (V(n - 6 + l) - V(n - 1));
    for (int l = 7; l <= 9; ++l) s2 += AR(l, n - 1) * (vbr2[l - 7] - V(n - 1));
    DV(n - 1) = s1 + s2;
    s1 = 0; s2 = 0;
    for (int l = 1; l <= 5; ++l) s1 += AR(l, n) * (V(n - 5 + l) - V(n));
    for (int l = 6; l <= 9; ++l) s2 += AR(l, n) * (vbr2[l - 6] - V(n));
    DV(n) = s1 + s2;
  }
#undef AR
}
#undef V
#undef DV

/* One line: RHS, solve, +v (compact_r4.f90:209-233), null options. */
static void op1_line(const ora_op1* o, const double* v, double* dv) {
  const int n = o->n;
  if (o->null_op) {
    for (int i = 0; i < n; ++i) dv[i] = (o->null_option == 1) ? v[i] : 0.0;
    return;
  }
  if (o->op == ORA_D1) rhs_d1(o, v, dv);
  else if (o->op == ORA_D2) rhs_r3(o, v, dv);
  else rhs_r4(o, v, dv);
  if (o->implicit_op) {
    if (o->periodic) line_ppentLUS3(o->al, dv, n);
    else line_bpentLUS3(o->al, dv, n);
  }
  if (o->null_option == 1)
    for (int i = 0; i < n; ++i) dv[i] = dv[i] + v[i];
}

/* ------------------------------------------------------------------------- */
/* Context: operator suite per direction (compact.f90:55-319 for np=1).       */
/* ------------------------------------------------------------------------- */
struct ora_ctx {
  int n[3];
  double d[3];      /* nominal spacing for the metric (compact_operators) */
  double mesh_d[3]; /* mesh d1,d2,d3 (mesh.f90:190-211) */
  ora_op1 ops[ORA_NOPS][3];
  /* second member of the symmetric/antisymmetric pair (compact.f90:78-84):
     variant -1 at each SYMM end, built only where a direction has one */
  int has_anti[3];
  ora_op1 anti[ORA_NOPS][3];
};

ora_ctx* ora_create(int nx, int ny, int nz, double dx, double dy, double dz, const int* bc6) {
  ora_ctx* c = (ora_ctx*)calloc(1, sizeof(ora_ctx));
  c->n[0] = nx; c->n[1] = ny; c->n[2] = nz;
  c->d[0] = dx; c->d[1] = dy; c->d[2] = dz;
  c->mesh_d[0] = (nx == 1) ? fmax(dy, dz) : dx;
  c->mesh_d[1] = (ny == 1) ? fmax(dx, dz) : dy;
  c->mesh_d[2] = (nz == 1) ? fmax(dx, dy) : dz;
  for (int dir = 0; dir < 3; ++dir) {
    const int b1 = bc6[2 * dir], b2 = bc6[2 * dir + 1];
    const int periodic = (b1 == ORA_BC_PERI);
    /* SYMM selects the +1 variant for iop=1 (compact.f90:79-84) */
    const int v1 = (b1 == ORA_BC_SYMM) ? 1 : 0;
    const int v2 = (b2 == ORA_BC_SYMM) ? 1 : 0;
    for (int op = 0; op < ORA_NOPS; ++op) op1_setup(&c->ops[op][dir], op, c->n[dir], periodic, v1, v2);
    c->has_anti[dir] = !periodic && (v1 || v2);
    if (c->has_anti[dir])
      for (int op = 0; op < ORA_NOPS; ++op) op1_setup(&c->anti[op][dir], op, c->n[dir], 0, -v1, -v2);
  }
  return c;
}

void ora_destroy(ora_ctx* c) {
  if (!c) return;
  for (int op = 0; op < ORA_NOPS; ++op)
    for (int d = 0; d < 3; ++d) {
      op1_free(&c->ops[op][d]);
      if (c->has_anti[d]) op1_free(&c->anti[op][d]);
    }
  free(c);
}

/* isym = -1 selects the antisymmetric member of a SYMM pair where the
   direction has one, as d1x/d8x(...,bc) do (compact_operators.f90:300-303);
   otherwise (and for isym = +1) the default operator. */
int ora_apply_sym(ora_ctx* c, int op, int dir, int isym, const double* v, double* dv) {
  if (op < 0 || op >= ORA_NOPS || dir < 0 || dir > 2) return -1;
  const int nx = c->n[0], ny = c->n[1], nz = c->n[2];
  const int n = c->n[dir];
  const ora_op1* o = (isym == -1 && c->has_anti[dir]) ? &c->anti[op][dir] : &c->ops[op][dir];
  const size_t stride = dir == 0 ? 1 : (dir == 1 ? (size_t)nx : (size_t)nx * ny);
  const int n1 = dir == 0 ? ny : nx;
  const int n2 = dir == 2 ? ny : nz;
  const long nlines = (long)n1 * n2;
  /* metric: d1 divides by dx (compact_operators.f90:41-45), d2 by dx**2 (:152) */
  const double d = c->d[dir];
  const int scale_kind = (op == ORA_D1) ? 1 : (op == ORA_D2 ? 2 : 0);
  const double d2 = d * d;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (long L = 0; L < nlines; ++L) {
    double vb[4096], db[4096];
    const long a = L % n1, b = L / n1;
    size_t base;
    if (dir == 0) base = (size_t)nx * (a + (size_t)ny * b);
    else if (dir == 1) base = a + (size_t)nx * ny * b;
    else base = a + (size_t)nx * b;
    for (int i = 0; i < n; ++i) vb[i] = v[base + stride * i];
    op1_line(o, vb, db);
    if (!o->null_op) {
      if (scale_kind == 1) for (int i = 0; i < n; ++i) db[i] = db[i] / d;
      else if (scale_kind == 2) for (int i = 0; i < n; ++i) db[i] = db[i] / d2;
    }
    for (int i = 0; i < n; ++i) dv[base + stride * i] = db[i];
  }
  return 0;
}

int ora_apply(ora_ctx* c, int op, int dir, const double* v, double* dv) {
  return ora_apply_sym(c, op, dir, 1, v, dv);
}

/* isymX/Y/Z of a patch: -1 where the direction has a SYMM end (patch.f90:89-91) */
static int isym(const ora_ctx* c, int dir) { return c->has_anti[dir] ? -1 : 1; }

static size_t npts(const ora_ctx* c) { return (size_t)c->n[0] * c->n[1] * c->n[2]; }

/* filter('spectral'|'smooth'), operators.f90:847-875: x -> y -> z */
static int filter3(ora_ctx* c, int op, const double* v, double* out) {
  double* tmp = (double*)malloc(sizeof(double) * npts(c));
  ora_apply(c, op, 0, v, out);
  ora_apply(c, op, 1, out, tmp);
  ora_apply(c, op, 2, tmp, out);
  free(tmp);
  return 0;
}
int ora_sfilter(ora_ctx* c, const double* v, double* out) { return filter3(c, ORA_SF, v, out); }
int ora_gfilter(ora_ctx* c, const double* v, double* out) { return filter3(c, ORA_GF, v, out); }
int ora_gfilterdir(ora_ctx* c, const double* v, double* out, int dir1) {
  return ora_apply(c, ORA_GF, dir1 - 1, v, out);
}

/* ringS with L=2, operators.f90:615-643 (ringx/y/z :701-753) */
int ora_ring(ora_ctx* c, const double* v, double* out) {
  const size_t N = npts(c);
  double* r[3];
  for (int d = 0; d < 3; ++d) {
    r[d] = (double*)malloc(sizeof(double) * N);
    if (c->n[d] == 1) memset(r[d], 0, sizeof(double) * N);
    else ora_apply(c, ORA_D8, d, v, r[d]);
  }
  const double s0 = c->mesh_d[0] * c->mesh_d[0], s1 = c->mesh_d[1] * c->mesh_d[1],
               s2 = c->mesh_d[2] * c->mesh_d[2];
  for (size_t i = 0; i < N; ++i) {
    double a = fabs(r[0][i]) * s0, b = fabs(r[1][i]) * s1, e = fabs(r[2][i]) * s2;
    double m = a > b ? a : b;
    out[i] = m > e ? m : e;
  }
  for (int d = 0; d < 3; ++d) free(r[d]);
  return 0;
}

/* ringV with L=1, operators.f90:645-699; the normal component of each
   direction takes the antisymmetric operator at SYMM ends (:661-671) */
int ora_ringv(ora_ctx* c, const double* vx, const double* vy, const double* vz, double* out) {
  const size_t N = npts(c);
  const double* f[3] = {vx, vy, vz};
  double* r[3][3];
  for (int comp = 0; comp < 3; ++comp)
    for (int d = 0; d < 3; ++d) {
      r[comp][d] = (double*)malloc(sizeof(double) * N);
      if (c->n[d] == 1) memset(r[comp][d], 0, sizeof(double) * N);
      else ora_apply_sym(c, ORA_D8, d, comp == d ? isym(c, d) : 1, f[comp], r[comp][d]);
    }
  for (size_t i = 0; i < N; ++i) {
    double L[3];
    for (int d = 0; d < 3; ++d) {
      double a = fabs(r[0][d][i]), b = fabs(r[1][d][i]), e = fabs(r[2][d][i]);
      double m = a > b ? a : b;
      m = m > e ? m : e;
      L[d] = m * c->mesh_d[d];
    }
    double m = L[0] > L[1] ? L[0] : L[1];
    out[i] = m > L[2] ? m : L[2];
  }
  for (int comp = 0; comp < 3; ++comp)
    for (int d = 0; d < 3; ++d) free(r[comp][d]);
  return 0;
}

/* LapS Cartesian, operators.f90:522-526 */
int ora_laplacian(ora_ctx* c, const double* v, double* out) {
  const size_t N = npts(c);
  double* t = (double*)malloc(sizeof(double) * N);
  double* u = (double*)malloc(sizeof(double) * N);
  ora_apply(c, ORA_D2, 0, v, out);
  ora_apply(c, ORA_D2, 1, v, t);
  ora_apply(c, ORA_D2, 2, v, u);
  for (size_t i = 0; i < N; ++i) out[i] = out[i] + t[i] + u[i];
  free(t); free(u);
  return 0;
}

/* divV Cartesian, operators.f90:38-53 (ddx(fx,..,isymX) etc.) */
int ora_divergence(ora_ctx* c, const double* fx, const double* fy, const double* fz, double* out) {
  const size_t N = npts(c);
  double* a = (double*)malloc(sizeof(double) * N);
  double* b = (double*)malloc(sizeof(double) * N);
  double* e = (double*)malloc(sizeof(double) * N);
  ora_apply_sym(c, ORA_D1, 0, isym(c, 0), fx, a);
  ora_apply_sym(c, ORA_D1, 1, isym(c, 1), fy, b);
  ora_apply_sym(c, ORA_D1, 2, isym(c, 2), fz, e);
  for (size_t i = 0; i < N; ++i) out[i] = a[i] + b[i] + e[i];
  free(a); free(b); free(e);
  return 0;
}

/* gradS Cartesian, operators.f90:184-212 */
int ora_grads(ora_ctx* c, const double* v, double* gx, double* gy, double* gz) {
  ora_apply(c, ORA_D1, 0, v, gx);
  ora_apply(c, ORA_D1, 1, v, gy);
  ora_apply(c, ORA_D1, 2, v, gz);
  return 0;
}
