// Compact-scheme weight tables (host side).
//
// The coefficient values and the symmetric-boundary folding are fixed by the
// reference schemes, so the tables are alike by necessity:
//   c10d1  pyranda/parcop/stencils.f90:207-277   (10th-order Pade d/dx)
//   c10d2  pyranda/parcop/stencils.f90:358-428   (10th-order d2/dx2)
//   c10d8  pyranda/parcop/stencils.f90:515-621   (8th derivative, LAD ring)
//   c8ff8  pyranda/parcop/stencils.f90:713-835   (8th-order compact filter, fbar)
//   cgfs4  pyranda/parcop/stencils.f90:1387-1476 (9-point Gaussian, gbar)
//   folding of the SYMM variants: stencils.f90:2390-2453
// Boundary tables are kept per variant v in {-1,0,+1} (index v+1); the XTND
// variant (2) is unreachable from pyranda's BC strings and is not built.
#pragma once

#include <cstring>

namespace pyr {

enum OpId { OP_D1 = 0, OP_D2 = 1, OP_D8 = 2, OP_SF = 3, OP_GF = 4, NOPS = 5 };

struct Scheme {
  int nol = 0, nor = 0, ncl = 0, ncr = 0, null_option = 0, implicit = 0;
  double ali[5] = {0, 0, 0, 0, 0};
  double ari[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double alb1[3][4][5] = {}, alb2[3][4][5] = {};
  double arb1[3][4][9] = {}, arb2[3][4][9] = {};
};

namespace detail {

// fold rows into the interior weights for a mirror at the face (lower end)
inline void fold_lower(double alb[4][5], double arb[4][9], const Scheme& s, int syml, int symr) {
  for (int r = 0; r < 4; ++r) {
    for (int c = 0; c < s.ncl; ++c) alb[r][c] = s.ali[c];
    for (int c = 0; c < s.ncr; ++c) arb[r][c] = s.ari[c];
  }
  for (int j = 1; j <= s.nol; ++j) {
    const int k = s.nol - j + 1;
    for (int i = 1; i <= k; ++i) {
      const int from = k - i, to = k + i - 1;  // 0-based columns
      alb[j - 1][to] += double(syml) * alb[j - 1][from];
      alb[j - 1][from] = 0.0;
    }
  }
  for (int j = 1; j <= s.nor; ++j) {
    const int k = s.nor - j + 1;
    for (int i = 1; i <= k; ++i) {
      const int from = k - i, to = k + i - 1;
      arb[j - 1][to] += double(symr) * arb[j - 1][from];
      arb[j - 1][from] = 0.0;
    }
  }
}

// same at the upper end (rows 1..4 are the last four rows, row 4 = last)
inline void fold_upper(double alb[4][5], double arb[4][9], const Scheme& s, int syml, int symr) {
  for (int r = 0; r < 4; ++r) {
    for (int c = 0; c < s.ncl; ++c) alb[r][c] = s.ali[c];
    for (int c = 0; c < s.ncr; ++c) arb[r][c] = s.ari[c];
  }
  for (int j = 1; j <= s.nol; ++j) {
    const int row = 4 - j, k = s.nol - j + 1, off = s.ncl - k;
    for (int i = 1; i <= k; ++i) {
      alb[row][off - i] += double(syml) * alb[row][off + i - 1];
      alb[row][off + i - 1] = 0.0;
    }
  }
  for (int j = 4 - s.nor + 1; j <= 4; ++j) {
    const int k = j - 4 + s.nor, off = s.ncr - k;
    for (int i = 1; i <= k; ++i) {
      arb[j - 1][off - i] += double(symr) * arb[j - 1][off + i - 1];
      arb[j - 1][off + i - 1] = 0.0;
    }
  }
}

inline void mirror_upper(Scheme& s, double rsign) {
  for (int j = 0; j < 4; ++j) {
    for (int c = 0; c < s.ncl; ++c) s.alb2[1][j][c] = s.alb1[1][3 - j][s.ncl - 1 - c];
    for (int c = 0; c < s.ncr; ++c) s.arb2[1][j][c] = rsign * s.arb1[1][3 - j][s.ncr - 1 - c];
  }
}

inline void symm_variants(Scheme& s, int lhs_sign_plus, int rhs_sign_plus) {
  // variant +1 (index 2) uses (lhs_sign_plus, rhs_sign_plus); variant -1 flips both
  fold_lower(s.alb1[2], s.arb1[2], s, lhs_sign_plus, rhs_sign_plus);
  fold_lower(s.alb1[0], s.arb1[0], s, -lhs_sign_plus, -rhs_sign_plus);
  fold_upper(s.alb2[2], s.arb2[2], s, lhs_sign_plus, rhs_sign_plus);
  fold_upper(s.alb2[0], s.arb2[0], s, -lhs_sign_plus, -rhs_sign_plus);
}

inline void set_rows5(double dst[4][5], const double src[4][5]) { std::memcpy(dst, src, sizeof(double) * 20); }
inline void set_rows9(double dst[4][9], const double src[4][9]) { std::memcpy(dst, src, sizeof(double) * 36); }

}  // namespace detail

inline Scheme make_scheme(int op) {
  using namespace detail;
  Scheme s;
  switch (op) {
    case OP_D1: {
      const double ali[5] = {0.45, 4.5, 9.0, 4.5, 0.45};
      const double ari[9] = {-0.015, -1.515, -6.375, 0.0, 6.375, 1.515, 0.015, 0, 0};
      const double alb[4][5] = {{0.0, 0.0, 4.725, 9.45, 0.0},
                                {0.0, 1.94578125, 7.783125, 1.94578125, 0.0},
                                {0.2964375, 4.743, 10.67175, 4.743, 0.2964375},
                                {0.451390625, 4.63271875, 9.38146875, 4.63271875, 0.451390625}};
      const double arb[4][9] = {{0.0, 0.0, 0.0, -11.8125, 9.45, 2.3625, 0.0},
                                {0.0, 0.0, -5.83734375, 0.0, 5.83734375, 0.0, 0.0},
                                {0.0, -1.23515625, -7.905, 0.0, 7.905, 1.23515625, 0.0},
                                {-0.015, -1.53, -6.66984375, 0.0, 6.66984375, 1.53, 0.015}};
      s.nol = 2; s.nor = 3; s.ncl = 5; s.ncr = 7; s.null_option = 0; s.implicit = 1;
      std::memcpy(s.ali, ali, sizeof ali); std::memcpy(s.ari, ari, sizeof ari);
      set_rows5(s.alb1[1], alb); set_rows9(s.arb1[1], arb);
      mirror_upper(s, -1.0);   // antisymmetric operator: upper rows change sign
      symm_variants(s, -1, +1);  // even function -> odd derivative
      break;
    }
    case OP_D2: {
      const double ali[5] = {387.0, 6012.0, 16182.0, 6012.0, 387.0};
      const double ari[9] = {79.0, 4671.0, 9585.0, -28670.0, 9585.0, 4671.0, 79.0, 0, 0};
      const double alb[4][5] = {{0.0, 0.0, 1.0, 11.0, 0.0},
                                {0.0, 1.0, 10.0, 1.0, 0.0},
                                {23.0, 688.0, 2358.0, 688.0, 23.0},
                                {387.0, 6012.0, 16182.0, 6012.0, 387.0}};
      const double arb[4][9] = {{0.0, 0.0, 0.0, 13.0, -27.0, 15.0, -1.0},
                                {0.0, 0.0, 12.0, -24.0, 12.0, 0.0, 0.0},
                                {0.0, 465.0, 1920.0, -4770.0, 1920.0, 465.0, 0.0},
                                {79.0, 4671.0, 9585.0, -28670.0, 9585.0, 4671.0, 79.0}};
      s.nol = 2; s.nor = 3; s.ncl = 5; s.ncr = 7; s.null_option = 0; s.implicit = 1;
      std::memcpy(s.ali, ali, sizeof ali); std::memcpy(s.ari, ari, sizeof ari);
      set_rows5(s.alb1[1], alb); set_rows9(s.arb1[1], arb);
      mirror_upper(s, 1.0);
      symm_variants(s, +1, +1);
      break;
    }
    case OP_D8: {
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
      s.nol = 2; s.nor = 4; s.ncl = 5; s.ncr = 9; s.null_option = 0; s.implicit = 1;
      std::memcpy(s.ali, ali, sizeof ali); std::memcpy(s.ari, ari, sizeof ari);
      set_rows5(s.alb1[1], alb); set_rows9(s.arb1[1], arb);
      mirror_upper(s, 1.0);
      symm_variants(s, +1, +1);
      break;
    }
    case OP_SF: {
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
      s.nol = 2; s.nor = 4; s.ncl = 5; s.ncr = 9; s.null_option = 1; s.implicit = 1;
      std::memcpy(s.ali, ali, sizeof ali); std::memcpy(s.ari, ari, sizeof ari);
      set_rows5(s.alb1[1], alb); set_rows9(s.arb1[1], arb);
      mirror_upper(s, 1.0);
      symm_variants(s, +1, +1);
      // stored as (rhs - lhs) so the filter is v + A^-1 (B-A) v
      for (int c = 0; c < 5; ++c) s.ari[2 + c] -= s.ali[c];
      for (int v = 0; v < 3; ++v)
        for (int j = 0; j < 4; ++j)
          for (int c = 0; c < 5; ++c) {
            s.arb1[v][j][2 + c] -= s.alb1[v][j][c];
            s.arb2[v][j][2 + c] -= s.alb2[v][j][c];
          }
      break;
    }
    case OP_GF: {
      const double agau = 3565.0 / 10368.0, bgau = 3091.0 / 12960.0, cgau = 1997.0 / 25920.0,
                   dgau = 149.0 / 12960.0, egau = 107.0 / 103680.0;
      const double dg14 = dgau + egau, cg13 = cgau + dgau, bg23 = bgau + egau, bg12 = bgau + cgau,
                   ag22 = agau + dgau, bg32 = bgau + egau, ag11 = agau + bgau, bg21 = bgau + cgau,
                   cg31 = cgau + dgau, dg41 = dgau + egau;
      const double ari[9] = {egau, dgau, cgau, bgau, agau, bgau, cgau, dgau, egau};
      const double arb[4][9] = {{0.0, 0.0, 0.0, 0.0, ag11, bg21, cg31, dg41, egau},
                                {0.0, 0.0, 0.0, bg12, ag22, bg32, cgau, dgau, egau},
                                {0.0, 0.0, cg13, bg23, agau, bgau, cgau, dgau, egau},
                                {0.0, dg14, cgau, bgau, agau, bgau, cgau, dgau, egau}};
      s.nol = 0; s.nor = 4; s.ncl = 1; s.ncr = 9; s.null_option = 1; s.implicit = 0;
      s.ali[0] = 1.0;
      std::memcpy(s.ari, ari, sizeof ari);
      for (int j = 0; j < 4; ++j) s.alb1[1][j][0] = 1.0;
      set_rows9(s.arb1[1], arb);
      mirror_upper(s, 1.0);
      symm_variants(s, +1, +1);
      s.ari[4] -= s.ali[0];
      for (int v = 0; v < 3; ++v)
        for (int j = 0; j < 4; ++j) {
          s.arb1[v][j][4] -= s.alb1[v][j][0];
          s.arb2[v][j][4] -= s.alb2[v][j][0];
        }
      break;
    }
    default:
      break;
  }
  return s;
}

}  // namespace pyr
