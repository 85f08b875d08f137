// Host-side plan of one compact operator along one direction.
//
// The reference assembles a global band per direction (compact_basetype.f90:
// 124-148) and factors it whole: a bordered cyclic LU for periodic lines
// (pentadiagonal.f90:64-131) or a plain banded LU (:18-39). On the GPU each
// line is instead split into P chunks of at most CHUNK rows; a thread keeps its
// chunk in registers and solves it with the chunk's own banded LU. Chunks are
// coupled only through the two rows at each end, so the exact solution is
//     x = x_loc - rc(:,0..3) . (y[s-2], y[s-1], y[e], y[e+1])
// where the four neighbour values come from a 4P x 4P reduced system whose
// inverse G is identical for every line and is precomputed here. The same
// construction handles the cyclic coupling of periodic lines (chunk 0's left
// neighbour is chunk P-1), so no Sherman-Morrison step is needed.
//
// Periodic lines have a circulant LHS, which factors exactly as
//     A = c Q^T Q,   Q = I + p1 S + p2 S^2   ((S x)_i = x_{i-1}, cyclic)
// with both roots of 1 + p1 z + p2 z^2 outside the unit circle. A solve is then
// a forward and a backward second-order recurrence with constant coefficients;
// a chunk runs each recurrence from a zero state to get its exit state, the
// true entry states of all chunks follow from a 2P x 2P carry system (rows
// `kf`, `kb`), and the chunk re-runs the recurrence from its true entry. No
// per-row tables are read.
//
// RHS rows: interior rows use the interior weights; the up-to-4 rows at each
// end of a bounded line are stored as term lists coef*(v[a]-v[b]) that keep
// the reference's difference forms (compact_d1.f90:113-207,
// compact_r3.f90:97-157, compact_r4.f90:106-207).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "pyr_weights.hpp"

namespace pyr {

constexpr int CHUNK = 32;   // rows per thread (register-resident)
constexpr int PMAX = 32;    // chunks per line -> N <= 1024
constexpr int NB = 4;       // boundary rows per side with their own RHS forms
constexpr int MAXT = 10;    // terms per boundary row

enum Kind { K_D1 = 0, K_R3 = 1, K_R4 = 2 };

inline int op_kind(int op) { return op == OP_D1 ? K_D1 : (op == OP_D2 ? K_R3 : K_R4); }
inline int kind_halo(int kind) { return kind == K_R4 ? 4 : 3; }

struct OpPlan {
  int op = 0, kind = 0, N = 0, P = 0, periodic = 0, implicit = 0, addv = 0;
  int null_op = 0, null_option = 0, R = 3;
  double scale = 1.0;       // metric applied to the solve result
  double w[9] = {};         // interior RHS weights, offsets -R..R
  std::vector<int> cstart;  // P+1 chunk starts
  std::vector<double> lu;   // N x 5: l(i-2), l(i-1), 1/u(i,i), u(i,i+1), u(i,i+2)
  std::vector<double> rc;   // N x 4: spikes for y[s-2], y[s-1], y[e], y[e+1]
  std::vector<double> gsel; // P x 4 x 4P: rows of G giving those four values
  std::vector<double> band; // N x 5: full LHS band (refinement residual)
  std::vector<double> bcoef;  // 2*NB x MAXT
  std::vector<short> bidx;    // 2*NB x MAXT x 2 (a,b); -1 reads as 0
  std::vector<int> bnt;       // 2*NB
  // circulant solve of periodic implicit lines
  int circ = 0;
  double p1 = 0, p2 = 0, invc = 0;  // A = c Q^T Q, Q = I + p1 S + p2 S^2
  double bd[3] = {};                // LHS row: A[i][i+-2], A[i][i+-1], A[i][i]
  std::vector<double> kf, kb;       // P x 2 x 2P: entry states from all exit states
};

namespace detail {

using ld = long double;

struct TermRow {
  std::vector<double> c;
  std::vector<int> a, b;
  void add(double coef, int ia, int ib) {
    if (coef == 0.0) return;
    c.push_back(coef); a.push_back(ia); b.push_back(ib);
  }
};

// Boundary-row RHS forms of a bounded line, 0-based rows; ar(l,i) is the
// reference's (1-based l, 1-based i) weight. Ghost columns read 0.
inline void boundary_terms(int kind, int n, int v1, int v2,
                           const std::vector<std::vector<double>>& ar,
                           TermRow lo[NB], TermRow hi[NB]) {
  auto AR = [&](int l, int i1) { return ar[i1 - 1][l - 1]; };
  auto sumr = [&](int i1, int nr) { double s = 0; for (int l = 1; l <= nr; ++l) s += AR(l, i1); return s; };
  if (kind == K_D1) {
    if (v1 == -1) {
      for (int l = 4; l <= 7; ++l) lo[0].add(AR(l, 1), l - 4, -1);
      for (int l = 3; l <= 7; ++l) lo[1].add(AR(l, 2), l - 3, -1);
      for (int l = 2; l <= 7; ++l) lo[2].add(AR(l, 3), l - 2, -1);
    } else {
      for (int l = 5; l <= 7; ++l) lo[0].add(AR(l, 1), l - 4, 0);
      for (int l = 4; l <= 7; ++l) lo[1].add(AR(l, 2), l - 3, 0);
      lo[2].add(AR(5, 3), 3, 1); lo[2].add(AR(6, 3), 4, 0); lo[2].add(AR(7, 3), 5, 0);
    }
    // row 4 (1-based): interior antisymmetric form with its own weights
    lo[3].add(AR(5, 4), 4, 2); lo[3].add(AR(6, 4), 5, 1); lo[3].add(AR(7, 4), 6, 0);
    // row n-3 (1-based)
    hi[0].add(AR(5, n - 3), n - 3, n - 5); hi[0].add(AR(6, n - 3), n - 2, n - 6);
    hi[0].add(AR(7, n - 3), n - 1, n - 7);
    if (v2 == -1) {
      for (int l = 1; l <= 6; ++l) hi[1].add(AR(l, n - 2), n - 7 + l, -1);
      for (int l = 1; l <= 5; ++l) hi[2].add(AR(l, n - 1), n - 6 + l, -1);
      for (int l = 1; l <= 4; ++l) hi[3].add(AR(l, n), n - 5 + l, -1);
    } else {
      hi[1].add(AR(1, n - 2), n - 6, n - 1); hi[1].add(AR(2, n - 2), n - 5, n - 1);
      hi[1].add(AR(3, n - 2), n - 4, n - 2);
      for (int l = 1; l <= 4; ++l) hi[2].add(AR(l, n - 1), n - 6 + l, n - 1);
      for (int l = 1; l <= 3; ++l) hi[3].add(AR(l, n), n - 5 + l, n - 1);
    }
  } else if (kind == K_R3) {
    // ghosts are zero on a bounded end (compact_r3.f90:84-99)
    for (int l = 1; l <= 3; ++l) lo[0].add(AR(l, 1), -1, 0);
    for (int l = 4; l <= 7; ++l) lo[0].add(AR(l, 1), l - 4, 0);
    for (int l = 1; l <= 2; ++l) lo[1].add(AR(l, 2), -1, 1);
    for (int l = 3; l <= 7; ++l) lo[1].add(AR(l, 2), l - 3, 1);
    lo[2].add(AR(1, 3), -1, 2);
    for (int l = 2; l <= 7; ++l) lo[2].add(AR(l, 3), l - 2, 2);
    for (int l = 1; l <= 7; ++l) lo[3].add(AR(l, 4), l - 1, 3);
    if (v1 == -1)
      for (int r = 0; r < 3; ++r) lo[r].add(sumr(r + 1, 7), r, -1);
    for (int l = 1; l <= 7; ++l) hi[0].add(AR(l, n - 3), n - 8 + l, n - 4);
    for (int l = 1; l <= 6; ++l) hi[1].add(AR(l, n - 2), n - 7 + l, n - 3);
    hi[1].add(AR(7, n - 2), -1, n - 3);
    for (int l = 1; l <= 5; ++l) hi[2].add(AR(l, n - 1), n - 6 + l, n - 2);
    for (int l = 6; l <= 7; ++l) hi[2].add(AR(l, n - 1), -1, n - 2);
    for (int l = 1; l <= 4; ++l) hi[3].add(AR(l, n), n - 5 + l, n - 1);
    for (int l = 5; l <= 7; ++l) hi[3].add(AR(l, n), -1, n - 1);
    if (v2 == -1)
      for (int r = 1; r < 4; ++r) hi[r].add(sumr(n - 3 + r, 7), n - 4 + r, -1);
  } else {
    if (v1 == -1) {
      for (int l = 5; l <= 9; ++l) lo[0].add(AR(l, 1), l - 5, -1);
      for (int l = 4; l <= 9; ++l) lo[1].add(AR(l, 2), l - 4, -1);
      for (int l = 3; l <= 9; ++l) lo[2].add(AR(l, 3), l - 3, -1);
      for (int l = 2; l <= 9; ++l) lo[3].add(AR(l, 4), l - 2, -1);
    } else {
      for (int l = 6; l <= 9; ++l) lo[0].add(AR(l, 1), l - 5, 0);
      for (int l = 5; l <= 9; ++l) lo[1].add(AR(l, 2), l - 4, 0);
      for (int l = 4; l <= 9; ++l) lo[2].add(AR(l, 3), l - 3, 0);
      for (int l = 3; l <= 9; ++l) lo[3].add(AR(l, 4), l - 2, 0);
    }
    if (v2 == -1) {
      for (int l = 1; l <= 8; ++l) hi[0].add(AR(l, n - 3), n - 9 + l, -1);
      for (int l = 1; l <= 7; ++l) hi[1].add(AR(l, n - 2), n - 8 + l, -1);
      for (int l = 1; l <= 6; ++l) hi[2].add(AR(l, n - 1), n - 7 + l, -1);
      for (int l = 1; l <= 5; ++l) hi[3].add(AR(l, n), n - 6 + l, -1);
    } else {
      for (int l = 1; l <= 7; ++l) hi[0].add(AR(l, n - 3), n - 9 + l, n - 1);
      for (int l = 1; l <= 6; ++l) hi[1].add(AR(l, n - 2), n - 8 + l, n - 1);
      for (int l = 1; l <= 5; ++l) hi[2].add(AR(l, n - 1), n - 7 + l, n - 1);
      for (int l = 1; l <= 4; ++l) hi[3].add(AR(l, n), n - 6 + l, n - 1);
    }
  }
}

// Gauss-Jordan inverse with partial pivoting (small dense, long double)
inline void invert(std::vector<ld>& a, int n) {
  std::vector<ld> inv(size_t(n) * n, 0.0L);
  for (int i = 0; i < n; ++i) inv[size_t(i) * n + i] = 1.0L;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[size_t(r) * n + c]) > std::fabs(a[size_t(piv) * n + c])) piv = r;
    if (a[size_t(piv) * n + c] == 0.0L) throw std::runtime_error("singular reduced system");
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[size_t(c) * n + k], a[size_t(piv) * n + k]);
        std::swap(inv[size_t(c) * n + k], inv[size_t(piv) * n + k]);
      }
    const ld d = 1.0L / a[size_t(c) * n + c];
    for (int k = 0; k < n; ++k) { a[size_t(c) * n + k] *= d; inv[size_t(c) * n + k] *= d; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const ld f = a[size_t(r) * n + c];
      if (f == 0.0L) continue;
      for (int k = 0; k < n; ++k) {
        a[size_t(r) * n + k] -= f * a[size_t(c) * n + k];
        inv[size_t(r) * n + k] -= f * inv[size_t(c) * n + k];
      }
    }
  }
  a.swap(inv);
}

// Spectral factor of the symmetric pentadiagonal symbol
//   b/z^2 + a/z + d + a z + b z^2 = c (1 + p1 z + p2 z^2)(1 + p1/z + p2/z^2)
// with the roots of 1 + p1 z + p2 z^2 outside the unit circle. With w = z + 1/z
// the quartic becomes b w^2 + a w + (d - 2b) = 0; each w gives the pair
// z, 1/z, of which the root inside the unit circle is kept. False if the
// symbol vanishes on the unit circle or b = 0.
inline bool spectral_factor(ld b, ld a, ld d, ld& p1, ld& p2, ld& c) {
  using cx = std::complex<ld>;
  if (b == 0.0L) return false;
  const cx disc = std::sqrt(cx(a * a - 4.0L * b * (d - 2.0L * b), 0.0L));
  const cx w[2] = {(cx(-a, 0.0L) + disc) / (2.0L * b), (cx(-a, 0.0L) - disc) / (2.0L * b)};
  cx r[2];
  for (int i = 0; i < 2; ++i) {
    const cx sq = std::sqrt(w[i] * w[i] - cx(4.0L, 0.0L));
    const cx z1 = (w[i] + sq) / 2.0L, z2 = (w[i] - sq) / 2.0L;
    r[i] = std::abs(z1) < std::abs(z2) ? z1 : z2;
    if (!(std::abs(r[i]) < 1.0L - 1e-9L)) return false;
  }
  const cx s1 = -(r[0] + r[1]), s2 = r[0] * r[1];
  if (std::fabs(s1.imag()) > 1e-15L || std::fabs(s2.imag()) > 1e-15L) return false;
  p1 = s1.real();
  p2 = s2.real();
  c = b / p2;
  const ld ed = c * (1.0L + p1 * p1 + p2 * p2) - d, ea = c * p1 * (1.0L + p2) - a;
  const ld scale = std::fabs(d) + 2.0L * std::fabs(a) + 2.0L * std::fabs(b);
  return std::fabs(ed) <= 1e-14L * scale && std::fabs(ea) <= 1e-14L * scale;
}

// Carry rows of the circulant solve. Forward: chunk q entered with
// (x[s_q-2], x[s_q-1]) leaves with its zero-entry exit e_q plus H_q times the
// entry, and chunk q+1 (cyclic) is entered with that exit. Backward likewise
// from the right with (u[e_q], u[e_q+1]). K maps the exits of all chunks
// (layout 2q + i) to the entry of each chunk (rows 2p + i).
inline void circ_carry_rows(int P, const std::vector<int>& cstart, ld p1, ld p2,
                            std::vector<double>& kf, std::vector<double>& kb) {
  const int n2 = 2 * P;
  std::vector<ld> Mf(size_t(n2) * n2, 0.0L), Mb(size_t(n2) * n2, 0.0L);
  for (int q = 0; q < P; ++q) {
    const int m = cstart[q + 1] - cstart[q];
    ld Hf[2][2], Hb[2][2];
    for (int j = 0; j < 2; ++j) {
      ld xm2 = j == 0, xm1 = j == 1;  // forward state (x[k-2], x[k-1])
      for (int k = 0; k < m; ++k) {
        const ld x = -p1 * xm1 - p2 * xm2;
        xm2 = xm1;
        xm1 = x;
      }
      Hf[0][j] = xm2;
      Hf[1][j] = xm1;
      ld u1 = j == 0, u2 = j == 1;  // backward state (u[k+1], u[k+2])
      for (int k = m - 1; k >= 0; --k) {
        const ld u = -p1 * u1 - p2 * u2;
        u2 = u1;
        u1 = u;
      }
      Hb[0][j] = u1;
      Hb[1][j] = u2;
    }
    const int qn = (q + 1) % P, qp = (q - 1 + P) % P;
    for (int i = 0; i < 2; ++i) {
      Mf[size_t(2 * qn + i) * n2 + 2 * qn + i] += 1.0L;
      Mb[size_t(2 * qp + i) * n2 + 2 * qp + i] += 1.0L;
      for (int j = 0; j < 2; ++j) {
        Mf[size_t(2 * qn + i) * n2 + 2 * q + j] -= Hf[i][j];
        Mb[size_t(2 * qp + i) * n2 + 2 * q + j] -= Hb[i][j];
      }
    }
  }
  invert(Mf, n2);
  invert(Mb, n2);
  // the exit of chunk q is the right-hand side of row block q+1 (forward) or
  // q-1 (backward)
  kf.assign(size_t(n2) * n2, 0.0);
  kb.assign(size_t(n2) * n2, 0.0);
  for (int r = 0; r < n2; ++r)
    for (int q = 0; q < P; ++q)
      for (int i = 0; i < 2; ++i) {
        kf[size_t(r) * n2 + 2 * q + i] = (double)Mf[size_t(r) * n2 + 2 * ((q + 1) % P) + i];
        kb[size_t(r) * n2 + 2 * q + i] = (double)Mb[size_t(r) * n2 + 2 * ((q - 1 + P) % P) + i];
      }
}

}  // namespace detail

// Build the plan. v1/v2 are the boundary variants (0 NONE, +1 SYMM).
inline OpPlan build_plan(int op, int n, int periodic, int v1, int v2, double metric) {
  using detail::ld;
  OpPlan pl;
  const Scheme sc = make_scheme(op);
  pl.op = op;
  pl.kind = op_kind(op);
  pl.N = n;
  pl.periodic = periodic;
  pl.implicit = sc.implicit;
  pl.addv = (sc.null_option == 1);
  pl.null_option = sc.null_option;
  pl.R = kind_halo(pl.kind);
  pl.null_op = (n < 4);  // compact_basetype.f90:86-90
  pl.scale = (op == OP_D1) ? 1.0 / metric : (op == OP_D2 ? 1.0 / (metric * metric) : 1.0);
  const int hr = sc.ncr / 2;
  for (int l = -pl.R; l <= pl.R; ++l) pl.w[l + pl.R] = sc.ari[l + hr];
  if (pl.null_op) return pl;
  // the reference's boundary forms read up to row 8 of a line (compact_r4.f90:
  // 117-120); shorter lines are outside what the schemes define
  if (n < 8) throw std::runtime_error("line length " + std::to_string(n) + " is below 8 (supported: 1-3 or 8-1024)");
  if (n > CHUNK * PMAX)
    throw std::runtime_error("line length " + std::to_string(n) + " exceeds " + std::to_string(CHUNK * PMAX));

  // assembled rows (compact_basetype.f90:124-133), 0-based
  std::vector<std::vector<double>> gal(n, std::vector<double>(5, 0.0));
  std::vector<std::vector<double>> gar(n, std::vector<double>(9, 0.0));
  for (int i = 0; i < n; ++i) {
    for (int c = 0; c < sc.ncl; ++c) gal[i][c] = sc.ali[c];
    for (int l = 0; l < sc.ncr; ++l) gar[i][l] = sc.ari[l];
  }
  if (!periodic) {
    for (int j = 0; j < 4; ++j) {
      for (int c = 0; c < sc.ncl; ++c) {
        gal[j][c] = sc.alb1[v1 + 1][j][c];
        gal[n - 4 + j][c] = sc.alb2[v2 + 1][j][c];
      }
      for (int l = 0; l < sc.ncr; ++l) {
        gar[j][l] = sc.arb1[v1 + 1][j][l];
        gar[n - 4 + j][l] = sc.arb2[v2 + 1][j][l];
      }
    }
  }
  // full LHS band as A[i][i-2..i+2] (explicit operators: identity)
  pl.band.assign(size_t(n) * 5, 0.0);
  for (int i = 0; i < n; ++i) {
    if (sc.implicit)
      for (int c = 0; c < 5; ++c) pl.band[size_t(i) * 5 + c] = gal[i][c];
    else
      pl.band[size_t(i) * 5 + 2] = 1.0;
  }
  // boundary RHS term lists
  pl.bcoef.assign(2 * NB * MAXT, 0.0);
  pl.bidx.assign(2 * NB * MAXT * 2, -1);
  pl.bnt.assign(2 * NB, 0);
  if (!periodic) {
    detail::TermRow lo[NB], hi[NB];
    detail::boundary_terms(pl.kind, n, v1, v2, gar, lo, hi);
    for (int r = 0; r < 2 * NB; ++r) {
      const detail::TermRow& t = r < NB ? lo[r] : hi[r - NB];
      if ((int)t.c.size() > MAXT) throw std::runtime_error("boundary row has too many terms");
      pl.bnt[r] = (int)t.c.size();
      for (size_t q = 0; q < t.c.size(); ++q)
        if (t.a[q] < -1 || t.a[q] >= n || t.b[q] < -1 || t.b[q] >= n)
          throw std::runtime_error("boundary term outside the line");
      for (size_t q = 0; q < t.c.size(); ++q) {
        pl.bcoef[r * MAXT + q] = t.c[q];
        pl.bidx[(r * MAXT + q) * 2 + 0] = (short)t.a[q];
        pl.bidx[(r * MAXT + q) * 2 + 1] = (short)t.b[q];
      }
    }
  }

  // chunk partition: P = ceil(n/CHUNK), balanced lengths (all >= 4)
  const int P = (n + CHUNK - 1) / CHUNK;
  pl.P = P;
  pl.cstart.resize(P + 1);
  {
    const int base = n / P, extra = n % P;
    int s = 0;
    for (int p = 0; p < P; ++p) { pl.cstart[p] = s; s += base + (p < extra ? 1 : 0); }
    pl.cstart[P] = n;
  }
  pl.lu.assign(size_t(n) * 5, 0.0);
  pl.rc.assign(size_t(n) * 4, 0.0);
  pl.gsel.assign(size_t(P) * 4 * 4 * P, 0.0);
  if (!sc.implicit) return pl;

  if (periodic && sc.ncl == 5 && sc.ali[0] == sc.ali[4] && sc.ali[1] == sc.ali[3]) {
    ld p1, p2, cc;
    if (detail::spectral_factor(sc.ali[0], sc.ali[1], sc.ali[2], p1, p2, cc)) {
      pl.circ = 1;
      pl.p1 = (double)p1;
      pl.p2 = (double)p2;
      pl.invc = (double)(1.0L / cc);
      pl.bd[0] = sc.ali[0];
      pl.bd[1] = sc.ali[1];
      pl.bd[2] = sc.ali[2];
      detail::circ_carry_rows(P, pl.cstart, p1, p2, pl.kf, pl.kb);
    }
  }

  auto Aent = [&](int i, int col) -> double {  // A[i][col], col may wrap
    const int off = col - i;
    if (periodic) {
      int d = col - i;
      // bring the offset into [-2,2] modulo n
      if (d > 2) d -= n;
      if (d < -2) d += n;
      if (d < -2 || d > 2) return 0.0;
      return gal[i][d + 2];
    }
    if (off < -2 || off > 2 || col < 0 || col >= n) return 0.0;
    return gal[i][off + 2];
  };

  std::vector<ld> S(size_t(4 * P) * 4 * P, 0.0L);
  for (int p = 0; p < P; ++p) {
    const int s = pl.cstart[p], e = pl.cstart[p + 1], m = e - s;
    // dense block, no pivoting LU (SPD-like blocks of the compact LHS)
    std::vector<ld> blk(size_t(m) * m, 0.0L);
    for (int i = 0; i < m; ++i)
      for (int j = std::max(0, i - 2); j <= std::min(m - 1, i + 2); ++j)
        blk[size_t(i) * m + j] = (ld)gal[s + i][j - i + 2];
    std::vector<ld> L(size_t(m) * 3, 0.0L);  // multipliers for i-2, i-1
    for (int k = 0; k < m; ++k) {
      const ld piv = blk[size_t(k) * m + k];
      if (piv == 0.0L) throw std::runtime_error("zero pivot in chunk LU");
      for (int i = k + 1; i <= std::min(m - 1, k + 2); ++i) {
        const ld f = blk[size_t(i) * m + k] / piv;
        L[size_t(i) * 3 + (i - k == 2 ? 0 : 1)] = f;
        for (int j = k; j <= std::min(m - 1, k + 2); ++j) blk[size_t(i) * m + j] -= f * blk[size_t(k) * m + j];
      }
    }
    for (int i = 0; i < m; ++i) {
      double* c = &pl.lu[size_t(s + i) * 5];
      c[0] = (double)L[size_t(i) * 3 + 0];
      c[1] = (double)L[size_t(i) * 3 + 1];
      c[2] = (double)(1.0L / blk[size_t(i) * m + i]);
      c[3] = (i + 1 < m) ? (double)blk[size_t(i) * m + i + 1] : 0.0;
      c[4] = (i + 2 < m) ? (double)blk[size_t(i) * m + i + 2] : 0.0;
    }
    // solve with the long-double factors for the spike columns
    auto solve = [&](std::vector<ld>& r) {
      for (int i = 0; i < m; ++i) {
        if (i >= 1) r[i] -= L[size_t(i) * 3 + 1] * r[i - 1];
        if (i >= 2) r[i] -= L[size_t(i) * 3 + 0] * r[i - 2];
      }
      for (int i = m - 1; i >= 0; --i) {
        ld v = r[i];
        if (i + 1 < m) v -= blk[size_t(i) * m + i + 1] * r[i + 1];
        if (i + 2 < m) v -= blk[size_t(i) * m + i + 2] * r[i + 2];
        r[i] = v / blk[size_t(i) * m + i];
      }
    };
    std::vector<ld> w0(m, 0.0L), w1(m, 0.0L), wm2(m, 0.0L), wm1(m, 0.0L);
    w0[0] = 1; w1[1] = 1; wm2[m - 2] = 1; wm1[m - 1] = 1;
    solve(w0); solve(w1); solve(wm2); solve(wm1);
    const bool has_lo = periodic || p > 0, has_hi = periodic || p < P - 1;
    // couplings to the previous chunk's last two unknowns (y[s-2], y[s-1])
    const ld a_s_m2 = has_lo ? (ld)Aent(s, s - 2) : 0.0L;
    const ld a_s_m1 = has_lo ? (ld)Aent(s, s - 1) : 0.0L;
    const ld a_s1_m1 = has_lo ? (ld)Aent(s + 1, s - 1) : 0.0L;
    // couplings to the next chunk's first two unknowns (y[e], y[e+1])
    const ld a_e2_e = has_hi ? (ld)Aent(e - 2, e) : 0.0L;
    const ld a_e1_e = has_hi ? (ld)Aent(e - 1, e) : 0.0L;
    const ld a_e1_e1 = has_hi ? (ld)Aent(e - 1, e + 1) : 0.0L;
    std::vector<ld> r0(m), r1(m), r2(m), r3(m);
    for (int i = 0; i < m; ++i) {
      r0[i] = a_s_m2 * w0[i];
      r1[i] = a_s_m1 * w0[i] + a_s1_m1 * w1[i];
      r2[i] = a_e2_e * wm2[i] + a_e1_e * wm1[i];
      r3[i] = a_e1_e1 * wm1[i];
      double* rr = &pl.rc[size_t(s + i) * 4];
      rr[0] = (double)r0[i]; rr[1] = (double)r1[i]; rr[2] = (double)r2[i]; rr[3] = (double)r3[i];
    }
    // reduced-system rows for this chunk's four edge unknowns
    const int pm = (p - 1 + P) % P, pp = (p + 1) % P;
    const int edge[4] = {0, 1, m - 2, m - 1};
    for (int q = 0; q < 4; ++q) {
      const int row = 4 * p + q, li = edge[q];
      S[size_t(row) * 4 * P + row] += 1.0L;
      if (has_lo) {
        S[size_t(row) * 4 * P + 4 * pm + 2] += r0[li];
        S[size_t(row) * 4 * P + 4 * pm + 3] += r1[li];
      }
      if (has_hi) {
        S[size_t(row) * 4 * P + 4 * pp + 0] += r2[li];
        S[size_t(row) * 4 * P + 4 * pp + 1] += r3[li];
      }
    }
  }
  detail::invert(S, 4 * P);
  for (int p = 0; p < P; ++p) {
    const bool has_lo = periodic || p > 0, has_hi = periodic || p < P - 1;
    const int pm = (p - 1 + P) % P, pp = (p + 1) % P;
    const int rows[4] = {4 * pm + 2, 4 * pm + 3, 4 * pp + 0, 4 * pp + 1};
    const bool use[4] = {has_lo, has_lo, has_hi, has_hi};
    for (int q = 0; q < 4; ++q)
      for (int c = 0; c < 4 * P; ++c)
        pl.gsel[(size_t(p) * 4 + q) * 4 * P + c] = use[q] ? (double)S[size_t(rows[q]) * 4 * P + c] : 0.0;
  }
  return pl;
}

}  // namespace pyr
