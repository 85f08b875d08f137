// Per-thread arithmetic of the line-operator kernels, written once for the
// device and for the host emulation used by CPU tests.
//
// A thread owns one chunk (rows s..s+m-1, m <= CHUNK) of one line. Phases:
//   A  RHS of its rows from the register window, local banded solve, and the
//      four edge values handed to the reduced system;
//   B  the four neighbour values from G . (edge values of all chunks), and the
//      spike correction; for the refined mode also the FP64 residual
//      r = Bv - A y32 and the local solve of the correction;
//   C  (refined mode) neighbour values of the correction and its spikes.
// All array indices are compile-time after unrolling, so the chunk state lives
// in VGPRs; chunk-uniform quantities (s, m, table rows) are scalar in the y/z
// kernels.
#pragma once

#include "pyr_plan.hpp"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PYR_HD __host__ __device__ __forceinline__
#else
#define PYR_HD inline
#endif

namespace pyr {

// MODE_F64: FP64 fields and solve. MODE_R32: FP32 fields, FP32 solve + one
// FP64-residual refinement. MODE_S32: FP32 fields, FP64 solve.
enum Mode { MODE_F64 = 0, MODE_R32 = 1, MODE_S32 = 2 };
enum Epi { EPI_STORE = 0, EPI_ADD = 1, EPI_ABS = 2, EPI_MAXABS = 3 };

// Everything a kernel needs about one operator/direction, passed by value.
struct LineArgs {
  int N, P, nlines, periodic;
  int nbytes;              // bytes of one field (buffer range of the y/z kernels)
  long long A, B, stride;  // y/z: line L starts at (L % A) + B*(L / A), rows `stride` apart
  int lx;                  // x kernel: lines per workgroup
  int addv, epi;
  double scale, epi_c;
  double w[9];             // interior RHS weights at offsets -R..R
  const int* cstart;
  const double* lu;
  const double* rc;
  const double* gsel;
  const double* band;
  const float* lu32;
  const float* rc32;
  const float* gsel32;
  // the same tables as contiguous blocks, staged into LDS by the kernels:
  // tab64 = lu|rc|gsel (doubles), tab32 = lu32|rc32|gsel32 (floats)
  const double* tab64;
  const float* tab32;
  int ntab;                // entries of one block: 9N + 16P^2
  int gtab;                // 1: the solve reads the tables from global memory
                           //    (too large for LDS: long lines, grouped launches)
  const double* bcoef;
  const short* bidx;
  const int* bnt;
  // x kernel: row pitch of the LDS tile, and whether rows move as 16-byte vectors
  int pitch, vec;
  // circulant solve (periodic lines): A = c Q^T Q, Q = I + p1 S + p2 S^2
  double p1, p2, invc;
  float p1f, p2f, invcf;
  double bd[3];            // LHS row A[i][i+-2], A[i][i+-1], A[i][i]
  const double* kf;        // P x 2 x 2P carry rows, forward and backward
  const double* kb;
  const float* kf32;
  const float* kb32;
  // FP64 sweeps with the store epilogue: when set, every stored value is also
  // written here rounded to FP32 (same element index as the output); with
  // only32 the FP64 result itself is not stored
  float* out32;
  int only32;
};

template <int KIND> struct Halo { static constexpr int R = (KIND == K_R4) ? 4 : 3; };

// --- RHS ----------------------------------------------------------------
// win[w] holds v[s - R + w]; rows k >= m get 0.
template <int KIND, typename TW>
PYR_HD void rhs_interior(const LineArgs& a, const TW* win, double* rhs, int m) {
  constexpr int R = Halo<KIND>::R;
#pragma unroll
  for (int k = 0; k < CHUNK; ++k) {
    double acc;
    if (KIND == K_D1) {
      // antisymmetric pairs (compact_d1.f90:156)
      acc = a.w[4] * ((double)win[k + R + 1] - (double)win[k + R - 1]) +
            a.w[5] * ((double)win[k + R + 2] - (double)win[k + R - 2]) +
            a.w[6] * ((double)win[k + R + 3] - (double)win[k + R - 3]);
    } else {
      // difference form sum_l w_l (v_{i+l} - v_i) (compact_r3.f90:124-132,
      // compact_r4.f90:144-156)
      const double vc = (double)win[k + R];
      acc = 0.0;
#pragma unroll
      for (int l = -R; l <= R; ++l)
        if (l != 0) acc += ((double)win[k + R + l] - vc) * a.w[l + R];
    }
    rhs[k] = (k < m) ? acc : 0.0;
  }
}

// One boundary row from its term list; rd(i) returns v[i] of this line.
template <typename Reader>
PYR_HD double boundary_row(const LineArgs& a, int slot, Reader rd) {
  double s = 0.0;
  const int nt = a.bnt[slot];
  for (int t = 0; t < nt; ++t) {
    const double c = a.bcoef[slot * MAXT + t];
    const int ia = a.bidx[(slot * MAXT + t) * 2 + 0];
    const int ib = a.bidx[(slot * MAXT + t) * 2 + 1];
    const double va = ia >= 0 ? (double)rd(ia) : 0.0;
    const double vb = ib >= 0 ? (double)rd(ib) : 0.0;
    s += c * (va - vb);
  }
  return s;
}

// Replace the rows of a bounded line's ends by their own RHS forms.
template <typename Reader>
PYR_HD void rhs_boundary(const LineArgs& a, int s, int m, double* rhs, Reader rd) {
  if (a.periodic) return;
  if (s == 0) {
#pragma unroll
    for (int k = 0; k < NB; ++k) rhs[k] = boundary_row(a, k, rd);
  }
  if (s + m > a.N - NB) {
    double g[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) g[j] = boundary_row(a, NB + j, rd);
#pragma unroll
    for (int k = 0; k < CHUNK; ++k) {
      const int j = s + k - (a.N - NB);
      if (k < m && j >= 0) rhs[k] = (j == 0) ? g[0] : (j == 1) ? g[1] : (j == 2) ? g[2] : g[3];
    }
  }
}

// --- chunk solves ---------------------------------------------------------
template <typename C, typename T>
PYR_HD void local_solve(const T* lu, int s, int m, C* x) {
#pragma unroll
  for (int k = 0; k < CHUNK; ++k) {
    if (k < m) {
      const T* c = lu + 5 * (s + k);
      C v = x[k];
      if (k >= 1) v -= (C)c[1] * x[k - 1];
      if (k >= 2) v -= (C)c[0] * x[k - 2];
      x[k] = v;
    }
  }
#pragma unroll
  for (int k = CHUNK - 1; k >= 0; --k) {
    if (k < m) {
      const T* c = lu + 5 * (s + k);
      C v = x[k];
      if (k + 1 < CHUNK) v -= (C)c[3] * x[k + 1];
      if (k + 2 < CHUNK) v -= (C)c[4] * x[k + 2];
      x[k] = v * (C)c[2];
    }
  }
}

template <typename C>
PYR_HD void edges(const C* x, int m, C* e4) {
  e4[0] = x[0];
  e4[1] = x[1];
  C a = 0, b = 0;
#pragma unroll
  for (int k = 2; k < CHUNK; ++k) {
    if (k == m - 2) a = x[k];
    if (k == m - 1) b = x[k];
  }
  e4[2] = a;
  e4[3] = b;
}

// z[q] = sum_c G[row_q(p)][c] * Y[c]; yr(c) returns the edge value c of this line
template <typename C, typename T, typename YR>
PYR_HD void neighbours(const T* gsel, int p, int P, YR yr, C* z) {
  const T* g = gsel + (size_t)p * 16 * P;
  const int n4 = 4 * P;
  C z0 = 0, z1 = 0, z2 = 0, z3 = 0;
  for (int c = 0; c < n4; ++c) {
    const C y = yr(c);
    z0 += (C)g[c] * y;
    z1 += (C)g[n4 + c] * y;
    z2 += (C)g[2 * n4 + c] * y;
    z3 += (C)g[3 * n4 + c] * y;
  }
  z[0] = z0; z[1] = z1; z[2] = z2; z[3] = z3;
}

template <typename C, typename T>
PYR_HD void spike_correct(const T* rc, int s, int m, const C* z, C* x) {
#pragma unroll
  for (int k = 0; k < CHUNK; ++k) {
    if (k < m) {
      const T* r = rc + 4 * (s + k);
      x[k] -= (C)r[0] * z[0] + (C)r[1] * z[1] + (C)r[2] * z[2] + (C)r[3] * z[3];
    }
  }
}

// --- circulant solve (periodic lines) -------------------------------------
// Forward recurrence x_k = r_k - p1 x_{k-1} - p2 x_{k-2} over rows k < m, in
// place, entered with (x_{-2}, x_{-1}) = (a0, a1).
template <typename C>
PYR_HD void circ_fwd(C* x, int m, C p1, C p2, C a0, C a1) {
#pragma unroll
  for (int k = 0; k < CHUNK; ++k) {
    if (k < m) {
      const C v = x[k] - p1 * a1 - p2 * a0;
      x[k] = v;
      a0 = a1;
      a1 = v;
    }
  }
}

// Exit state (x_{m-2}, x_{m-1}) of the forward recurrence entered with zeros.
template <typename C>
PYR_HD void circ_fwd_exit(const C* x, int m, C p1, C p2, C& e0, C& e1) {
  C a0 = 0, a1 = 0;
#pragma unroll
  for (int k = 0; k < CHUNK; ++k) {
    if (k < m) {
      const C v = x[k] - p1 * a1 - p2 * a0;
      a0 = a1;
      a1 = v;
    }
  }
  e0 = a0;
  e1 = a1;
}

// Backward recurrence u_k = w_k - p1 u_{k+1} - p2 u_{k+2}, in place, entered
// with (u_m, u_{m+1}) = (b1, b2).
template <typename C>
PYR_HD void circ_bwd(C* x, int m, C p1, C p2, C b1, C b2) {
#pragma unroll
  for (int k = CHUNK - 1; k >= 0; --k) {
    if (k < m) {
      const C v = x[k] - p1 * b1 - p2 * b2;
      x[k] = v;
      b2 = b1;
      b1 = v;
    }
  }
}

// Final backward recurrence: each finished row goes to sink(k, u_k) instead
// of back into x, so no result stays live in registers.
template <typename C, typename Sink>
PYR_HD void circ_bwd_sink(const C* x, int m, C p1, C p2, C b1, C b2, Sink sink) {
#pragma unroll
  for (int k = CHUNK - 1; k >= 0; --k) {
    if (k < m) {
      const C v = x[k] - p1 * b1 - p2 * b2;
      b2 = b1;
      b1 = v;
      sink(k, v);
    }
  }
}

// Exit state (u_0, u_1) of the backward recurrence entered with zeros.
template <typename C>
PYR_HD void circ_bwd_exit(const C* x, int m, C p1, C p2, C& e0, C& e1) {
  C b1 = 0, b2 = 0;
#pragma unroll
  for (int k = CHUNK - 1; k >= 0; --k) {
    if (k < m) {
      const C v = x[k] - p1 * b1 - p2 * b2;
      b2 = b1;
      b1 = v;
    }
  }
  e0 = b1;
  e1 = b2;
}

// Entry state of chunk p: rows 2p, 2p+1 of K times the exits yr(c), c < 2P.
template <typename C, typename T, typename YR>
PYR_HD void circ_carry(const T* K, int p, int P, YR yr, C& s0, C& s1) {
  const int n2 = 2 * P;
  const T* k0 = K + (size_t)(2 * p) * n2;
  const T* k1 = k0 + n2;
  C a = 0, b = 0;
  for (int c = 0; c < n2; ++c) {
    const C y = yr(c);
    a += (C)k0[c] * y;
    b += (C)k1[c] * y;
  }
  s0 = a;
  s1 = b;
}

// FP64 residual r = rhs - A y for rows of the chunk; y outside the chunk comes
// from the neighbour values z = (y[s-2], y[s-1], y[e], y[e+1]). row(k) gives
// the five band entries A[s+k][s+k-2 .. s+k+2].
template <typename Row>
PYR_HD void residual64(Row row, int m, const double* rhs, const float* y, const float* z,
                       double* res) {
#pragma unroll
  for (int k = 0; k < CHUNK; ++k) {
    double yy[5];
#pragma unroll
    for (int o = -2; o <= 2; ++o) {
      const int kk = k + o;
      double v;
      if (kk < 0) {
        v = (kk == -2) ? (double)z[0] : (double)z[1];
      } else if (kk < CHUNK) {
        v = (kk < m) ? (double)y[kk] : ((kk == m) ? (double)z[2] : (double)z[3]);
      } else {
        v = (kk == m) ? (double)z[2] : (double)z[3];
      }
      yy[o + 2] = v;
    }
    double r = 0.0;
    if (k < m) {
      const double* b = row(k);
      r = rhs[k] - (b[0] * yy[0] + b[1] * yy[1] + b[2] * yy[2] + b[3] * yy[3] + b[4] * yy[4]);
    }
    res[k] = r;
  }
}

template <typename TS>
PYR_HD TS epilogue(int epi, double c, double val, TS old) {
  switch (epi) {
    case EPI_ADD: return (TS)((double)old + val);
    case EPI_ABS: return (TS)(fabs(val) * c);
    case EPI_MAXABS: {
      const double t = fabs(val) * c;
      const double o = (double)old;
      return (TS)(o > t ? o : t);
    }
    default: return (TS)val;
  }
}

}  // namespace pyr
