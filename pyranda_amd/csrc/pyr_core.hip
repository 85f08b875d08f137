// MI355X (gfx950) kernels and C-ABI of the compact finite-difference core.
//
// One launch applies one compact operator along one direction to a whole
// (nx,ny,nz) field. Each thread owns one chunk of one line (pyr_plan.hpp):
//   y/z lines: lanes of a wave are adjacent lines, i.e. adjacent x, so every
//     row load/store of a wave is one contiguous 256-512 B segment;
//   x lines:   a workgroup stages LX contiguous lines in LDS with coalesced
//     copies, threads then read their chunk window from LDS.
// The reduced system couples the chunks of a line through LDS within the
// workgroup, so every field element is read once and written once from HBM.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pyranda_amd.h"
#include "pyr_line.hpp"

using namespace pyr;

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
// The file compiles as several translation units that build in parallel:
// PYR_PART 0 holds the host code and the C-ABI, parts 1-3 the line kernels
// of one solve mode each (FP64, FP32-refined, FP32 fields with FP64 solves),
// through explicit instantiations of the launch functions.
#ifndef PYR_PART
#define PYR_PART 0
#endif
#if PYR_PART == 0
static thread_local std::string g_err;
static int fail(const std::string& m) {
  g_err = m;
  return 1;
}
namespace pyr {
// errors of the other translation units (pyr_mlp.hip, the kernel parts) go
// to pyr_last_error too
int report_error(const std::string& m) { return fail(m); }
}  // namespace pyr
#else
namespace pyr {
int report_error(const std::string& m);
}  // namespace pyr
static int fail(const std::string& m) { return pyr::report_error(m); }
#endif
#define HIPCHECK(x)                                                              \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) return fail(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <int MODE> struct SolveT { using type = double; };
template <> struct SolveT<MODE_R32> { using type = float; };  // MODE_S32 solves in double

constexpr int WG_MAX = 512;

// Coefficient tables of one operator, as the solve reads them (LDS copies on
// the device, plan vectors in the host emulation).
struct Tabs {
  const double* lu = nullptr;
  const double* rc = nullptr;
  const double* gsel = nullptr;
  const float* lu32 = nullptr;
  const float* rc32 = nullptr;
  const float* gsel32 = nullptr;
  const double* band = nullptr;
};

__host__ __device__ inline size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }

// bytes of the staged tables: FP64 lu|rc|gsel, or FP32 lu|rc|gsel + FP64 band
template <int MODE>
__host__ __device__ inline size_t stage_bytes(int ntab, int N) {
  return MODE != MODE_R32 ? align16((size_t)ntab * sizeof(double))
                          : align16((size_t)ntab * sizeof(float)) + align16((size_t)5 * N * sizeof(double));
}

template <int MODE>
__device__ __forceinline__ Tabs stage_tables(const LineArgs& a, unsigned char* dst, int tid, int nthr) {
  Tabs t;
  if (MODE != MODE_R32) {
    double* d = reinterpret_cast<double*>(dst);
    for (int e = tid; e < a.ntab; e += nthr) d[e] = a.tab64[e];
    t.lu = d;
    t.rc = d + 5 * a.N;
    t.gsel = d + 9 * a.N;
  } else {
    float* f = reinterpret_cast<float*>(dst);
    for (int e = tid; e < a.ntab; e += nthr) f[e] = a.tab32[e];
    double* bd = reinterpret_cast<double*>(dst + align16((size_t)a.ntab * sizeof(float)));
    for (int e = tid; e < 5 * a.N; e += nthr) bd[e] = a.band[e];
    t.lu32 = f;
    t.rc32 = f + 5 * a.N;
    t.gsel32 = f + 9 * a.N;
    t.band = bd;
  }
  return t;
}

// The tables as the solve reads them from global memory (lines whose tables
// exceed the LDS budget, grouped launches).
template <int MODE>
__device__ __forceinline__ Tabs global_tables(const LineArgs& a) {
  Tabs t;
  t.lu = a.lu;
  t.rc = a.rc;
  t.gsel = a.gsel;
  t.lu32 = a.lu32;
  t.rc32 = a.rc32;
  t.gsel32 = a.gsel32;
  t.band = a.band;
  return t;
}

template <int MODE>
__device__ __forceinline__ Tabs get_tables(const LineArgs& a, unsigned char* dst, int tid, int nthr) {
  return a.gtab ? global_tables<MODE>(a) : stage_tables<MODE>(a, dst, tid, nthr);
}

// staged tables above this size are read from global memory instead
constexpr size_t STAGE_MAX_BYTES = 64 * 1024;

enum Solve { SOLVE_NONE = 0, SOLVE_LU = 1, SOLVE_CIRC = 2 };

// Solve one chunk given its RHS with the chunked LU; the LDS exchange of edge
// values goes through Y (4P x ly entries of the solve type), column `col`
// belongs to this line. The solution replaces the RHS in `rhs`.
// act: the thread's chunk exists (a grouped launch shapes its workgroups for
// the member with the most chunks; other threads only take the barriers).
template <int MODE, typename CS>
__device__ __forceinline__ void solve_chunk_lu(const LineArgs& a, const Tabs& t, int p, int s, int m,
                                               CS* Y, int ly, int col, double* rhs, bool act = true) {
  auto yr = [&](int c) { return Y[c * ly + col]; };
  if (MODE != MODE_R32) {
    double* x = rhs;  // solve in place
    double e4[4];
    if (act) {
      local_solve<double>(t.lu, s, m, x);
      edges(x, m, e4);
#pragma unroll
      for (int q = 0; q < 4; ++q) Y[(4 * p + q) * ly + col] = (CS)e4[q];
    }
    __syncthreads();
    if (act) {
      double z[4];
      neighbours<double>(t.gsel, p, a.P, yr, z);
      spike_correct(t.rc, s, m, z, x);
    }
  } else {
    float x[CHUNK];
    float z[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < CHUNK; ++k) x[k] = (float)rhs[k];
    if (act) {
      local_solve<float>(t.lu32, s, m, x);
      float e4[4];
      edges(x, m, e4);
#pragma unroll
      for (int q = 0; q < 4; ++q) Y[(4 * p + q) * ly + col] = (CS)e4[q];
    }
    __syncthreads();
    float d[CHUNK];
    if (act) {
      neighbours<float>(t.gsel32, p, a.P, yr, z);
      spike_correct(t.rc32, s, m, z, x);
      // one refinement step: FP64 residual against the same banded operator
      residual64([&](int k) { return t.band + 5 * (s + k); }, m, rhs, x, z, rhs);
#pragma unroll
      for (int k = 0; k < CHUNK; ++k) d[k] = (float)rhs[k];
      local_solve<float>(t.lu32, s, m, d);
    }
    __syncthreads();
    if (act) {
      float f4[4];
      edges(d, m, f4);
#pragma unroll
      for (int q = 0; q < 4; ++q) Y[(4 * p + q) * ly + col] = (CS)f4[q];
    }
    __syncthreads();
    if (act) {
      float zd[4];
      neighbours<float>(t.gsel32, p, a.P, yr, zd);
      spike_correct(t.rc32, s, m, zd, d);
#pragma unroll
      for (int k = 0; k < CHUNK; ++k) rhs[k] = (double)x[k] + (double)d[k];
    }
  }
}

// One circulant solve A y = r of the chunk (without the 1/c): forward
// recurrence with its carry through Xa, backward recurrence with its carry
// through Xb; the rows of the result go to sink(k, y_k). Two barriers. A
// thread whose chunk index p is not below P (a workgroup shaped for a longer
// line of another member of a grouped launch) only takes the barriers.
template <bool GRP, typename C, typename Sink>
__device__ __forceinline__ void circ_solve(const C* kf, const C* kb, C p1, C p2, int p, int P,
                                           int m, C* Xa, C* Xb, int ly, int col, C* x, Sink sink) {
  const bool act = !GRP || p < P;
  C e0 = 0, e1 = 0;
  if (act) {
    circ_fwd_exit(x, m, p1, p2, e0, e1);
    Xa[(2 * p) * ly + col] = e0;
    Xa[(2 * p + 1) * ly + col] = e1;
  }
  __syncthreads();
  if (act) {
    circ_carry(kf, p, P, [&](int c) { return Xa[c * ly + col]; }, e0, e1);
    circ_fwd(x, m, p1, p2, e0, e1);
    circ_bwd_exit(x, m, p1, p2, e0, e1);
    Xb[(2 * p) * ly + col] = e0;
    Xb[(2 * p + 1) * ly + col] = e1;
  }
  __syncthreads();
  if (act) {
    circ_carry(kb, p, P, [&](int c) { return Xb[c * ly + col]; }, e0, e1);
    circ_bwd_sink(x, m, p1, p2, e0, e1, sink);
  }
}

// Circulant solve of one chunk; X0 and X1 hold 4P x ly entries each. Every
// buffer is written only after a barrier that follows its last read. The
// solution rows go to emit(k, y_k) in FP64 (before the metric scale).
// carry rows of the circulant solve in the solve type: 0 forward, 1 backward
template <int MODE>
__device__ __forceinline__ const typename SolveT<MODE>::type* kmat(const LineArgs& a, int which) {
  if constexpr (MODE == MODE_R32) return which ? a.kb32 : a.kf32;
  else return which ? a.kb : a.kf;
}

// Entries per chunk of each of the two exchange buffers of solve_chunk_circ
// (per line): the forward and backward carries take 2; the refined FP32 solve
// also exchanges the 4 edge values of its first solution.
template <int MODE>
constexpr int circ_words() { return MODE == MODE_R32 ? 4 : 2; }

// kf/kb: the carry rows in the solve type (global memory or an LDS copy)
template <int MODE, typename CS, bool GRP = false, typename Emit>
__device__ __forceinline__ void solve_chunk_circ(const LineArgs& a, const CS* kf, const CS* kb, int p,
                                                 int m, CS* X0, CS* X1, int ly, int col, double* rhs,
                                                 Emit emit) {
  const int P = a.P;
  if constexpr (MODE != MODE_R32) {
    circ_solve<GRP, double>(kf, kb, a.p1, a.p2, p, P, m, X0, X1, ly, col, rhs,
                       [&](int k, double v) { emit(k, v * a.invc); });
  } else {
    float y[CHUNK];
#pragma unroll
    for (int k = 0; k < CHUNK; ++k) y[k] = (float)rhs[k];
    circ_solve<GRP, float>(kf, kb, a.p1f, a.p2f, p, P, m, X0, X1, ly, col, y,
                      [&](int k, float v) { y[k] = v * a.invcf; });
    // neighbour values of the FP32 solution for the residual
    const bool act = !GRP || p < P;
    if (act) {
      float e4[4];
      edges(y, m, e4);
#pragma unroll
      for (int q = 0; q < 4; ++q) X0[(4 * p + q) * ly + col] = e4[q];
    }
    __syncthreads();
    if (act) {
      const int pm = p == 0 ? P - 1 : p - 1, pp = p == P - 1 ? 0 : p + 1;
      const float z[4] = {X0[(4 * pm + 2) * ly + col], X0[(4 * pm + 3) * ly + col],
                          X0[(4 * pp) * ly + col], X0[(4 * pp + 1) * ly + col]};
      const double row[5] = {a.bd[0], a.bd[1], a.bd[2], a.bd[1], a.bd[0]};
      residual64([&](int) { return row; }, m, rhs, y, z, rhs);
    }
    float d[CHUNK];
#pragma unroll
    for (int k = 0; k < CHUNK; ++k) d[k] = (float)rhs[k];
    circ_solve<GRP, float>(kf, kb, a.p1f, a.p2f, p, P, m, X1, X0, ly, col, d,
                      [&](int k, float v) { emit(k, (double)y[k] + (double)(v * a.invcf)); });
  }
}

// y/z lines. block = (LY lanes, P chunks). With LY = 64 the chunk index is
// wave-uniform; with LY = 32 a wave holds two chunks of 32 lines each.
// What the epilogue of a y/z row needs besides the solution, fetched into
// registers before the solve so that no store waits on a load: the input
// value (filters add v back) or the current output (accumulating epilogues).
// PRE_LATE loads them in the epilogue instead (fewer registers, a wait per row).
enum Pre { PRE_NONE = 0, PRE_V = 1, PRE_OLD = 2, PRE_LATE = 3 };

// Build-time tuning knobs (register pressure against latency hiding):
//   PYR_FULL_CHUNKS  compile-time row counts when CHUNK divides N
//   PYR_PREFETCH     PRE_V / PRE_OLD (1) or PRE_LATE (0)
//   PYR_YZ_WAVES_PER_EU, PYR_X_WAVES_PER_EU  minimum waves per SIMD the y/z
//                    and x kernels are built for
//   XU               copies in flight per thread in the x-tile loads and stores
//                    (8; 16 measured 0.7% faster on the FP64 Taylor-Green step
//                    at 256^3 and 1.3% slower on the FP32-refined one)
#ifndef PYR_FULL_CHUNKS
#define PYR_FULL_CHUNKS 1
#endif
#ifndef PYR_PREFETCH
#define PYR_PREFETCH 1
#endif
#ifndef PYR_YZ_WAVES_PER_EU
#define PYR_YZ_WAVES_PER_EU 1
#endif
#ifndef PYR_X_WAVES_PER_EU
#define PYR_X_WAVES_PER_EU 1
#endif

// Raw buffer access to one field: a wave's row of 64 adjacent lines is one
// load/store whose lane part sits in a VGPR and whose row part (uniform across
// the wave) sits in an SGPR, so no per-row 64-bit address is formed. Fields
// are below 2 GiB (at most 512^3 FP64). dword3 0x00020000: 32-bit data format,
// the raw (untyped) configuration of gfx9-family buffer descriptors.
constexpr int BUF_DWORD3 = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, int nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, nbytes, BUF_DWORD3);
}
__device__ __forceinline__ double buf_load(double, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ float buf_load(float, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void buf_store(double v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  using U2 = unsigned int __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, voff, soff, 0);
}
__device__ __forceinline__ void buf_store(float v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, 0);
}

// Occupancy the y/z kernels are built for (waves per SIMD). Measured on
// gfx950 at 256^3: capping the FP64-solve circulant kernels without epilogue
// prefetch and the FP32 explicit filter at 128 VGPRs (4 waves) is faster
// despite a few spilled registers (d1 y 39 -> 33 us, gbar y 62 -> 54 us);
// for every other instance the spills cost more than the occupancy gains.
template <typename TS, int SOLVE, int MODE, int PRE>
struct YzWaves {
  static constexpr int v =
      PYR_YZ_WAVES_PER_EU > 1 ? PYR_YZ_WAVES_PER_EU
      : (((MODE == MODE_S32 && SOLVE == SOLVE_CIRC && PRE == PRE_NONE) ||
          (SOLVE == SOLVE_NONE && sizeof(TS) == 4 && PRE == PRE_NONE)) ? 4 : 1);
};

// FULL: every chunk has CHUNK rows (N a multiple of CHUNK), so the row
// predicates vanish at compile time. The circulant solve implies periodic lines.
// The body takes the workgroup's line-group index `blk` (blockIdx.x for one
// field, the member-local index for a grouped launch over several fields).
template <typename TS, int KIND, int SOLVE, int MODE, int LY, bool FULL, int PRE, bool GRP = false>
__device__ __forceinline__ void line_yz_body(const LineArgs& a, long long blk, const TS* __restrict__ in,
                                             TS* __restrict__ out) {
  constexpr int R = Halo<KIND>::R;
  constexpr bool PER = SOLVE == SOLVE_CIRC;
  using CS = typename SolveT<MODE>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x;
  int p = threadIdx.y;
  if (LY == 64) p = __builtin_amdgcn_readfirstlane(p);
  const int tid = threadIdx.x + LY * threadIdx.y;
  const int nthr = LY * blockDim.y;
  CS* Y = reinterpret_cast<CS*>(smem);
  Tabs t;
  if (SOLVE == SOLVE_LU) t = get_tables<MODE>(a, smem + align16((size_t)4 * a.P * LY * sizeof(CS)), tid, nthr);
  const long long L = blk * LY + lane;
  const bool valid = L < a.nlines;
  const long long Lc = valid ? L : (long long)a.nlines - 1;
  const long long base = (Lc % a.A) + a.B * (Lc / a.A);
  // act: this thread's chunk exists (a grouped launch shapes its workgroups
  // for the member with the most chunks)
  const bool act = FULL || !GRP || p < a.P;
  const int s = FULL ? p * CHUNK : (act ? a.cstart[p] : 0);
  const int m = FULL ? CHUNK : (act ? a.cstart[p + 1] - s : 0);
  const __amdgpu_buffer_rsrc_t rin = buf_rsrc(in, a.nbytes), rout = buf_rsrc(out, a.nbytes);
  // the FP32 copy of an FP64 result (LineArgs::out32), half the bytes per element
  [[maybe_unused]] const __amdgpu_buffer_rsrc_t rout32 = buf_rsrc(a.out32, a.nbytes / 2);
  const int lane_off = (int)(base * (long long)sizeof(TS));
  const int row_bytes = (int)(a.stride * (long long)sizeof(TS));
  // row `row` of this line; the row offset is wave-uniform when LY == 64
  auto ld = [&](__amdgpu_buffer_rsrc_t r, int row) -> TS {
    if constexpr (LY == 64) return buf_load(TS(), r, lane_off, row * row_bytes);
    else return buf_load(TS(), r, lane_off + row * row_bytes, 0);
  };
  auto st = [&](TS v, int row) {
    if constexpr (LY == 64) buf_store(v, rout, lane_off, row * row_bytes);
    else buf_store(v, rout, lane_off + row * row_bytes, 0);
  };
  // With fewer than 64 lines per workgroup a wave holds several chunks, so a
  // row index is not wave-uniform; for whole chunks the row offset splits
  // into a per-lane base (the chunk's first row, or the first row of a halo)
  // and a wave-uniform row count, one VGPR in all instead of one per row.
  constexpr bool SPLIT = LY != 64 && FULL;
  const int vmid = lane_off + s * row_bytes;
  // row s + k of this line (0 <= k < CHUNK)
  auto ldk = [&](__amdgpu_buffer_rsrc_t r, int k) -> TS {
    if constexpr (SPLIT) return buf_load(TS(), r, vmid, k * row_bytes);
    else return ld(r, s + k);
  };
  auto stk = [&](TS v, int k) {
    if constexpr (SPLIT) buf_store(v, rout, vmid, k * row_bytes);
    else st(v, s + k);
  };
  TS pre[CHUNK];
  if constexpr (PRE == PRE_OLD) {
#pragma unroll
    for (int k = 0; k < CHUNK; ++k) pre[k] = k < m ? ldk(rout, k) : (TS)0;
  }

  TS win[CHUNK + 2 * R];
  if constexpr (SPLIT) {
    // rows s-R .. s-1 (wrapped to N-R .. N-1 for the first chunk of a
    // periodic line), s .. s+CHUNK-1, s+CHUNK .. s+CHUNK+R-1 (wrapped to
    // 0 .. R-1 for the last chunk); a bounded line's missing halo rows are 0
    const bool wrap = PER || a.periodic;
    const int vlo = lane_off + (s >= R ? s - R : a.N - R) * row_bytes;
    const int vhi = lane_off + (s + CHUNK < a.N ? s + CHUNK : 0) * row_bytes;
    const bool lo_ok = act && (wrap || s >= R);
    const bool hi_ok = act && (wrap || s + CHUNK < a.N);
#pragma unroll
    for (int w = 0; w < CHUNK + 2 * R; ++w) {
      TS v = (TS)0;
      if (w < R) {
        if (lo_ok) v = buf_load(TS(), rin, vlo, w * row_bytes);
      } else if (w < CHUNK + R) {
        if (act) v = buf_load(TS(), rin, vmid, (w - R) * row_bytes);
      } else {
        if (hi_ok) v = buf_load(TS(), rin, vhi, (w - CHUNK - R) * row_bytes);
      }
      win[w] = v;
    }
  } else {
#pragma unroll
    for (int w = 0; w < CHUNK + 2 * R; ++w) {
      int row = s - R + w;
      TS v = (TS)0;
      if (act && w < m + 2 * R) {
        if (PER || a.periodic) {
          row = row < 0 ? row + a.N : (row >= a.N ? row - a.N : row);
          v = ld(rin, row);
        } else if (row >= 0 && row < a.N) {
          v = ld(rin, row);
        }
      }
      win[w] = v;
    }
  }
  // grouped launches (runtime row counts, longer register live ranges) re-read
  // v when it is added back instead of holding the chunk across the solve
  constexpr bool HOLDV = PRE == PRE_V && !GRP;
  if constexpr (HOLDV) {
#pragma unroll
    for (int k = 0; k < CHUNK; ++k) pre[k] = win[k + R];
  }
  double rhs[CHUNK];
  rhs_interior<KIND>(a, win, rhs, m);
  if (!PER && act) rhs_boundary(a, s, m, rhs, [&](int i) { return ld(rin, i); });
  // row k of the result: metric scale, filters add v back, fused epilogue, store
  auto emit = [&](int k, double y) {
    if (!valid) return;
    double val = y * a.scale;
    if constexpr (HOLDV) val += (double)pre[k];
    if constexpr (PRE == PRE_V && GRP) val += (double)ldk(rin, k);
    if constexpr (PRE == PRE_LATE)
      if (a.addv) val += (double)ldk(rin, k);
    TS old = (TS)0;
    if constexpr (PRE == PRE_OLD) old = pre[k];
    if constexpr (PRE == PRE_LATE) old = (a.epi == EPI_ADD || a.epi == EPI_MAXABS) ? ldk(rout, k) : (TS)0;
    // a filter with an accumulating epilogue (only through pyr_apply) reads
    // the output in place
    if constexpr (PRE == PRE_V) old = (a.epi == EPI_ADD || a.epi == EPI_MAXABS) ? ldk(rout, k) : (TS)0;
    const TS r = epilogue<TS>(a.epi, a.epi_c, val, old);
    if (sizeof(TS) != 8 || GRP || !a.only32) stk(r, k);
    // (store epilogues only: the accumulating variants keep their registers)
    if constexpr (sizeof(TS) == 8 && PRE != PRE_OLD && !GRP)
      if (a.out32) {
        if constexpr (LY == 64) buf_store((float)r, rout32, lane_off / 2, (s + k) * (row_bytes / 2));
        else if constexpr (SPLIT) buf_store((float)r, rout32, vmid / 2, k * (row_bytes / 2));
        else buf_store((float)r, rout32, lane_off / 2 + (s + k) * (row_bytes / 2), 0);
      }
  };
  if (SOLVE == SOLVE_CIRC) {
    solve_chunk_circ<MODE, CS, GRP>(a, kmat<MODE>(a, 0), kmat<MODE>(a, 1), p, m, Y, Y + 4 * a.P * LY, LY, lane,
                                    rhs, emit);
  } else {
    if (SOLVE == SOLVE_LU) {
      __syncthreads();  // staged tables visible
      solve_chunk_lu<MODE, CS>(a, t, p, s, m, Y, LY, lane, rhs, act);
    }
#pragma unroll
    for (int k = 0; k < CHUNK; ++k)
      if (k < m) emit(k, rhs[k]);
  }
}

template <typename TS, int KIND, int SOLVE, int MODE, int LY, bool FULL, int PRE>
__global__ void __launch_bounds__(WG_MAX) __attribute__((amdgpu_waves_per_eu(YzWaves<TS, SOLVE, MODE, PRE>::v, 8)))
k_line_yz(LineArgs a, const TS* __restrict__ in, TS* __restrict__ out) {
  line_yz_body<TS, KIND, SOLVE, MODE, LY, FULL, PRE>(a, blockIdx.x, in, out);
}

// 16-byte copy from HBM straight into LDS (global_load_lds_dwordx4): lane i's
// bytes land at lds + 16 i. The builtin is given to the device pass only:
// with it in a kernel body the host pass drops the kernel's launch stub.
__device__ __forceinline__ void lds_copy16(const void* g, void* lds) {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)lds, 16, 0, 0);
#endif
}

// y/z lines, software-pipelined. A persistent workgroup walks groups of LG
// adjacent lines (LG lanes x P chunks). The N rows of the next group are
// copied from HBM into one of two LDS tiles by global->LDS DMA (no registers,
// no waits) while the current group is solved from the other tile, so HBM
// reads overlap the solve even with few waves per CU. Results leave as
// coalesced buffer stores. Requirements, checked on the host: periodic lines,
// CHUNK | N, the LG lines of a group adjacent in memory (LG | A), 16-byte
// aligned rows and both tiles within the LDS budget.
template <typename TS, int KIND, int SOLVE, int MODE, int LG, int PRE>
__global__ void __launch_bounds__(WG_MAX) k_line_yz_pipe(LineArgs a, const TS* __restrict__ in,
                                                         TS* __restrict__ out) {
  constexpr int R = Halo<KIND>::R;
  using CS = typename SolveT<MODE>::type;
  constexpr int VE = 16 / (int)sizeof(TS);        // elements one lane moves
  constexpr int LPR = LG * (int)sizeof(TS) / 16;  // lanes per tile row
  constexpr int RPI = 64 / LPR;                   // tile rows per wave instruction (1 KiB)
  static_assert(LPR >= 1 && 64 % LPR == 0, "tile row must be a whole number of 16-byte pieces");
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = a.N, P = a.P;
  const int lane = threadIdx.x;
  const int p = threadIdx.y;
  const int flat = threadIdx.x + LG * threadIdx.y;
  const int wave = __builtin_amdgcn_readfirstlane(flat >> 6);
  const int nwaves = (LG * (int)blockDim.y) >> 6;
  const int lane64 = flat & 63;
  // tile rows: R wrapped rows, the N rows of the lines, R wrapped rows, and
  // up to RPI-1 surplus rows (copies are issued in runs of RPI rows)
  const int NT = ((N + 2 * R + RPI - 1) / RPI) * RPI;
  TS* tiles = reinterpret_cast<TS*>(smem);
  const int tile_elems = NT * LG;
  CS* X0 = reinterpret_cast<CS*>(smem + align16(2 * (size_t)tile_elems * sizeof(TS)));
  CS* X1 = X0 + 4 * P * LG;
  CS* Kl = X1 + 4 * P * LG;  // carry rows, forward then backward (4P^2 each)
  if constexpr (SOLVE == SOLVE_CIRC) {
    const CS* kf = kmat<MODE>(a, 0);
    const CS* kb = kmat<MODE>(a, 1);
    const int nk = 4 * P * P;
    for (int e = flat; e < nk; e += LG * (int)blockDim.y) {
      Kl[e] = kf[e];
      Kl[nk + e] = kb[e];
    }
  }
  const int ngroups = a.nlines / LG;
  const __amdgpu_buffer_rsrc_t rout = buf_rsrc(out, a.nbytes);
  const int row_bytes = (int)(a.stride * (long long)sizeof(TS));
  const int s = p * CHUNK;
  constexpr int m = CHUNK;

  // tile row j (0 <= j < NT) holds line row (j - R) mod N
  auto issue = [&](int g, int t) {
    const long long L0 = (long long)g * LG;
    const TS* src = in + (L0 % a.A) + a.B * (L0 / a.A);
    TS* dst = tiles + t * tile_elems;
    const int c = (lane64 % LPR) * VE;
    for (int j0 = wave * RPI; j0 < NT; j0 += nwaves * RPI) {
      int r = j0 + lane64 / LPR - R;
      r = r < 0 ? r + N : (r >= N ? r - N : r);
      lds_copy16(src + (long long)a.stride * r + c, dst + j0 * LG);
    }
  };

  int t = 0;
  if ((int)blockIdx.x < ngroups) issue(blockIdx.x, 0);
  for (int g = blockIdx.x; g < ngroups; g += gridDim.x, t ^= 1) {
    // this wave's copies (and last group's stores) are done; the barrier makes
    // every wave's copies visible and retires all reads of the other tile
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    const long long L = (long long)g * LG + lane;
    const int lane_off = (int)(((L % a.A) + a.B * (L / a.A)) * (long long)sizeof(TS));
    // accumulating epilogues read the current output before the next copies are
    // queued, so waiting for them never waits for the copies
    TS pre[CHUNK];
    if constexpr (PRE == PRE_OLD) {
#pragma unroll
      for (int k = 0; k < CHUNK; ++k) pre[k] = buf_load(TS(), rout, lane_off + (s + k) * row_bytes, 0);
    }
    const int gn = g + gridDim.x;
    if (gn < ngroups) issue(gn, t ^ 1);
    // window rows s-R .. s+CHUNK+R-1 are tile rows s .. s+CHUNK+2R-1 (no wrap)
    const TS* col0 = tiles + t * tile_elems + s * LG + lane;
    TS win[CHUNK + 2 * R];
#pragma unroll
    for (int w = 0; w < CHUNK + 2 * R; ++w) win[w] = col0[w * LG];
    double rhs[CHUNK];
    rhs_interior<KIND>(a, win, rhs, m);
    auto emit = [&](int k, double y) {
      double val = y * a.scale;
      // filters add v back: re-read from the tile (intact until the copies
      // of the group after next) rather than holding CHUNK values across the
      // solve, which would spill
      if constexpr (PRE == PRE_V) val += (double)col0[(k + R) * LG];
      TS old = (TS)0;
      if constexpr (PRE == PRE_OLD) old = pre[k];
      buf_store(epilogue<TS>(a.epi, a.epi_c, val, old), rout, lane_off + (s + k) * row_bytes, 0);
    };
    if constexpr (SOLVE == SOLVE_CIRC) {
      solve_chunk_circ<MODE, CS>(a, Kl, Kl + 4 * P * P, p, m, X0, X1, LG, lane, rhs, emit);
    } else {
#pragma unroll
      for (int k = 0; k < CHUNK; ++k) emit(k, rhs[k]);
    }
  }
}

// 16-byte vector of the storage type
template <typename TS> struct Vec16;
template <> struct Vec16<double> { using type = double2; static constexpr int W = 2; };
template <> struct Vec16<float> { using type = float4; static constexpr int W = 4; };

// r <- f(r, o) per component, without taking the vectors' addresses (which
// would move them to scratch)
template <typename F>
__device__ __forceinline__ void vzip(double2& r, const double2& o, F f) {
  r.x = f(r.x, o.x);
  r.y = f(r.y, o.y);
}
template <typename F>
__device__ __forceinline__ void vzip(float4& r, const float4& o, F f) {
  r.x = f(r.x, o.x);
  r.y = f(r.y, o.y);
  r.z = f(r.z, o.z);
  r.w = f(r.w, o.w);
}

#ifndef XU
#define XU 8  // unroll of the tile load/store loops (a build-time knob, above)
#endif

// x lines: LX contiguous lines per workgroup staged in LDS (row pitch
// a.pitch). Rows move between HBM and LDS as 16-byte vectors when the line
// length and pointers allow it, XU copies in flight per thread. The tile is
// reused for the results, which leave through the same coalesced path.
// LU solve: the tile region also holds the tables and edge exchange once the
// windows are in registers. Circulant solve: the carry exchange sits after
// the tile.
template <typename TS, int KIND, int SOLVE, int MODE, bool FULL, bool GRP = false>
__device__ __forceinline__ void line_x_body(const LineArgs& a, long long blk, const TS* __restrict__ in,
                                            TS* __restrict__ out) {
  constexpr int R = Halo<KIND>::R;
  constexpr bool PER = SOLVE == SOLVE_CIRC;
  using CS = typename SolveT<MODE>::type;
  using V = typename Vec16<TS>::type;
  constexpr int VW = Vec16<TS>::W;
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = a.N, pitch = a.pitch, LX = a.lx;
  TS* tile = reinterpret_cast<TS*>(smem);
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  const int l = tid % LX;
  const int p = tid / LX;
  const long long L0 = blk * LX;
  const long long rem = (long long)a.nlines - L0;
  const int nval = (int)(rem < LX ? rem : LX);
  const TS* src = in + L0 * N;
  if (a.vec) {
    const int vpl = N / VW, nvec = nval * vpl;
    const V* sv = reinterpret_cast<const V*>(src);
    // global and LDS never alias, so the unrolled loads issue back to back
#pragma unroll XU
    for (int e = tid; e < nvec; e += nthr) {
      const int ll = e / vpl, i = e - ll * vpl;
      *reinterpret_cast<V*>(tile + ll * pitch + i * VW) = sv[e];
    }
  } else {
    const int ne = nval * N;
#pragma unroll XU
    for (int e = tid; e < ne; e += nthr) {
      const int ll = e / N, i = e - ll * N;
      tile[ll * pitch + i] = src[e];
    }
  }
  __syncthreads();
  const int lr = l < nval ? l : 0;
  const TS* row0 = tile + lr * pitch;
  // act: this thread's chunk exists (a grouped launch sizes its workgroups
  // for the member with the most threads per workgroup)
  const bool act = !GRP || p < a.P;
  const int s = FULL ? p * CHUNK : (act ? a.cstart[p] : 0);
  const int m = FULL ? CHUNK : (act ? a.cstart[p + 1] - s : 0);
  TS win[CHUNK + 2 * R];
#pragma unroll
  for (int w = 0; w < CHUNK + 2 * R; ++w) {
    int row = s - R + w;
    TS v = (TS)0;
    if (act && w < m + 2 * R) {
      if (PER || a.periodic) {
        row = row < 0 ? row + N : (row >= N ? row - N : row);
        v = row0[row];
      } else if (row >= 0 && row < N) {
        v = row0[row];
      }
    }
    win[w] = v;
  }
  double rhs[CHUNK];
  rhs_interior<KIND>(a, win, rhs, m);
  if (!PER && act) rhs_boundary(a, s, m, rhs, [&](int i) { return row0[i]; });
  if (SOLVE == SOLVE_LU) {
    __syncthreads();  // the tile is consumed; reuse its LDS
    CS* Y = reinterpret_cast<CS*>(smem);
    const Tabs t = get_tables<MODE>(a, smem + align16((size_t)4 * a.P * LX * sizeof(CS)), tid, nthr);
    __syncthreads();
    solve_chunk_lu<MODE, CS>(a, t, p, s, m, Y, LX, l, rhs, act);
    __syncthreads();  // tables and edge values no longer read
  } else if (SOLVE == SOLVE_NONE) {
    __syncthreads();  // the tile is consumed
  }
  TS* rowo = tile + l * pitch;
  // filters add v back here, read from the tile row before the result takes
  // its place (the windows are in registers by now; with the LU solve the
  // tile holds the tables instead and v is added on the way out)
  constexpr bool VTILE = SOLVE != SOLVE_LU;
  auto emit = [&](int k, double y) {
    if (act && l < nval) {
      double val = y * a.scale;
      if (VTILE && a.addv) val += (double)rowo[s + k];
      rowo[s + k] = (TS)val;
    }
  };
  if (SOLVE == SOLVE_CIRC) {
    // the first carry barrier also retires every read of the tile
    CS* X0 = reinterpret_cast<CS*>(smem + align16((size_t)LX * pitch * sizeof(TS)));
    solve_chunk_circ<MODE, CS, GRP>(a, kmat<MODE>(a, 0), kmat<MODE>(a, 1), p, m, X0,
                                    X0 + circ_words<MODE>() * a.P * LX, LX, l, rhs, emit);
  } else {
#pragma unroll
    for (int k = 0; k < CHUNK; ++k)
      if (k < m) emit(k, rhs[k]);
  }
  __syncthreads();
  TS* dst = out + L0 * N;
  const bool need_old = (a.epi == EPI_ADD || a.epi == EPI_MAXABS);
  if constexpr (sizeof(TS) == 8 && !GRP) {
    if (a.out32) {
      // FP64 result rounded to FP32 (store epilogue), with or without the
      // FP64 store
      float* d32 = a.out32 + L0 * N;
      const int ne = nval * N;
#pragma unroll XU
      for (int e = tid; e < ne; e += nthr) {
        const int ll = e / N, i = e - ll * N;
        double val = (double)tile[ll * pitch + i];
        if (!VTILE && a.addv) val = (double)(TS)(val + (double)src[e]);
        if (!a.only32) dst[e] = (TS)val;
        d32[e] = (float)val;
      }
      return;
    }
  }
  if (a.vec) {
    const int vpl = N / VW, nvec = nval * vpl;
    const V* sv = reinterpret_cast<const V*>(src);
    V* dv = reinterpret_cast<V*>(dst);
#pragma unroll XU
    for (int e = tid; e < nvec; e += nthr) {
      const int ll = e / vpl, i = e - ll * vpl;
      V r = *reinterpret_cast<const V*>(tile + ll * pitch + i * VW);
      if (!VTILE && a.addv) vzip(r, sv[e], [](TS y, TS v) { return (TS)((double)y + (double)v); });
      if (a.epi != EPI_STORE) {
        const V o = need_old ? dv[e] : V{};
        vzip(r, o, [&](TS y, TS old) { return epilogue<TS>(a.epi, a.epi_c, (double)y, old); });
      }
      dv[e] = r;
    }
  } else {
    const int ne = nval * N;
#pragma unroll XU
    for (int e = tid; e < ne; e += nthr) {
      const int ll = e / N, i = e - ll * N;
      double val = (double)tile[ll * pitch + i];
      if (!VTILE && a.addv) val = (double)(TS)(val + (double)src[e]);
      const TS old = need_old ? dst[e] : (TS)0;
      dst[e] = epilogue<TS>(a.epi, a.epi_c, val, old);
    }
  }
}

template <typename TS, int KIND, int SOLVE, int MODE, bool FULL>
__global__ void __launch_bounds__(WG_MAX) __attribute__((amdgpu_waves_per_eu(PYR_X_WAVES_PER_EU, 8)))
k_line_x(LineArgs a, const TS* __restrict__ in, TS* __restrict__ out) {
  line_x_body<TS, KIND, SOLVE, MODE, FULL>(a, blockIdx.x, in, out);
}

// x lines, software-pipelined (periodic lines of whole 32-row chunks whose
// length is a whole number of KiB, circulant solve or explicit filter). A
// persistent workgroup walks groups of LX lines; the next group's lines are
// copied HBM -> LDS by global->LDS DMA into the second of two tiles while the
// current group is solved from the first, and the results leave through the
// current tile with coalesced vector stores. Tile row pitch: N + one 16-byte
// vector, so the windows (lanes over lines) read conflict-free.
template <typename TS, int KIND, int SOLVE, int MODE>
__global__ void __launch_bounds__(WG_MAX) __attribute__((amdgpu_waves_per_eu(PYR_X_WAVES_PER_EU, 8)))
k_line_x_pipe(LineArgs a, const TS* __restrict__ in, TS* __restrict__ out) {
  constexpr int R = Halo<KIND>::R;
  using CS = typename SolveT<MODE>::type;
  using V = typename Vec16<TS>::type;
  constexpr int VW = Vec16<TS>::W;
  constexpr int EPK = 1024 / (int)sizeof(TS);  // elements of one wave's 1 KiB copy
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = a.N, LX = a.lx, pitch = a.pitch;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int l = tid % LX, p = tid / LX;
  const int lane64 = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nwaves = nthr >> 6;
  const int tile_elems = LX * pitch;
  TS* tiles = reinterpret_cast<TS*>(smem);
  CS* X0 = reinterpret_cast<CS*>(smem + align16(2 * (size_t)tile_elems * sizeof(TS)));
  CS* X1 = X0 + 4 * a.P * LX;
  const int ngroups = (a.nlines + LX - 1) / LX;
  const int ppl = N / EPK;  // 1 KiB pieces per line
  auto issue = [&](int g, int t) {
    const long long L0 = (long long)g * LX;
    const long long rem = (long long)a.nlines - L0;
    const int nval = (int)(rem < LX ? rem : LX);
    const TS* src = in + L0 * N;
    TS* dst = tiles + t * tile_elems;
    for (int j = wave; j < nval * ppl; j += nwaves) {
      const int ll = j / ppl, q = j - ll * ppl;
      lds_copy16(src + (long long)ll * N + q * EPK + lane64 * VW, dst + ll * pitch + q * EPK);
    }
  };
  const int s = p * CHUNK;
  constexpr int m = CHUNK;
  const bool need_old = (a.epi == EPI_ADD || a.epi == EPI_MAXABS);
  int t = 0;
  if ((int)blockIdx.x < ngroups) issue(blockIdx.x, 0);
  for (int g = blockIdx.x; g < ngroups; g += gridDim.x, t ^= 1) {
    // this wave's copies and last group's stores are done; the barrier makes
    // every wave's copies visible and retires all reads of the other tile
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    const int gn = g + gridDim.x;
    if (gn < ngroups) issue(gn, t ^ 1);
    const long long L0 = (long long)g * LX;
    const long long rem = (long long)a.nlines - L0;
    const int nval = (int)(rem < LX ? rem : LX);
    TS* tile = tiles + t * tile_elems;
    const int lr = l < nval ? l : 0;
    const TS* row0 = tile + lr * pitch;
    TS win[CHUNK + 2 * R];
#pragma unroll
    for (int w = 0; w < CHUNK + 2 * R; ++w) {
      int row = s - R + w;
      row = row < 0 ? row + N : (row >= N ? row - N : row);
      win[w] = row0[row];
    }
    double rhs[CHUNK];
    rhs_interior<KIND>(a, win, rhs, m);
    TS* rowo = tile + l * pitch;
    auto emit = [&](int k, double y) {
      if (l < nval) {
        double val = y * a.scale;
        if (a.addv) val += (double)rowo[s + k];
        rowo[s + k] = (TS)val;
      }
    };
    if constexpr (SOLVE == SOLVE_CIRC) {
      // the first carry barrier also retires every window read of the tile
      solve_chunk_circ<MODE, CS>(a, kmat<MODE>(a, 0), kmat<MODE>(a, 1), p, m, X0, X1, LX, l, rhs, emit);
    } else {
      __syncthreads();  // every window read before any result lands in the tile
#pragma unroll
      for (int k = 0; k < CHUNK; ++k) emit(k, rhs[k]);
    }
    __syncthreads();
    const int vpl = N / VW, nvec = nval * vpl;
    V* dv = reinterpret_cast<V*>(out + L0 * N);
#pragma unroll XU
    for (int e = tid; e < nvec; e += nthr) {
      const int ll = e / vpl, i = e - ll * vpl;
      V r = *reinterpret_cast<const V*>(tile + ll * pitch + i * VW);
      if (a.epi != EPI_STORE) {
        const V o = need_old ? dv[e] : V{};
        vzip(r, o, [&](TS y, TS old) { return epilogue<TS>(a.epi, a.epi_c, (double)y, old); });
      }
      dv[e] = r;
    }
  }
}

// x lines, software-pipelined through registers (periodic lines of whole
// 32-row chunks, 16-byte aligned rows, circulant solve or explicit filter).
// A persistent workgroup walks groups of LX lines through ONE LDS tile: while
// it solves group g from the tile, the rows of its next group are already in
// flight from HBM into registers (CHUNK elements per thread, as 16-byte
// vectors), and they go into the tile once group g's results have left it.
// Against k_line_x_pipe (a second LDS tile filled by DMA) this keeps the LDS
// footprint, hence the workgroups per CU, of the one-pass kernel, and pays
// for the prefetch in registers, which the LDS-bound occupancy leaves free.
// Every LX * P threads move LX * N elements: CHUNK per thread exactly. The
// prefetch is issued after the windows have been turned into right-hand
// sides, so that it overlaps the solve and the stores.
template <typename TS, int KIND, int SOLVE, int MODE>
__global__ void __launch_bounds__(WG_MAX) __attribute__((amdgpu_waves_per_eu(2, 8)))
k_line_x_rpipe(LineArgs a, const TS* __restrict__ in, TS* __restrict__ out) {
  constexpr int R = Halo<KIND>::R;
  using CS = typename SolveT<MODE>::type;
  using V = typename Vec16<TS>::type;
  constexpr int VW = Vec16<TS>::W;
  constexpr int NV = CHUNK / VW;  // prefetched vectors per thread
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = a.N, LX = a.lx, pitch = a.pitch;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int l = tid % LX, p = tid / LX;
  const int vpl = N / VW;
  TS* tile = reinterpret_cast<TS*>(smem);
  CS* X0 = reinterpret_cast<CS*>(smem + align16((size_t)LX * pitch * sizeof(TS)));
  CS* X1 = X0 + circ_words<MODE>() * a.P * LX;
  const int ngroups = (a.nlines + LX - 1) / LX;
  auto nval_of = [&](int g) {
    const long long rem = (long long)a.nlines - (long long)g * LX;
    return (int)(rem < LX ? rem : LX);
  };
  V pf[NV];
  const int s = p * CHUNK;
  constexpr int m = CHUNK;
  const bool need_old = (a.epi == EPI_ADD || a.epi == EPI_MAXABS);
  int g = blockIdx.x;
  if (g < ngroups) {
    const int nvec = nval_of(g) * vpl;
    const V* sv = reinterpret_cast<const V*>(in + (long long)g * LX * N);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const int e = tid + j * nthr;
      pf[j] = e < nvec ? sv[e] : V{};
    }
  }
  for (; g < ngroups; g += gridDim.x) {
    const int nval = nval_of(g);
    const int nvec = nval * vpl;
    // group g's rows into the tile (the previous group's stores have read it:
    // barrier at the end of the last iteration)
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const int e = tid + j * nthr;
      if (e < nvec) {
        const int ll = e / vpl, i = e - ll * vpl;
        *reinterpret_cast<V*>(tile + ll * pitch + i * VW) = pf[j];
      }
    }
    __syncthreads();
    const int lr = l < nval ? l : 0;
    const TS* row0 = tile + lr * pitch;
    double rhs[CHUNK];
    {
      TS win[CHUNK + 2 * R];
#pragma unroll
      for (int w = 0; w < CHUNK + 2 * R; ++w) {
        int row = s - R + w;
        row = row < 0 ? row + N : (row >= N ? row - N : row);
        win[w] = row0[row];
      }
      rhs_interior<KIND>(a, win, rhs, m);
    }
    // the next group's rows travel while this group is solved and stored
    // (issued once the windows are consumed: their registers take the rows)
    const int gn = g + gridDim.x;
    __builtin_amdgcn_sched_barrier(0);  // keep the prefetch below the windows
    if (gn < ngroups) {
      const int nvn = nval_of(gn) * vpl;
      const V* sv = reinterpret_cast<const V*>(in + (long long)gn * LX * N);
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const int e = tid + j * nthr;
        pf[j] = e < nvn ? sv[e] : V{};
      }
    }
    TS* rowo = tile + l * pitch;
    auto emit = [&](int k, double y) {
      if (l < nval) {
        double val = y * a.scale;
        if (a.addv) val += (double)rowo[s + k];
        rowo[s + k] = (TS)val;
      }
    };
    if constexpr (SOLVE == SOLVE_CIRC) {
      // the first carry barrier also retires every window read of the tile
      solve_chunk_circ<MODE, CS>(a, kmat<MODE>(a, 0), kmat<MODE>(a, 1), p, m, X0, X1, LX, l, rhs, emit);
    } else {
      __syncthreads();  // every window read before any result lands in the tile
#pragma unroll
      for (int k = 0; k < CHUNK; ++k) emit(k, rhs[k]);
    }
    __syncthreads();
    const long long L0 = (long long)g * LX;
    if constexpr (sizeof(TS) == 8) {
      if (a.out32) {
        // FP64 result rounded to FP32, with or without the FP64 store
        double* dst = reinterpret_cast<double*>(out) + L0 * N;
        float* d32 = a.out32 + L0 * N;
#pragma unroll XU
        for (int e = tid; e < nvec; e += nthr) {
          const int ll = e / vpl, i = e - ll * vpl;
          const double2 r = *reinterpret_cast<const double2*>(tile + ll * pitch + i * VW);
          if (!a.only32) reinterpret_cast<double2*>(dst)[e] = r;
          reinterpret_cast<float2*>(d32)[e] = make_float2((float)r.x, (float)r.y);
        }
        __syncthreads();  // every read of the tile before the next group lands
        continue;
      }
    }
    V* dv = reinterpret_cast<V*>(out + L0 * N);
#pragma unroll XU
    for (int e = tid; e < nvec; e += nthr) {
      const int ll = e / vpl, i = e - ll * vpl;
      V r = *reinterpret_cast<const V*>(tile + ll * pitch + i * VW);
      if (a.epi != EPI_STORE) {
        const V o = need_old ? dv[e] : V{};
        vzip(r, o, [&](TS y, TS old) { return epilogue<TS>(a.epi, a.epi_c, (double)y, old); });
      }
      dv[e] = r;
    }
    __syncthreads();  // every read of the tile before the next group lands
  }
}

// Grouped launches: one launch applies the same operator to the fields of
// several simulations (an ensemble's members, each at `off` elements into
// one allocation), with lines of any lengths. Entry j covers workgroups
// [blk0_j, blk0_{j+1}); a workgroup finds its member by binary search
// (wave-uniform) and runs the single-field body with that member's line
// arguments. Workgroups are shaped for the member with the most chunks (or
// threads); the bodies leave threads without a chunk idle at the barriers.
struct GrpEntry {
  LineArgs a;
  long long off;
  long long blk0;
};

__device__ __forceinline__ int grp_member(const GrpEntry* __restrict__ g, int M, long long b) {
  int lo = 0, hi = M - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g[mid].blk0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename TS, int KIND, int SOLVE, int MODE>
__global__ void __launch_bounds__(WG_MAX) __attribute__((amdgpu_waves_per_eu(PYR_X_WAVES_PER_EU, 8)))
k_line_x_grp(const GrpEntry* __restrict__ g, int M, const TS* __restrict__ in, TS* __restrict__ out) {
  const int j = grp_member(g, M, blockIdx.x);
  // a private copy: its fields load once (wave-uniform) and no store of the
  // body can alias them
  const LineArgs a = g[j].a;
  const long long off = g[j].off, b0 = g[j].blk0;
  line_x_body<TS, KIND, SOLVE, MODE, false, true>(a, (long long)blockIdx.x - b0, in + off, out + off);
}

template <typename TS, int KIND, int SOLVE, int MODE, int LY, int PRE>
__global__ void __launch_bounds__(WG_MAX) __attribute__((amdgpu_waves_per_eu(YzWaves<TS, SOLVE, MODE, PRE>::v, 8)))
k_line_yz_grp(const GrpEntry* __restrict__ g, int M, const TS* __restrict__ in, TS* __restrict__ out) {
  const int j = grp_member(g, M, blockIdx.x);
  const LineArgs a = g[j].a;
  const long long off = g[j].off, b0 = g[j].blk0;
  line_yz_body<TS, KIND, SOLVE, MODE, LY, false, PRE, true>(a, (long long)blockIdx.x - b0, in + off, out + off);
}

// null operator of a direction with fewer than 4 points (zero or copy)
template <typename TS>
__global__ void k_null(long long n, int copy, int epi, double c, const TS* __restrict__ in,
                       TS* __restrict__ out, float* __restrict__ out32) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double val = copy ? (double)in[i] : 0.0;
  const TS old = (epi == EPI_ADD || epi == EPI_MAXABS) ? out[i] : (TS)0;
  const TS r = epilogue<TS>(epi, c, val, old);
  if (out) out[i] = r;
  if (out32) out32[i] = (float)r;
}

// RK4 stage update (pyranda.py:800-804), one pass over F, PHI and U. With
// A == 0 (the first stage) A*PHI is zero and PHI is not read, so it needs no
// clearing before the step. dtp (if not null) holds dt in device memory: one
// value, or with mid (the member index of every point of an ensemble's
// fields) one value per member.
template <typename TF, typename TS>
__global__ void k_rk_update(long long n, const TF* __restrict__ F, double* __restrict__ PHI,
                            const TS* __restrict__ Uin, TS* __restrict__ Uout,
                            double* __restrict__ U64, const double* __restrict__ dtp,
                            const int* __restrict__ mid, double dt, double A, double B) {
  if (dtp && !mid) dt = *dtp;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double dti = mid ? dtp[mid[i]] : dt;
    const double tmp1 = A == 0.0 ? 0.0 : A * PHI[i];
    // explicit FMAs: the generated kernels that fuse a right-hand side into
    // this update round the same way
    const double phi = fma(dti, (double)F[i], tmp1);
    PHI[i] = phi;
    if (U64) {
      const double u = fma(B, phi, U64[i]);
      U64[i] = u;
      if (Uout) Uout[i] = (TS)u;
    } else {
      Uout[i] = (TS)fma(B, phi, (double)Uin[i]);
    }
  }
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct DevPlan {
  void* mem = nullptr;  // one allocation holding every table
  const int* cstart = nullptr;
  const double *lu = nullptr, *rc = nullptr, *gsel = nullptr, *band = nullptr, *bcoef = nullptr;
  const float *lu32 = nullptr, *rc32 = nullptr, *gsel32 = nullptr;
  const short* bidx = nullptr;
  const int* bnt = nullptr;
  int ntab = 0;
  const double *kf = nullptr, *kb = nullptr;
  const float *kf32 = nullptr, *kb32 = nullptr;
};

struct pyr_ctx {
  int n[3] = {1, 1, 1};
  int bc[6] = {PYR_BC_NONE, PYR_BC_NONE, PYR_BC_NONE, PYR_BC_NONE, PYR_BC_NONE, PYR_BC_NONE};
  double x1[3] = {0, 0, 0};
  double dpatch[3] = {1, 1, 1};  // patch dx (metric of derivatives)
  double dmesh[3] = {1, 1, 1};   // mesh d1..d3 (ring length scales)
  double gridlen = 1.0;
  bool device = false;
  hipStream_t stream = nullptr;
  OpPlan plan[NOPS][3];
  DevPlan dplan[NOPS][3];
  // second member of the symmetric/antisymmetric pair (compact.f90:78-84):
  // variant -1 at each SYMM end of a direction that has one
  bool has_anti[3] = {false, false, false};
  OpPlan plan_a[NOPS][3];
  DevPlan dplan_a[NOPS][3];
  void* field_tmp[2] = {nullptr, nullptr};  // one field-sized temp per precision
  int periodic_solver = PYR_PSOLVE_CIRCULANT;
  int r32_solve = PYR_R32_SOLVE_REFINED;
  int x_lds_kb = 40;                        // LDS budget of one x-line workgroup
  int yz_pipe = 1;                           // pipelined y/z kernels where the shape allows
  // per operator, direction and solve mode (MODE_*): tuned values of the two
  // knobs above, -1 = the context default (pyr_set_tuning)
  int yz_lines = 0;                          // lines per plain y/z workgroup (0: 64, or 32 past 8 chunks; 16, 32)
  int tune_xlds[NOPS][3][3];
  int tune_pipe[NOPS][3][3];
  int tune_ly[NOPS][3][3];
  // per operator and direction: FP32 solve arithmetic (PYR_R32_SOLVE_*), -1 = r32_solve
  int tune_solve[NOPS][3];
  pyr_ctx() {
    for (int o = 0; o < NOPS; ++o)
      for (int d = 0; d < 3; ++d) {
        tune_solve[o][d] = -1;
        for (int m = 0; m < 3; ++m) tune_xlds[o][d][m] = tune_pipe[o][d][m] = tune_ly[o][d][m] = -1;
      }
  }
  int ncu = 256;                             // compute units of the device
  void* noalloc_tmp = nullptr;               // this device's copy of `scratch` (pyr_sfilter_noalloc)
  float* out32 = nullptr;                    // FP32 copy of the next FP64 sweep's output
  int only32 = 0;                            // ... and no FP64 store
  size_t npts() const { return (size_t)n[0] * n[1] * n[2]; }
};

static int upload_plan(const OpPlan& pl, DevPlan& d) {
  if (pl.null_op) return 0;
  const size_t n = pl.N, P = pl.P;
  std::vector<std::pair<const void*, size_t>> parts;
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t off = 0;
  size_t o_cs = off; off = align(off + (P + 1) * sizeof(int));
  // lu|rc|gsel contiguous (one block copied to LDS by the kernels)
  size_t o_lu = off;
  size_t o_rc = o_lu + n * 5 * sizeof(double);
  size_t o_gs = o_rc + n * 4 * sizeof(double);
  off = align(o_gs + P * 16 * P * sizeof(double));
  size_t o_bd = off; off = align(off + n * 5 * sizeof(double));
  size_t o_bc = off; off = align(off + 2 * NB * MAXT * sizeof(double));
  size_t o_bi = off; off = align(off + 2 * NB * MAXT * 2 * sizeof(short));
  size_t o_bn = off; off = align(off + 2 * NB * sizeof(int));
  size_t o_lu32 = off;
  size_t o_rc32 = o_lu32 + n * 5 * sizeof(float);
  size_t o_gs32 = o_rc32 + n * 4 * sizeof(float);
  off = align(o_gs32 + P * 16 * P * sizeof(float));
  const size_t nk = pl.circ ? 4 * P * P : 0;  // 2P x 2P carry rows
  size_t o_kf = off; off = align(off + nk * sizeof(double));
  size_t o_kb = off; off = align(off + nk * sizeof(double));
  size_t o_kf32 = off; off = align(off + nk * sizeof(float));
  size_t o_kb32 = off; off = align(off + nk * sizeof(float));
  std::vector<unsigned char> host(off, 0);
  std::memcpy(&host[o_cs], pl.cstart.data(), (P + 1) * sizeof(int));
  std::memcpy(&host[o_lu], pl.lu.data(), n * 5 * sizeof(double));
  std::memcpy(&host[o_rc], pl.rc.data(), n * 4 * sizeof(double));
  std::memcpy(&host[o_gs], pl.gsel.data(), P * 16 * P * sizeof(double));
  std::memcpy(&host[o_bd], pl.band.data(), n * 5 * sizeof(double));
  std::memcpy(&host[o_bc], pl.bcoef.data(), 2 * NB * MAXT * sizeof(double));
  std::memcpy(&host[o_bi], pl.bidx.data(), 2 * NB * MAXT * 2 * sizeof(short));
  std::memcpy(&host[o_bn], pl.bnt.data(), 2 * NB * sizeof(int));
  float* lu32 = reinterpret_cast<float*>(&host[o_lu32]);
  float* rc32 = reinterpret_cast<float*>(&host[o_rc32]);
  float* gs32 = reinterpret_cast<float*>(&host[o_gs32]);
  for (size_t i = 0; i < n * 5; ++i) lu32[i] = (float)pl.lu[i];
  for (size_t i = 0; i < n * 4; ++i) rc32[i] = (float)pl.rc[i];
  for (size_t i = 0; i < P * 16 * P; ++i) gs32[i] = (float)pl.gsel[i];
  if (nk) {
    std::memcpy(&host[o_kf], pl.kf.data(), nk * sizeof(double));
    std::memcpy(&host[o_kb], pl.kb.data(), nk * sizeof(double));
    float* kf32 = reinterpret_cast<float*>(&host[o_kf32]);
    float* kb32 = reinterpret_cast<float*>(&host[o_kb32]);
    for (size_t i = 0; i < nk; ++i) {
      kf32[i] = (float)pl.kf[i];
      kb32[i] = (float)pl.kb[i];
    }
  }
  HIPCHECK(hipMalloc(&d.mem, off));
  HIPCHECK(hipMemcpy(d.mem, host.data(), off, hipMemcpyHostToDevice));
  unsigned char* b = static_cast<unsigned char*>(d.mem);
  d.cstart = reinterpret_cast<const int*>(b + o_cs);
  d.lu = reinterpret_cast<const double*>(b + o_lu);
  d.rc = reinterpret_cast<const double*>(b + o_rc);
  d.gsel = reinterpret_cast<const double*>(b + o_gs);
  d.band = reinterpret_cast<const double*>(b + o_bd);
  d.bcoef = reinterpret_cast<const double*>(b + o_bc);
  d.bidx = reinterpret_cast<const short*>(b + o_bi);
  d.bnt = reinterpret_cast<const int*>(b + o_bn);
  d.lu32 = reinterpret_cast<const float*>(b + o_lu32);
  d.rc32 = reinterpret_cast<const float*>(b + o_rc32);
  d.gsel32 = reinterpret_cast<const float*>(b + o_gs32);
  d.ntab = (int)(n * 9 + P * 16 * P);
  if (nk) {
    d.kf = reinterpret_cast<const double*>(b + o_kf);
    d.kb = reinterpret_cast<const double*>(b + o_kb);
    d.kf32 = reinterpret_cast<const float*>(b + o_kf32);
    d.kb32 = reinterpret_cast<const float*>(b + o_kb32);
  }
  return 0;
}

// host pointers of the same tables (host emulation)
static void host_tables(const OpPlan& pl, DevPlan& d, std::vector<float>& lu32,
                        std::vector<float>& rc32, std::vector<float>& gs32,
                        std::vector<float>& kf32, std::vector<float>& kb32) {
  lu32.assign(pl.lu.begin(), pl.lu.end());
  rc32.assign(pl.rc.begin(), pl.rc.end());
  gs32.assign(pl.gsel.begin(), pl.gsel.end());
  kf32.assign(pl.kf.begin(), pl.kf.end());
  kb32.assign(pl.kb.begin(), pl.kb.end());
  d.kf = pl.kf.data();
  d.kb = pl.kb.data();
  d.kf32 = kf32.data();
  d.kb32 = kb32.data();
  d.cstart = pl.cstart.data();
  d.lu = pl.lu.data();
  d.rc = pl.rc.data();
  d.gsel = pl.gsel.data();
  d.band = pl.band.data();
  d.bcoef = pl.bcoef.data();
  d.bidx = pl.bidx.data();
  d.bnt = pl.bnt.data();
  d.lu32 = lu32.data();
  d.rc32 = rc32.data();
  d.gsel32 = gs32.data();
}

static LineArgs make_args(const pyr_ctx* c, const OpPlan& pl, const DevPlan& d, int dir, int epi,
                          double ec, int prec) {
  LineArgs a;
  std::memset(&a, 0, sizeof(a));
  const long long nx = c->n[0], ny = c->n[1], nz = c->n[2];
  a.N = pl.N;
  a.P = pl.P;
  a.periodic = pl.periodic;
  if (dir == 0) {
    a.nlines = (int)(ny * nz);
    a.A = 1; a.B = nx; a.stride = 1;
  } else if (dir == 1) {
    a.nlines = (int)(nx * nz);
    a.A = nx; a.B = nx * ny; a.stride = nx;
  } else {
    a.nlines = (int)(nx * ny);
    a.A = nx * ny; a.B = 0; a.stride = nx * ny;
  }
  a.nbytes = (int)(c->npts() * (prec == PYR_F64 ? sizeof(double) : sizeof(float)));
  a.addv = pl.addv;
  a.epi = epi;
  a.scale = pl.scale;
  a.epi_c = ec;
  for (int i = 0; i < 9; ++i) a.w[i] = pl.w[i];
  a.cstart = d.cstart; a.lu = d.lu; a.rc = d.rc; a.gsel = d.gsel; a.band = d.band;
  a.lu32 = d.lu32; a.rc32 = d.rc32; a.gsel32 = d.gsel32;
  a.tab64 = d.lu; a.tab32 = d.lu32; a.ntab = d.ntab;
  a.bcoef = d.bcoef; a.bidx = d.bidx; a.bnt = d.bnt;
  a.p1 = pl.p1; a.p2 = pl.p2; a.invc = pl.invc;
  a.p1f = (float)pl.p1; a.p2f = (float)pl.p2; a.invcf = (float)pl.invc;
  for (int i = 0; i < 3; ++i) a.bd[i] = pl.bd[i];
  a.kf = d.kf; a.kb = d.kb; a.kf32 = d.kf32; a.kb32 = d.kb32;
  return a;
}

// solve used for a plan: explicit, chunked LU, or circulant (periodic lines)
static int plan_solve(const pyr_ctx* c, const OpPlan& pl) {
  if (!pl.implicit) return SOLVE_NONE;
  if (pl.circ && c->periodic_solver == PYR_PSOLVE_CIRCULANT) return SOLVE_CIRC;
  return SOLVE_LU;
}

// ---------------------------------------------------------------------------
// launch
// ---------------------------------------------------------------------------
template <typename K>
static int set_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    HIPCHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return 0;
}

template <typename TS, int KIND, int SOLVE, int MODE, int LY, bool FULL>
static int launch_yz(pyr_ctx* c, int pre, dim3 grid, dim3 block, size_t shm, const LineArgs& a,
                     const TS* in, TS* out) {
  auto k = k_line_yz<TS, KIND, SOLVE, MODE, LY, FULL, PRE_NONE>;
#if PYR_PREFETCH
  // filters only exist for the r4 kind; accumulating epilogues for every kind
  if (pre == PRE_V) {
    if constexpr (KIND == K_R4) k = k_line_yz<TS, KIND, SOLVE, MODE, LY, FULL, PRE_V>;
    else return fail("filter epilogue on a non-filter kernel");
  } else if (pre == PRE_OLD) {
    k = k_line_yz<TS, KIND, SOLVE, MODE, LY, FULL, PRE_OLD>;
  }
#else
  if (pre != PRE_NONE) k = k_line_yz<TS, KIND, SOLVE, MODE, LY, FULL, PRE_LATE>;
#endif
  if (set_lds(k, shm)) return 1;
  hipLaunchKernelGGL(k, grid, block, shm, c->stream, a, in, out);
  return 0;
}

// Pipelined y/z launch when the line shape allows it (see k_line_yz_pipe);
// returns 1 if launched, 0 if the caller should use k_line_yz, -1 on error.
template <typename TS, int KIND, int SOLVE, int MODE>
static int launch_yz_pipe(pyr_ctx* c, int pre, const LineArgs& a, const TS* in, TS* out) {
  if constexpr (SOLVE != SOLVE_CIRC && SOLVE != SOLVE_NONE) {
    return 0;
  } else {
    using CS = typename SolveT<MODE>::type;
    constexpr int LG = 32;
    // Measured on gfx950 at 256^3: with FP32 fields the pipelined kernel wins
    // for the explicit filter (gbar y 54 -> 28 us) and the refined circulant
    // solve (d1 y 51 -> 41 us), but not with accumulating epilogues; FP64
    // fields leave too few waves per CU once two tiles fit in LDS.
    if constexpr (sizeof(TS) != 4 || (SOLVE == SOLVE_CIRC && MODE != MODE_R32)) {
      return 0;
    } else {
    if (pre == PRE_OLD) return 0;
    if (LG * a.P > WG_MAX) return 0;
    // whole waves (the copy loop strides by the wave count)
    if (!c->yz_pipe || !a.periodic || a.N % CHUNK != 0 || (LG * a.P) % 64 != 0) return 0;
    if (a.A % LG != 0 || a.nlines % LG != 0) return 0;
    if ((uintptr_t)in % 16 != 0 || (a.stride * sizeof(TS)) % 16 != 0 || (a.B * sizeof(TS)) % 16 != 0) return 0;
    constexpr int R = Halo<KIND>::R;
    constexpr int RPI = 64 / (LG * (int)sizeof(TS) / 16);
    // NT tile rows as in the kernel: N lines rows + 2R wrapped rows, in runs of RPI
    const size_t tile_rows = (size_t)((a.N + 2 * R + RPI - 1) / RPI) * RPI;
    const size_t tiles = align16(2 * tile_rows * LG * sizeof(TS));
    // carry exchange and carry rows only for the circulant solve
    const size_t shm = tiles + (SOLVE == SOLVE_CIRC ? (2 * (size_t)4 * a.P * LG + 2 * (size_t)4 * a.P * a.P) * sizeof(CS) : 0);
    if (shm > 80 * 1024) return 0;
    const int per_cu = (int)((160 * 1024) / shm);
    const int ngroups = a.nlines / LG;
    const int nblocks = std::min(ngroups, std::max(1, c->ncu * per_cu));
    auto k = k_line_yz_pipe<TS, KIND, SOLVE, MODE, LG, PRE_NONE>;
    if (pre == PRE_V) {
      if constexpr (KIND == K_R4) k = k_line_yz_pipe<TS, KIND, SOLVE, MODE, LG, PRE_V>;
      else return 0;
    }
    if (set_lds(k, shm)) return -1;
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(LG, a.P), shm, c->stream, a, in, out);
    return 1;
    }
  }
}

template <typename TS, int KIND, int SOLVE, int MODE, bool FULL>
static int launch_typed(pyr_ctx* c, int dir, LineArgs a, const TS* in, TS* out) {
  using CS = typename SolveT<MODE>::type;
  const int P = a.P;
  // LU tables go to LDS unless they are too large (long lines): then the
  // solve reads them from global memory (L2-resident, shared by all lines)
  const size_t stage_full = SOLVE == SOLVE_LU ? stage_bytes<MODE>(a.ntab, a.N) : 0;
  a.gtab = stage_full > STAGE_MAX_BYTES;
  const size_t stage = a.gtab ? 0 : stage_full;
  if (dir == 0) {
    constexpr int VW = Vec16<TS>::W;
    a.vec = (a.N % VW == 0) && ((uintptr_t)in % 16 == 0) && ((uintptr_t)out % 16 == 0);
    a.pitch = a.vec ? a.N + VW : a.N + 1;
    auto lds = [&](int lx) {
      const size_t tile = align16((size_t)lx * a.pitch * sizeof(TS));
      if (SOLVE == SOLVE_CIRC) return tile + 2 * align16((size_t)circ_words<MODE>() * P * lx * sizeof(CS));
      if (SOLVE == SOLVE_LU) {
        const size_t work = align16((size_t)4 * P * lx * sizeof(CS)) + stage;
        return tile > work ? tile : work;
      }
      return tile;
    };
    if constexpr (FULL && SOLVE == SOLVE_NONE) {
      // register-pipelined persistent kernel (yz_pipe 2 on x; explicit
      // filters only: with a solve the prefetched rows and the solve's
      // registers do not fit together without spills): one tile of LX lines
      // with 16-byte aligned rows, LX * P threads in whole waves
      const bool rok = c->yz_pipe == 2 && a.periodic && a.N % VW == 0 && ((uintptr_t)in % 16 == 0) &&
                       ((uintptr_t)out % 16 == 0) && (!a.out32 || (uintptr_t)a.out32 % 8 == 0);
      if (rok) {
        a.pitch = a.N + VW;
        a.vec = 1;
        int lx = WG_MAX / P;
        if (lx > 64) lx = 64;
        while (lx > 1 && lds(lx) > (size_t)c->x_lds_kb * 1024) lx /= 2;
        if ((lx * P) % 64 == 0) {
          a.lx = lx;
          const size_t shm = lds(lx);
          const int ngroups = (a.nlines + lx - 1) / lx;
          const int per_cu = std::max(1, (int)((160 * 1024) / shm));
          const int nblocks = std::min(ngroups, c->ncu * per_cu);
          auto k = k_line_x_rpipe<TS, KIND, SOLVE, MODE>;
          if (set_lds(k, shm)) return 1;
          hipLaunchKernelGGL(k, dim3(nblocks), dim3(lx * P), shm, c->stream, a, in, out);
          HIPCHECK(hipGetLastError());
          return 0;
        }
        a.pitch = a.vec ? a.N + VW : a.N + 1;
      }
    }
    if constexpr (FULL && (SOLVE == SOLVE_CIRC || SOLVE == SOLVE_NONE)) {
      // pipelined persistent kernel: two tiles of LX lines (16-byte aligned
      // rows, whole KiB per line), LX * P threads in whole waves
      const bool ok = c->yz_pipe == 1 && !a.out32 && a.periodic && (a.N * (int)sizeof(TS)) % 1024 == 0 &&
                      ((uintptr_t)in % 16 == 0) && ((uintptr_t)out % 16 == 0);
      if (ok) {
        const int pitch = a.N + VW;
        auto lds2 = [&](int lx) {
          size_t t = align16(2 * (size_t)lx * pitch * sizeof(TS));
          if (SOLVE == SOLVE_CIRC) t += 2 * align16((size_t)4 * P * lx * sizeof(CS));
          return t;
        };
        int lx = WG_MAX / P;
        if (lx > 64) lx = 64;
        while (lx > 1 && lds2(lx) > (size_t)c->x_lds_kb * 1024) lx /= 2;
        if ((lx * P) % 64 == 0) {
          a.lx = lx;
          a.pitch = pitch;
          a.vec = 1;
          const size_t shm = lds2(lx);
          const int ngroups = (a.nlines + lx - 1) / lx;
          const int per_cu = std::max(1, (int)((160 * 1024) / shm));
          const int nblocks = std::min(ngroups, c->ncu * per_cu);
          auto k = k_line_x_pipe<TS, KIND, SOLVE, MODE>;
          if (set_lds(k, shm)) return 1;
          hipLaunchKernelGGL(k, dim3(nblocks), dim3(lx * P), shm, c->stream, a, in, out);
          HIPCHECK(hipGetLastError());
          return 0;
        }
      }
    }
    int lx = WG_MAX / P;
    if (lx > 64) lx = 64;
    // keep the workgroup's LDS within the budget so several workgroups share a CU
    while (lx > 1 && lds(lx) > (size_t)c->x_lds_kb * 1024) lx /= 2;
    a.lx = lx;
    const size_t shm = lds(lx);
    const dim3 grid((a.nlines + lx - 1) / lx), block(lx * P);
    auto k = k_line_x<TS, KIND, SOLVE, MODE, FULL>;
    if (set_lds(k, shm)) return 1;
    hipLaunchKernelGGL(k, grid, block, shm, c->stream, a, in, out);
  } else {
    // lines per workgroup: 64 x P chunk threads within WG_MAX (32 or 16 lines
    // for longer lines); tuned: 32 or 16 lines (more, smaller workgroups per CU)
    const int ly = c->yz_lines == 16 ? 16
                 : ((P <= 8 && c->yz_lines != 32) ? 64 : (P <= WG_MAX / 32 ? 32 : 16));
    const size_t ex = align16((size_t)4 * P * ly * sizeof(CS));
    const size_t shm = SOLVE == SOLVE_CIRC ? 2 * ex : (SOLVE == SOLVE_LU ? ex + stage : 0);
    const dim3 grid((a.nlines + ly - 1) / ly), block(ly, P);
    const int pre = a.addv ? PRE_V : ((a.epi == EPI_ADD || a.epi == EPI_MAXABS) ? PRE_OLD : PRE_NONE);
    if constexpr (FULL) {
      // a filter with an accumulating epilogue keeps the plain kernel
      if (!(a.addv && pre == PRE_V && (a.epi == EPI_ADD || a.epi == EPI_MAXABS))) {
        const int r = launch_yz_pipe<TS, KIND, SOLVE, MODE>(c, pre, a, in, out);
        if (r < 0) return 1;
        if (r > 0) {
          HIPCHECK(hipGetLastError());
          return 0;
        }
      }
    }
    if (ly == 64) {
      if (launch_yz<TS, KIND, SOLVE, MODE, 64, FULL>(c, pre, grid, block, shm, a, in, out)) return 1;
    } else if (ly == 32) {
      if (launch_yz<TS, KIND, SOLVE, MODE, 32, FULL>(c, pre, grid, block, shm, a, in, out)) return 1;
    } else {
      if (launch_yz<TS, KIND, SOLVE, MODE, 16, FULL>(c, pre, grid, block, shm, a, in, out)) return 1;
    }
  }
  HIPCHECK(hipGetLastError());
  return 0;
}

template <typename TS, int MODE, bool FULL>
int launch_kind(pyr_ctx* c, int dir, int kind, int sv, const LineArgs& a, const TS* in, TS* out) {
  if constexpr (FULL && PYR_FULL_CHUNKS) {
    switch (kind) {
      case K_D1: return launch_typed<TS, K_D1, SOLVE_CIRC, MODE, true>(c, dir, a, in, out);
      case K_R3: return launch_typed<TS, K_R3, SOLVE_CIRC, MODE, true>(c, dir, a, in, out);
      default:
        if (sv == SOLVE_CIRC) return launch_typed<TS, K_R4, SOLVE_CIRC, MODE, true>(c, dir, a, in, out);
        return launch_typed<TS, K_R4, SOLVE_NONE, MODE, true>(c, dir, a, in, out);
    }
  } else if constexpr (!FULL) {
    switch (kind) {
      case K_D1:
        return sv == SOLVE_CIRC ? launch_typed<TS, K_D1, SOLVE_CIRC, MODE, false>(c, dir, a, in, out)
                                : launch_typed<TS, K_D1, SOLVE_LU, MODE, false>(c, dir, a, in, out);
      case K_R3:
        return sv == SOLVE_CIRC ? launch_typed<TS, K_R3, SOLVE_CIRC, MODE, false>(c, dir, a, in, out)
                                : launch_typed<TS, K_R3, SOLVE_LU, MODE, false>(c, dir, a, in, out);
      default:
        if (sv == SOLVE_CIRC) return launch_typed<TS, K_R4, SOLVE_CIRC, MODE, false>(c, dir, a, in, out);
        if (sv == SOLVE_LU) return launch_typed<TS, K_R4, SOLVE_LU, MODE, false>(c, dir, a, in, out);
        return launch_typed<TS, K_R4, SOLVE_NONE, MODE, false>(c, dir, a, in, out);
    }
  }
  return fail("full-chunk kernels are not built");
}

#define PYR_LAUNCH_KIND(TS, MODE, FULL) \
  template int launch_kind<TS, MODE, FULL>(pyr_ctx*, int, int, int, const LineArgs&, const TS*, TS*);
#if PYR_PART == 1
PYR_LAUNCH_KIND(double, MODE_F64, true)
PYR_LAUNCH_KIND(double, MODE_F64, false)
#elif PYR_PART == 2
PYR_LAUNCH_KIND(float, MODE_R32, true)
PYR_LAUNCH_KIND(float, MODE_R32, false)
#elif PYR_PART == 3
PYR_LAUNCH_KIND(float, MODE_S32, true)
PYR_LAUNCH_KIND(float, MODE_S32, false)
#else
extern PYR_LAUNCH_KIND(double, MODE_F64, true)
extern PYR_LAUNCH_KIND(double, MODE_F64, false)
extern PYR_LAUNCH_KIND(float, MODE_R32, true)
extern PYR_LAUNCH_KIND(float, MODE_R32, false)
extern PYR_LAUNCH_KIND(float, MODE_S32, true)
extern PYR_LAUNCH_KIND(float, MODE_S32, false)
#endif
#undef PYR_LAUNCH_KIND

#if PYR_PART == 0

template <typename TS, int MODE>
static int launch_op(pyr_ctx* c, int op, int dir, const TS* in, TS* out, int epi, double ec, bool anti) {
  // anti: the antisymmetric member where the direction has a SYMM end, as
  // d1x/d8x(v, dv, bc=-1) select it (compact_operators.f90:300-303)
  const bool ua = anti && c->has_anti[dir];
  const OpPlan& pl = ua ? c->plan_a[op][dir] : c->plan[op][dir];
  if (pl.null_op) {
    const long long n = (long long)c->npts();
    float* o32 = sizeof(TS) == 8 ? c->out32 : nullptr;
    hipLaunchKernelGGL((k_null<TS>), dim3((n + 255) / 256), dim3(256), 0, c->stream, n,
                       pl.null_option == 1 ? 1 : 0, epi, ec, in, (o32 && c->only32) ? nullptr : out, o32);
    HIPCHECK(hipGetLastError());
    return 0;
  }
  LineArgs a = make_args(c, pl, ua ? c->dplan_a[op][dir] : c->dplan[op][dir], dir, epi, ec,
                         MODE == MODE_F64 ? PYR_F64 : PYR_R32);
  if (c->out32) {
    // the plain kernels of FP64 fields with the store epilogue write the
    // rounded copy
    if (sizeof(TS) != 8 || epi != EPI_STORE) return fail("an FP32 copy of the output needs an FP64 sweep that stores");
    a.out32 = c->out32;
    a.only32 = c->only32;
  }
  const int sv = plan_solve(c, pl);
  // the balanced partition gives every chunk CHUNK rows exactly when CHUNK | N;
  // the LU solve keeps runtime row counts (with compile-time counts the
  // compiler hoists every table read of the chunk and runs out of registers)
  const int sx = c->x_lds_kb, sp = c->yz_pipe, sl = c->yz_lines;
  if (c->tune_xlds[op][dir][MODE] >= 0) c->x_lds_kb = c->tune_xlds[op][dir][MODE];
  if (c->tune_pipe[op][dir][MODE] >= 0) c->yz_pipe = c->tune_pipe[op][dir][MODE];
  else if (dir == 0) c->yz_pipe = 0;  // the pipelined x kernel only where tuning chose it
  if (c->tune_ly[op][dir][MODE] >= 0) c->yz_lines = c->tune_ly[op][dir][MODE];
  // yz_pipe 3 on y/z: the plain kernel with runtime row counts even when
  // CHUNK | N (fewer registers, more waves per SIMD; the same results)
  const bool full = PYR_FULL_CHUNKS && pl.N % CHUNK == 0 && sv != SOLVE_LU && !(dir != 0 && c->yz_pipe == 3);
  const int r = full
                    ? launch_kind<TS, MODE, true>(c, dir, pl.kind, sv, a, in, out)
                    : launch_kind<TS, MODE, false>(c, dir, pl.kind, sv, a, in, out);
  c->x_lds_kb = sx;
  c->yz_pipe = sp;
  c->yz_lines = sl;
  return r;
}

static int check_ctx(pyr_ctx* c) {
  if (!c) return fail("null context");
  if (!c->device) return fail("context was set up host-only (pyr_setup_host)");
  return 0;
}

// FP32 solve arithmetic of one operator and direction: its tuned choice, else
// the context's
static int r32_mode(const pyr_ctx* c, int op, int dir) {
  return c->tune_solve[op][dir] >= 0 ? c->tune_solve[op][dir] : c->r32_solve;
}

static int apply_any(pyr_ctx* c, int op, int dir, int prec, const void* in, void* out, int epi,
                     double ec, bool anti = false) {
  if (check_ctx(c)) return 1;
  if (op < 0 || op >= NOPS || dir < 0 || dir > 2) return fail("bad operator or direction");
  // (an FP64 sweep that stores only the FP32 rounding has no FP64 output)
  const bool only32 = c->out32 && c->only32;
  if (!in || (!out && !only32)) return fail("null field pointer");
  if (in == out || (void*)in == (void*)c->out32) return fail("input and output must not alias");
  if (only32 && prec != PYR_F64) return fail("an FP32-only result needs an FP64 sweep");
  if (prec == PYR_F64)
    return launch_op<double, MODE_F64>(c, op, dir, (const double*)in, (double*)out, epi, ec, anti);
  if (prec == PYR_R32) {
    if (r32_mode(c, op, dir) == PYR_R32_SOLVE_FP64)
      return launch_op<float, MODE_S32>(c, op, dir, (const float*)in, (float*)out, epi, ec, anti);
    return launch_op<float, MODE_R32>(c, op, dir, (const float*)in, (float*)out, epi, ec, anti);
  }
  return fail("bad precision");
}

static int field_temp(pyr_ctx* c, int prec, void** p) {
  const size_t bytes = c->npts() * (prec == PYR_F64 ? sizeof(double) : sizeof(float));
  if (!c->field_tmp[prec]) HIPCHECK(hipMalloc(&c->field_tmp[prec], bytes));
  *p = c->field_tmp[prec];
  return 0;
}

// filter x -> y -> z (operators.f90:847-875) through the field temp `tmp`
static int filter3_with(pyr_ctx* c, int op, int prec, const void* in, void* out, void* tmp) {
  if (apply_any(c, op, 0, prec, in, out, EPI_STORE, 0.0)) return 1;
  if (apply_any(c, op, 1, prec, out, tmp, EPI_STORE, 0.0)) return 1;
  return apply_any(c, op, 2, prec, tmp, out, EPI_STORE, 0.0);
}

// filter x -> y -> z (operators.f90:847-875)
static int filter3(pyr_ctx* c, int op, int prec, const void* in, void* out) {
  void* tmp;
  if (check_ctx(c) || field_temp(c, prec, &tmp)) return 1;
  return filter3_with(c, op, prec, in, out, tmp);
}

// ring L=2 (operators.f90:615-643): max_d |d8_d f| * d_d^2, fused into the stores
static int ring(pyr_ctx* c, int prec, const void* in, void* out) {
  for (int d = 0; d < 3; ++d) {
    const double s = c->dmesh[d] * c->dmesh[d];
    if (apply_any(c, OP_D8, d, prec, in, out, d == 0 ? EPI_ABS : EPI_MAXABS, s)) return 1;
  }
  return 0;
}

// sum over directions of op_d(f_d); anti: each f_d is the normal component of
// direction d and takes the antisymmetric member at SYMM ends (divV,
// operators.f90:38-53)
static int sum3(pyr_ctx* c, int op, int prec, const void* f0, const void* f1, const void* f2,
                void* out, bool anti = false) {
  const void* f[3] = {f0, f1, f2};
  for (int d = 0; d < 3; ++d)
    if (apply_any(c, op, d, prec, f[d], out, d == 0 ? EPI_STORE : EPI_ADD, 0.0, anti)) return 1;
  return 0;
}

// ---------------------------------------------------------------------------
// device memory of the entry point that allocates nothing
// ---------------------------------------------------------------------------
// pyr_sfilter_noalloc keeps the field temporary of the x -> y -> z filter in
// this file-scope array (one copy per GPU, resident until the process ends)
// instead of allocating it. Host code never passes `scratch` itself to a
// kernel: its device address on each GPU comes from hipGetSymbolAddress at
// the first pyr_setup on that GPU (noalloc_init) and is kept per device.
constexpr size_t PYR_NOALLOC_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[PYR_NOALLOC_BYTES];
constexpr int PYR_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[PYR_MAX_DEVICES] = {};

static int noalloc_init(int dev, void** ptr) {
  if (dev < 0 || dev >= PYR_MAX_DEVICES) return fail("device ordinal out of range for the filter scratch");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_noalloc_ptr[dev] = p;
  }
  *ptr = g_noalloc_ptr[dev];
  return 0;
}

// ---------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------
static int setup_common(pyr_ctx** out, int nx, int ny, int nz, double x1, double xn, double y1,
                        double yn, double z1, double zn, const int* bc6, bool device) {
  if (!out) return fail("null output handle");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1) return fail("grid sizes must be >= 1");
  // the y/z kernels address a field through a 32-bit buffer range
  if ((double)nx * ny * nz * sizeof(double) > 2147483647.0) return fail("field exceeds 2 GiB");
  if (!bc6) return fail("null bc array");
  std::unique_ptr<pyr_ctx> c(new pyr_ctx);
  c->n[0] = nx; c->n[1] = ny; c->n[2] = nz;
  if (const char* e = std::getenv("PYRANDA_AMD_YZ_PIPE")) c->yz_pipe = std::atoi(e) != 0;
  if (const char* e = std::getenv("PYRANDA_AMD_X_LDS_KB")) {
    const int kb = std::atoi(e);
    if (kb >= 8 && kb <= 160) c->x_lds_kb = kb;
  }
  for (int i = 0; i < 6; ++i) {
    if (bc6[i] < PYR_BC_PERI || bc6[i] > PYR_BC_SYMM) return fail("bad boundary code");
    c->bc[i] = bc6[i];
  }
  const double lo[3] = {x1, y1, z1}, hi[3] = {xn, yn, zn};
  for (int d = 0; d < 3; ++d) {
    // faces from nodes (parcop.f90:46-56), patch spacing (patch.f90:83-85)
    const double du = (hi[d] - lo[d]) / (double)(c->n[d] - 1 > 1 ? c->n[d] - 1 : 1);
    const double f1 = lo[d] - du / 2.0, fn = hi[d] + du / 2.0;
    c->x1[d] = f1;
    c->dpatch[d] = (fn - f1) / (double)c->n[d];
  }
  // mesh length scales (mesh.f90:190-211)
  c->dmesh[0] = nx == 1 ? std::fmax(c->dpatch[1], c->dpatch[2]) : c->dpatch[0];
  c->dmesh[1] = ny == 1 ? std::fmax(c->dpatch[0], c->dpatch[2]) : c->dpatch[1];
  c->dmesh[2] = nz == 1 ? std::fmax(c->dpatch[0], c->dpatch[1]) : c->dpatch[2];
  c->gridlen = std::fmin(c->dmesh[0], std::fmin(c->dmesh[1], c->dmesh[2]));
  try {
    for (int d = 0; d < 3; ++d) {
      const int periodic = c->bc[2 * d] == PYR_BC_PERI;
      const int v1 = c->bc[2 * d] == PYR_BC_SYMM ? 1 : 0;
      const int v2 = c->bc[2 * d + 1] == PYR_BC_SYMM ? 1 : 0;
      for (int op = 0; op < NOPS; ++op)
        c->plan[op][d] = build_plan(op, c->n[d], periodic, v1, v2, c->dpatch[d]);
      c->has_anti[d] = !periodic && (v1 || v2) && c->n[d] >= 4;
      if (c->has_anti[d])
        for (int op = 0; op < NOPS; ++op)
          c->plan_a[op][d] = build_plan(op, c->n[d], periodic, -v1, -v2, c->dpatch[d]);
    }
  } catch (const std::exception& e) {
    return fail(std::string("setup: ") + e.what());
  }
  if (device) {
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
      c->ncu = prop.multiProcessorCount;
    // once per GPU, with that GPU current: its copy of the scratch array
    if (noalloc_init(dev, &c->noalloc_tmp)) return 1;
    for (int op = 0; op < NOPS; ++op)
      for (int d = 0; d < 3; ++d) {
        if (upload_plan(c->plan[op][d], c->dplan[op][d])) return 1;
        if (c->has_anti[d] && upload_plan(c->plan_a[op][d], c->dplan_a[op][d])) return 1;
      }
    c->device = true;
  }
  *out = c.release();
  return 0;
}

// ---------------------------------------------------------------------------
// host emulation of the kernels (tests only): the same per-thread arithmetic,
// chunk by chunk, with the LDS exchanges as plain arrays
// ---------------------------------------------------------------------------
template <typename TS, int KIND, int SOLVE, int MODE>
static void emulate_line(const LineArgs& a, const TS* vin, long long stride, TS* vout) {
  constexpr int R = Halo<KIND>::R;
  using CS = typename SolveT<MODE>::type;
  const int P = a.P, N = a.N;
  struct St {
    TS win[CHUNK + 2 * 4];
    double rhs[CHUNK];
    CS x[CHUNK], d[CHUNK], z[4], zd[4];
    int s, m;
  };
  std::vector<St> st(P);
  std::vector<CS> Y(4 * P), Yb(4 * P);
  auto rd = [&](int i) { return vin[stride * i]; };
  for (int p = 0; p < P; ++p) {
    St& t = st[p];
    t.s = a.cstart[p];
    t.m = a.cstart[p + 1] - t.s;
    for (int w = 0; w < CHUNK + 2 * R; ++w) {
      int row = t.s - R + w;
      TS v = (TS)0;
      if (w < t.m + 2 * R) {
        if (a.periodic) {
          row = row < 0 ? row + N : (row >= N ? row - N : row);
          v = rd(row);
        } else if (row >= 0 && row < N) {
          v = rd(row);
        }
      }
      t.win[w] = v;
    }
    rhs_interior<KIND>(a, t.win, t.rhs, t.m);
    rhs_boundary(a, t.s, t.m, t.rhs, rd);
  }
  // circulant solve of every chunk's vector `v` (without 1/c), as circ_solve
  auto circ_all = [&](auto getv, const CS* kf, const CS* kb, CS p1, CS p2) {
    for (int p = 0; p < P; ++p) circ_fwd_exit(getv(p), st[p].m, p1, p2, Y[2 * p], Y[2 * p + 1]);
    for (int p = 0; p < P; ++p) {
      CS e0, e1;
      circ_carry(kf, p, P, [&](int cc) { return Y[cc]; }, e0, e1);
      circ_fwd(getv(p), st[p].m, p1, p2, e0, e1);
      circ_bwd_exit(getv(p), st[p].m, p1, p2, Yb[2 * p], Yb[2 * p + 1]);
    }
    for (int p = 0; p < P; ++p) {
      CS e0, e1;
      circ_carry(kb, p, P, [&](int cc) { return Yb[cc]; }, e0, e1);
      circ_bwd(getv(p), st[p].m, p1, p2, e0, e1);
    }
  };
  if (SOLVE == SOLVE_LU) {
    const CS* G = MODE != MODE_R32 ? (const CS*)a.gsel : (const CS*)a.gsel32;
    const CS* RC = MODE != MODE_R32 ? (const CS*)a.rc : (const CS*)a.rc32;
    const CS* LU = MODE != MODE_R32 ? (const CS*)a.lu : (const CS*)a.lu32;
    for (int p = 0; p < P; ++p) {
      St& t = st[p];
      for (int k = 0; k < CHUNK; ++k) t.x[k] = (CS)t.rhs[k];
      local_solve<CS>(LU, t.s, t.m, t.x);
      CS e4[4];
      edges(t.x, t.m, e4);
      for (int q = 0; q < 4; ++q) Y[4 * p + q] = e4[q];
    }
    for (int p = 0; p < P; ++p) {
      St& t = st[p];
      neighbours<CS>(G, p, P, [&](int cc) { return Y[cc]; }, t.z);
      spike_correct(RC, t.s, t.m, t.z, t.x);
    }
    if (MODE == MODE_R32) {
      std::vector<CS> Yd(4 * P);
      for (int p = 0; p < P; ++p) {
        St& t = st[p];
        double res[CHUNK];
        residual64([&](int k) { return a.band + 5 * (t.s + k); }, t.m, t.rhs, (const float*)t.x,
                   (const float*)t.z, res);
        for (int k = 0; k < CHUNK; ++k) t.d[k] = (CS)res[k];
        local_solve<CS>(LU, t.s, t.m, t.d);
        CS f4[4];
        edges(t.d, t.m, f4);
        for (int q = 0; q < 4; ++q) Yd[4 * p + q] = f4[q];
      }
      for (int p = 0; p < P; ++p) {
        St& t = st[p];
        neighbours<CS>(G, p, P, [&](int cc) { return Yd[cc]; }, t.zd);
        spike_correct(RC, t.s, t.m, t.zd, t.d);
      }
    }
  } else if (SOLVE == SOLVE_CIRC) {
    for (int p = 0; p < P; ++p)
      for (int k = 0; k < CHUNK; ++k) st[p].x[k] = (CS)st[p].rhs[k];
    auto gx = [&](int p) { return st[p].x; };
    if (MODE != MODE_R32) {
      circ_all(gx, (const CS*)a.kf, (const CS*)a.kb, (CS)a.p1, (CS)a.p2);
      for (int p = 0; p < P; ++p)
        for (int k = 0; k < CHUNK; ++k) st[p].x[k] *= (CS)a.invc;
    } else {
      const CS *kf = (const CS*)a.kf32, *kb = (const CS*)a.kb32;
      const CS p1 = (CS)a.p1f, p2 = (CS)a.p2f, ic = (CS)a.invcf;
      circ_all(gx, kf, kb, p1, p2);
      std::vector<CS> E(4 * P);
      for (int p = 0; p < P; ++p) {
        for (int k = 0; k < CHUNK; ++k) st[p].x[k] *= ic;
        edges(st[p].x, st[p].m, &E[4 * p]);
      }
      const double row[5] = {a.bd[0], a.bd[1], a.bd[2], a.bd[1], a.bd[0]};
      for (int p = 0; p < P; ++p) {
        St& t = st[p];
        const int pm = (p - 1 + P) % P, pp = (p + 1) % P;
        const float z[4] = {(float)E[4 * pm + 2], (float)E[4 * pm + 3], (float)E[4 * pp], (float)E[4 * pp + 1]};
        double res[CHUNK];
        residual64([&](int) { return row; }, t.m, t.rhs, (const float*)t.x, z, res);
        for (int k = 0; k < CHUNK; ++k) t.d[k] = (CS)res[k];
      }
      circ_all([&](int p) { return st[p].d; }, kf, kb, p1, p2);
      for (int p = 0; p < P; ++p)
        for (int k = 0; k < CHUNK; ++k) st[p].d[k] *= ic;
    }
  }
  for (int p = 0; p < P; ++p) {
    St& t = st[p];
    for (int k = 0; k < t.m; ++k) {
      double y;
      if (SOLVE == SOLVE_NONE) y = t.rhs[k];
      else if (MODE != MODE_R32) y = (double)t.x[k];
      else y = (double)t.x[k] + (double)t.d[k];
      double val = y * a.scale;
      if (a.addv) val += (double)t.win[k + R];
      vout[stride * (t.s + k)] = (TS)val;
    }
  }
}

template <typename TS, int MODE>
static int emulate(pyr_ctx* c, int op, int dir, const TS* in, TS* out) {
  const OpPlan& pl = c->plan[op][dir];
  const size_t n = c->npts();
  if (pl.null_op) {
    for (size_t i = 0; i < n; ++i) out[i] = pl.null_option == 1 ? in[i] : (TS)0;
    return 0;
  }
  DevPlan d;
  std::vector<float> lu32, rc32, gs32, kf32, kb32;
  host_tables(pl, d, lu32, rc32, gs32, kf32, kb32);
  const LineArgs a = make_args(c, pl, d, dir, EPI_STORE, 0.0, MODE == MODE_F64 ? PYR_F64 : PYR_R32);
  const int sv = plan_solve(c, pl);
  for (long long L = 0; L < a.nlines; ++L) {
    const long long base = (dir == 0) ? L * pl.N : (L % a.A) + a.B * (L / a.A);
    const TS* vi = in + base;
    TS* vo = out + base;
    switch (pl.kind) {
      case K_D1:
        if (sv == SOLVE_CIRC) emulate_line<TS, K_D1, SOLVE_CIRC, MODE>(a, vi, a.stride, vo);
        else emulate_line<TS, K_D1, SOLVE_LU, MODE>(a, vi, a.stride, vo);
        break;
      case K_R3:
        if (sv == SOLVE_CIRC) emulate_line<TS, K_R3, SOLVE_CIRC, MODE>(a, vi, a.stride, vo);
        else emulate_line<TS, K_R3, SOLVE_LU, MODE>(a, vi, a.stride, vo);
        break;
      default:
        if (sv == SOLVE_CIRC) emulate_line<TS, K_R4, SOLVE_CIRC, MODE>(a, vi, a.stride, vo);
        else if (sv == SOLVE_LU) emulate_line<TS, K_R4, SOLVE_LU, MODE>(a, vi, a.stride, vo);
        else emulate_line<TS, K_R4, SOLVE_NONE, MODE>(a, vi, a.stride, vo);
    }
  }
  return 0;
}

#endif  // PYR_PART == 0

// ---------------------------------------------------------------------------
// groups: one operator over the fields of several simulations
// ---------------------------------------------------------------------------
// Members of a group are contexts of any grid sizes whose fields live in one
// allocation (member m at off[m] elements). An operator along one direction
// runs as ONE grouped launch over every member (workgroups shaped for the
// member with the most chunks per line; threads without a chunk only take
// the barriers), plus a k_null launch for a member whose direction is a null
// operator. (Splitting the members by chunk count into separate launches
// idles fewer threads but measured slower: the launches are latency-bound
// and run one after another.) The per-member line arguments of a
// launch are built once per (operator, direction, precision, epilogue) and
// kept on the device.
struct GrpLaunch {
  GrpEntry* dev = nullptr;
  int M = 0;
  long long nblocks = 0;
  dim3 block;
  size_t shm = 0;
  int kind = 0, solve = 0, ly = 0, pre = 0;
  bool x = false;
};
struct GrpNull {
  int member;
  int copy;
  double ec;
};
struct GrpPlan {
  std::vector<GrpLaunch> launches;
  std::vector<GrpNull> nulls;
};
// A plan depends on the solve arithmetic (carry types and LDS sizes) and on
// the members' launch knobs at build time; all of them are part of the key so
// that a change of tuning builds a new plan instead of reusing a stale one.
struct GrpKey {
  int op, dir, prec, epi, power, aligned;
  int mode, psolve, xlds;
  double ec;
  bool operator<(const GrpKey& o) const {
    if (op != o.op) return op < o.op;
    if (dir != o.dir) return dir < o.dir;
    if (prec != o.prec) return prec < o.prec;
    if (mode != o.mode) return mode < o.mode;
    if (psolve != o.psolve) return psolve < o.psolve;
    if (xlds != o.xlds) return xlds < o.xlds;
    if (epi != o.epi) return epi < o.epi;
    if (power != o.power) return power < o.power;
    if (aligned != o.aligned) return aligned < o.aligned;
    return ec < o.ec;
  }
};
struct pyr_group {
  std::vector<pyr_ctx*> mem;
  std::vector<long long> off;
  hipStream_t stream = nullptr;
  std::map<GrpKey, GrpPlan> plans;
  ~pyr_group() {
    for (auto& kv : plans)
      for (auto& l : kv.second.launches)
        if (l.dev) (void)hipFree(l.dev);
  }
};

#if PYR_PART == 0
template <typename TS, int MODE>
static int grp_build(pyr_group* g, const GrpKey& key, GrpPlan& gp) {
  using CS = typename SolveT<MODE>::type;
  constexpr int VW = Vec16<TS>::W;
  const int M = (int)g->mem.size();
  std::vector<int> ms;
  for (int m = 0; m < M; ++m) {
    pyr_ctx* c = g->mem[m];
    const OpPlan& pl = c->plan[key.op][key.dir];
    const double ec = key.ec * std::pow(c->dmesh[key.dir], (double)key.power);
    if (pl.null_op) gp.nulls.push_back({m, pl.null_option == 1 ? 1 : 0, ec});
    else ms.push_back(m);
  }
  if (ms.empty()) return 0;
  const OpPlan& pl0 = g->mem[ms[0]]->plan[key.op][key.dir];
  const int sv = plan_solve(g->mem[ms[0]], pl0);
  std::vector<GrpEntry> all(ms.size());
  for (size_t j = 0; j < ms.size(); ++j) {
    pyr_ctx* c = g->mem[ms[j]];
    const OpPlan& pl = c->plan[key.op][key.dir];
    if (plan_solve(c, pl) != sv || pl.kind != pl0.kind)
      return fail("group members differ in boundary conditions along a direction");
    const double ec = key.ec * std::pow(c->dmesh[key.dir], (double)key.power);
    all[j].a = make_args(c, pl, c->dplan[key.op][key.dir], key.dir, key.epi, ec, key.prec);
    // bounded lines: every member's LU tables are read from global memory
    // (members differ in length; the workgroup's LDS holds the exchange only)
    all[j].a.gtab = sv == SOLVE_LU;
    all[j].off = g->off[ms[j]];
  }
  auto upload = [&](GrpLaunch& L, std::vector<GrpEntry>& host) -> int {
    L.kind = pl0.kind;
    L.solve = sv;
    L.M = (int)host.size();
    HIPCHECK(hipMalloc(&L.dev, host.size() * sizeof(GrpEntry)));
    HIPCHECK(hipMemcpy(L.dev, host.data(), host.size() * sizeof(GrpEntry), hipMemcpyHostToDevice));
    gp.launches.push_back(L);
    return 0;
  };
  if (key.dir == 0) {
    // each member keeps its own lines per workgroup (its LDS within the
    // budget), so lx * P stays near WG_MAX for every member; the launch
    // takes the most threads and LDS of any member
    GrpLaunch L;
    L.x = true;
    int nthr = 0;
    size_t shm = 0;
    long long blk = 0;
    for (auto& e : all) {
      const int P = e.a.P;
      e.a.vec = key.aligned && (e.a.N % VW == 0) && ((e.off * (long long)sizeof(TS)) % 16 == 0);
      e.a.pitch = e.a.vec ? e.a.N + VW : e.a.N + 1;
      auto lds = [&](int lx) {
        const size_t tile = align16((size_t)lx * e.a.pitch * sizeof(TS));
        if (sv == SOLVE_CIRC) return tile + 2 * align16((size_t)4 * P * lx * sizeof(CS));
        if (sv == SOLVE_LU) return std::max(tile, align16((size_t)4 * P * lx * sizeof(CS)));
        return tile;
      };
      int lx = WG_MAX / P;
      if (lx > 64) lx = 64;
      while (lx > 1 && lds(lx) > (size_t)g->mem[0]->x_lds_kb * 1024) lx /= 2;
      e.a.lx = lx;
      e.blk0 = blk;
      blk += (e.a.nlines + lx - 1) / lx;
      nthr = std::max(nthr, lx * P);
      shm = std::max(shm, lds(lx));
    }
    L.block = dim3(nthr);
    L.shm = shm;
    L.nblocks = blk;
    return upload(L, all);
  }
  // y/z: workgroups are (ly lanes, P chunks) shaped for the longest member of
  // the launch; members are split into launches by chunk count (<= 8, <= 16,
  // <= 32) so that short lines do not run in workgroups shaped for 1024 points
  const int pre = all[0].a.addv ? PRE_V : ((key.epi == EPI_ADD || key.epi == EPI_MAXABS) ? PRE_OLD : PRE_NONE);
  const int bounds[3] = {8, 16, 32};
  for (int cl = 0; cl < 3; ++cl) {
    std::vector<GrpEntry> host;
    for (auto& e : all)
      if (e.a.P <= bounds[cl] && (cl == 0 || e.a.P > bounds[cl - 1])) host.push_back(e);
    if (host.empty()) continue;
    int Pmax = 1;
    for (auto& e : host) Pmax = std::max(Pmax, e.a.P);
    const int ly = Pmax <= 8 ? 64 : (Pmax <= WG_MAX / 32 ? 32 : 16);
    long long blk = 0;
    for (auto& e : host) {
      e.blk0 = blk;
      blk += (e.a.nlines + ly - 1) / ly;
    }
    GrpLaunch L;
    L.x = false;
    L.ly = ly;
    L.block = dim3(ly, Pmax);
    const size_t ex = align16((size_t)4 * Pmax * ly * sizeof(CS));
    L.shm = sv == SOLVE_CIRC ? 2 * ex : (sv == SOLVE_LU ? ex : 0);
    L.pre = pre;
    L.nblocks = blk;
    if (upload(L, host)) return 1;
  }
  return 0;
}

#endif  // PYR_PART == 0

template <typename TS, int KIND, int SOLVE, int MODE>
static int grp_launch_typed(pyr_group* g, const GrpLaunch& L, const TS* in, TS* out) {
  if (L.nblocks <= 0) return 0;
  const dim3 grid((unsigned)L.nblocks);
  if (L.x) {
    auto k = k_line_x_grp<TS, KIND, SOLVE, MODE>;
    if (set_lds(k, L.shm)) return 1;
    hipLaunchKernelGGL(k, grid, L.block, L.shm, g->stream, L.dev, L.M, in, out);
  } else {
    auto pick = [&](auto k64, auto k32, auto k16) {
      return L.ly == 64 ? k64 : (L.ly == 32 ? k32 : k16);
    };
    auto k = pick(k_line_yz_grp<TS, KIND, SOLVE, MODE, 64, PRE_NONE>, k_line_yz_grp<TS, KIND, SOLVE, MODE, 32, PRE_NONE>,
                  k_line_yz_grp<TS, KIND, SOLVE, MODE, 16, PRE_NONE>);
    if (L.pre == PRE_OLD)
      k = pick(k_line_yz_grp<TS, KIND, SOLVE, MODE, 64, PRE_OLD>, k_line_yz_grp<TS, KIND, SOLVE, MODE, 32, PRE_OLD>,
               k_line_yz_grp<TS, KIND, SOLVE, MODE, 16, PRE_OLD>);
    if (L.pre == PRE_V) {
      if constexpr (KIND == K_R4)
        k = pick(k_line_yz_grp<TS, KIND, SOLVE, MODE, 64, PRE_V>, k_line_yz_grp<TS, KIND, SOLVE, MODE, 32, PRE_V>,
                 k_line_yz_grp<TS, KIND, SOLVE, MODE, 16, PRE_V>);
      else
        return fail("filter epilogue on a non-filter kernel");
    }
    if (set_lds(k, L.shm)) return 1;
    hipLaunchKernelGGL(k, grid, L.block, L.shm, g->stream, L.dev, L.M, in, out);
  }
  HIPCHECK(hipGetLastError());
  return 0;
}

template <typename TS, int MODE>
int grp_launch_mode(pyr_group* g, const GrpLaunch& L, const TS* in, TS* out) {
  switch (L.kind) {
    case K_D1:
      return L.solve == SOLVE_CIRC ? grp_launch_typed<TS, K_D1, SOLVE_CIRC, MODE>(g, L, in, out)
                                   : grp_launch_typed<TS, K_D1, SOLVE_LU, MODE>(g, L, in, out);
    case K_R3:
      return L.solve == SOLVE_CIRC ? grp_launch_typed<TS, K_R3, SOLVE_CIRC, MODE>(g, L, in, out)
                                   : grp_launch_typed<TS, K_R3, SOLVE_LU, MODE>(g, L, in, out);
    default:
      if (L.solve == SOLVE_CIRC) return grp_launch_typed<TS, K_R4, SOLVE_CIRC, MODE>(g, L, in, out);
      if (L.solve == SOLVE_LU) return grp_launch_typed<TS, K_R4, SOLVE_LU, MODE>(g, L, in, out);
      return grp_launch_typed<TS, K_R4, SOLVE_NONE, MODE>(g, L, in, out);
  }
}

#define PYR_GRP_MODE(TS, MODE) template int grp_launch_mode<TS, MODE>(pyr_group*, const GrpLaunch&, const TS*, TS*);
#if PYR_PART == 1
PYR_GRP_MODE(double, MODE_F64)
#elif PYR_PART == 2
PYR_GRP_MODE(float, MODE_R32)
#elif PYR_PART == 3
PYR_GRP_MODE(float, MODE_S32)
#else
extern PYR_GRP_MODE(double, MODE_F64)
extern PYR_GRP_MODE(float, MODE_R32)
extern PYR_GRP_MODE(float, MODE_S32)
#endif
#undef PYR_GRP_MODE

#if PYR_PART == 0
template <typename TS, int MODE>
static int grp_apply(pyr_group* g, const GrpKey& key, const TS* in, TS* out) {
  auto it = g->plans.find(key);
  if (it == g->plans.end()) {
    GrpPlan gp;
    if (grp_build<TS, MODE>(g, key, gp)) {
      for (auto& l : gp.launches)
        if (l.dev) (void)hipFree(l.dev);
      return 1;
    }
    it = g->plans.emplace(key, std::move(gp)).first;
  }
  for (const GrpLaunch& L : it->second.launches)
    if (grp_launch_mode<TS, MODE>(g, L, in, out)) return 1;
  for (const GrpNull& z : it->second.nulls) {
    const long long n = (long long)g->mem[z.member]->npts();
    const long long o = g->off[z.member];
    hipLaunchKernelGGL((k_null<TS>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g->stream, n, z.copy,
                       key.epi, z.ec, in + o, out + o, nullptr);
    HIPCHECK(hipGetLastError());
  }
  return 0;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* pyr_last_error(void) { return g_err.c_str(); }

int pyr_setup(pyr_ctx** ctx, int nx, int ny, int nz, double x1, double xn, double y1, double yn,
              double z1, double zn, const int* bc6) {
  return setup_common(ctx, nx, ny, nz, x1, xn, y1, yn, z1, zn, bc6, true);
}

int pyr_setup_host(pyr_ctx** ctx, int nx, int ny, int nz, double x1, double xn, double y1,
                   double yn, double z1, double zn, const int* bc6) {
  return setup_common(ctx, nx, ny, nz, x1, xn, y1, yn, z1, zn, bc6, false);
}

int pyr_destroy(pyr_ctx* c) {
  if (!c) return 0;
  if (c->device) {
    for (int op = 0; op < NOPS; ++op)
      for (int d = 0; d < 3; ++d)
        for (DevPlan* dp : {&c->dplan[op][d], &c->dplan_a[op][d]})
          if (dp->mem) (void)hipFree(dp->mem);
    for (int i = 0; i < 2; ++i)
      if (c->field_tmp[i]) (void)hipFree(c->field_tmp[i]);
  }
  delete c;
  return 0;
}

int pyr_set_stream(pyr_ctx* c, void* stream) {
  if (!c) return fail("null context");
  c->stream = (hipStream_t)stream;
  return 0;
}

int pyr_set_r32_solve(pyr_ctx* c, int mode) {
  if (!c) return fail("null context");
  if (mode != PYR_R32_SOLVE_REFINED && mode != PYR_R32_SOLVE_FP64) return fail("bad r32 solve mode");
  c->r32_solve = mode;
  return 0;
}

int pyr_set_tuning(pyr_ctx* c, int op, int dir, int prec, int x_lds_kb, int yz_pipe, int yz_lines) {
  if (!c) return fail("null context");
  if (op < 0 || op >= NOPS || dir < 0 || dir > 2) return fail("bad operator or direction");
  if (x_lds_kb >= 0 && (x_lds_kb < 8 || x_lds_kb > 160)) return fail("x_lds_kb must be 8..160 (or -1)");
  if (yz_pipe > 3 || (yz_pipe == 2 && dir != 0) || (yz_pipe == 3 && dir == 0))
    return fail("yz_pipe must be 0, 1, 2 (x) or 3 (y/z) (or -1)");
  if (yz_lines >= 0 && yz_lines != 16 && yz_lines != 32 && yz_lines != 64)
    return fail("yz_lines must be 16, 32 or 64 (or -1)");
  int mode;
  if (prec == PYR_F64) mode = MODE_F64;
  else if (prec == PYR_R32) mode = r32_mode(c, op, dir) == PYR_R32_SOLVE_FP64 ? MODE_S32 : MODE_R32;
  else return fail("bad precision");
  c->tune_xlds[op][dir][mode] = x_lds_kb < 0 ? -1 : x_lds_kb;
  c->tune_pipe[op][dir][mode] = yz_pipe < 0 ? -1 : yz_pipe;
  c->tune_ly[op][dir][mode] = yz_lines < 0 ? -1 : yz_lines;
  return 0;
}

int pyr_set_solve_tuning(pyr_ctx* c, int op, int dir, int r32_solve) {
  if (!c) return fail("null context");
  if (op < 0 || op >= NOPS || dir < 0 || dir > 2) return fail("bad operator or direction");
  if (r32_solve != -1 && r32_solve != PYR_R32_SOLVE_REFINED && r32_solve != PYR_R32_SOLVE_FP64)
    return fail("bad r32 solve mode");
  c->tune_solve[op][dir] = r32_solve;
  return 0;
}

int pyr_set_periodic_solver(pyr_ctx* c, int mode) {
  if (!c) return fail("null context");
  if (mode != PYR_PSOLVE_CIRCULANT && mode != PYR_PSOLVE_LU) return fail("bad periodic solver");
  c->periodic_solver = mode;
  return 0;
}

int pyr_axis_coords(pyr_ctx* c, int axis, double* coords) {
  if (!c || !coords || axis < 0 || axis > 2) return fail("bad arguments to pyr_axis_coords");
  // xgrid = x1f + (2i-1)*half*dx (mesh.f90:173-177)
  for (int i = 1; i <= c->n[axis]; ++i)
    coords[i - 1] = c->x1[axis] + (double)(2 * i - 1) * 0.5 * c->dpatch[axis];
  return 0;
}

int pyr_mesh_scalars(pyr_ctx* c, double* dpatch3, double* dmesh3, double* gridlen) {
  if (!c) return fail("null context");
  for (int d = 0; d < 3; ++d) {
    if (dpatch3) dpatch3[d] = c->dpatch[d];
    if (dmesh3) dmesh3[d] = c->dmesh[d];
  }
  if (gridlen) *gridlen = c->gridlen;
  return 0;
}

int pyr_apply(pyr_ctx* c, int op, int dir, int prec, const void* in, void* out, int epi,
              double epi_scale) {
  if (epi < PYR_EPI_STORE || epi > PYR_EPI_MAXABS_SCALED) return fail("bad epilogue");
  const bool anti = (op & PYR_OP_ANTISYM) != 0;
  return apply_any(c, op & ~PYR_OP_ANTISYM, dir, prec, in, out, epi, epi_scale, anti);
}

#define PYR_DIR_OP(name, op, dir)                                                     \
  int name##_f64(pyr_ctx* c, const double* v, double* o) {                           \
    return apply_any(c, op, dir, PYR_F64, v, o, EPI_STORE, 0.0);                     \
  }
PYR_DIR_OP(pyr_ddx, OP_D1, 0)
PYR_DIR_OP(pyr_ddy, OP_D1, 1)
PYR_DIR_OP(pyr_ddz, OP_D1, 2)
PYR_DIR_OP(pyr_dd8x, OP_D8, 0)
PYR_DIR_OP(pyr_dd8y, OP_D8, 1)
PYR_DIR_OP(pyr_dd8z, OP_D8, 2)
PYR_DIR_OP(pyr_d2x, OP_D2, 0)
PYR_DIR_OP(pyr_d2y, OP_D2, 1)
PYR_DIR_OP(pyr_d2z, OP_D2, 2)
#undef PYR_DIR_OP

int pyr_ddx_r32(pyr_ctx* c, const float* v, float* o) { return apply_any(c, OP_D1, 0, PYR_R32, v, o, EPI_STORE, 0.0); }
int pyr_ddy_r32(pyr_ctx* c, const float* v, float* o) { return apply_any(c, OP_D1, 1, PYR_R32, v, o, EPI_STORE, 0.0); }
int pyr_ddz_r32(pyr_ctx* c, const float* v, float* o) { return apply_any(c, OP_D1, 2, PYR_R32, v, o, EPI_STORE, 0.0); }

int pyr_plaplacian_f64(pyr_ctx* c, const double* v, double* o) { return sum3(c, OP_D2, PYR_F64, v, v, v, o); }
int pyr_plaplacian_r32(pyr_ctx* c, const float* v, float* o) { return sum3(c, OP_D2, PYR_R32, v, v, v, o); }
int pyr_pring_f64(pyr_ctx* c, const double* v, double* o) { return ring(c, PYR_F64, v, o); }
int pyr_pring_r32(pyr_ctx* c, const float* v, float* o) { return ring(c, PYR_R32, v, o); }

int pyr_pringv_f64(pyr_ctx* c, const double* vx, const double* vy, const double* vz, double* o) {
  // ringV L=1 (operators.f90:645-699): max over components and directions of |d8| * d
  const double* f[3] = {vx, vy, vz};
  bool first = true;
  for (int d = 0; d < 3; ++d)
    for (int comp = 0; comp < 3; ++comp) {
      // the normal component takes the antisymmetric member at SYMM ends
      // (ringx(f,..,isymX) etc., operators.f90:661-671)
      if (apply_any(c, OP_D8, d, PYR_F64, f[comp], o, first ? EPI_ABS : EPI_MAXABS, c ? c->dmesh[d] : 0.0,
                    comp == d))
        return 1;
      first = false;
    }
  return 0;
}

int pyr_sfilter_f64(pyr_ctx* c, const double* v, double* o) { return filter3(c, OP_SF, PYR_F64, v, o); }
int pyr_apply_rounded(pyr_ctx* c, int op, int dir, const double* in, double* out, float* o32) {
  if (check_ctx(c)) return 1;
  if (!o32 || !out) return fail("null field pointer");
  c->out32 = o32;
  const int r = apply_any(c, op, dir, PYR_F64, in, out, EPI_STORE, 0.0);
  c->out32 = nullptr;
  return r;
}
int pyr_apply_to_r32(pyr_ctx* c, int op, int dir, const double* in, float* o32) {
  if (check_ctx(c)) return 1;
  if (!o32 || !in) return fail("null field pointer");
  c->out32 = o32;
  c->only32 = 1;
  const int r = apply_any(c, op, dir, PYR_F64, in, nullptr, EPI_STORE, 0.0);
  c->out32 = nullptr;
  c->only32 = 0;
  return r;
}
int pyr_sfilter_f64_rounded(pyr_ctx* c, const double* v, double* o, float* o32) {
  if (check_ctx(c)) return 1;
  if (!o32) return fail("null field pointer");
  void* tmp;
  if (field_temp(c, PYR_F64, &tmp)) return 1;
  if (apply_any(c, OP_SF, 0, PYR_F64, v, o, EPI_STORE, 0.0)) return 1;
  if (apply_any(c, OP_SF, 1, PYR_F64, o, tmp, EPI_STORE, 0.0)) return 1;
  // the z sweep also stores the FP32 rounding of the result
  c->out32 = o32;
  const int r = apply_any(c, OP_SF, 2, PYR_F64, tmp, o, EPI_STORE, 0.0);
  c->out32 = nullptr;
  return r;
}
int pyr_sfilter_r32(pyr_ctx* c, const float* v, float* o) { return filter3(c, OP_SF, PYR_R32, v, o); }
int pyr_sfilter_noalloc(pyr_ctx* c, int prec, const void* v, void* o) {
  // pyr_sfilter_f64/_r32 with the field temporary in `scratch` (this device's
  // copy, from pyr_setup): no allocation; overlapping calls on one GPU race
  if (check_ctx(c)) return 1;
  if (prec != PYR_F64 && prec != PYR_R32) return fail("bad precision");
  if (!v || !o) return fail("null field pointer");
  const size_t bytes = c->npts() * (prec == PYR_F64 ? sizeof(double) : sizeof(float));
  if (bytes > PYR_NOALLOC_BYTES)
    return fail("pyr_sfilter_noalloc: the field needs " + std::to_string(bytes) +
                " bytes of temporary, above the 64 MiB scratch (use pyr_sfilter_f64/_r32)");
  if (!c->noalloc_tmp) return fail("pyr_sfilter_noalloc: no scratch address for this context's device");
  return filter3_with(c, OP_SF, prec, v, o, c->noalloc_tmp);
}
int pyr_gfilter_f64(pyr_ctx* c, const double* v, double* o) { return filter3(c, OP_GF, PYR_F64, v, o); }
int pyr_gfilter_r32(pyr_ctx* c, const float* v, float* o) { return filter3(c, OP_GF, PYR_R32, v, o); }
int pyr_gfilterdir_f64(pyr_ctx* c, const double* v, double* o, int dir) {
  if (dir < 1 || dir > 3) return fail("gfilterdir direction must be 1, 2 or 3");
  return apply_any(c, OP_GF, dir - 1, PYR_F64, v, o, EPI_STORE, 0.0);
}
int pyr_gfilterdir_r32(pyr_ctx* c, const float* v, float* o, int dir) {
  if (dir < 1 || dir > 3) return fail("gfilterdir direction must be 1, 2 or 3");
  return apply_any(c, OP_GF, dir - 1, PYR_R32, v, o, EPI_STORE, 0.0);
}
int pyr_grads_f64(pyr_ctx* c, const double* v, double* a, double* b, double* e) {
  if (apply_any(c, OP_D1, 0, PYR_F64, v, a, EPI_STORE, 0.0)) return 1;
  if (apply_any(c, OP_D1, 1, PYR_F64, v, b, EPI_STORE, 0.0)) return 1;
  return apply_any(c, OP_D1, 2, PYR_F64, v, e, EPI_STORE, 0.0);
}
int pyr_grads_r32(pyr_ctx* c, const float* v, float* a, float* b, float* e) {
  if (apply_any(c, OP_D1, 0, PYR_R32, v, a, EPI_STORE, 0.0)) return 1;
  if (apply_any(c, OP_D1, 1, PYR_R32, v, b, EPI_STORE, 0.0)) return 1;
  return apply_any(c, OP_D1, 2, PYR_R32, v, e, EPI_STORE, 0.0);
}
int pyr_divergence_f64(pyr_ctx* c, const double* fx, const double* fy, const double* fz, double* o) {
  return sum3(c, OP_D1, PYR_F64, fx, fy, fz, o, true);
}
int pyr_divergence_r32(pyr_ctx* c, const float* fx, const float* fy, const float* fz, float* o) {
  return sum3(c, OP_D1, PYR_R32, fx, fy, fz, o, true);
}

static int rk_update(hipStream_t stream, int prec, long long n, const void* F, double* PHI, const void* Uin,
                     void* Uout, double* U64, const double* dtp, const int* mid, double dt, double A,
                     double B) {
  // with an FP64 state the rounded field may be left unwritten (Uout null)
  if (!F || !PHI || (!U64 && (!Uin || !Uout))) return fail("null pointer to pyr_rk_update");
  if (n <= 0) return 0;
  long long blocks = (n + 1023) / 1024;
  if (blocks > 8192) blocks = 8192;
  if (prec == PYR_F64)
    hipLaunchKernelGGL((k_rk_update<double, double>), dim3((unsigned)blocks), dim3(256), 0, stream, n,
                       (const double*)F, PHI, (const double*)Uin, (double*)Uout, U64, dtp, mid, dt, A, B);
  else if (prec == PYR_R32)
    hipLaunchKernelGGL((k_rk_update<float, float>), dim3((unsigned)blocks), dim3(256), 0, stream, n,
                       (const float*)F, PHI, (const float*)Uin, (float*)Uout, U64, dtp, mid, dt, A, B);
  else if (prec == PYR_RK_F32_U64)
    hipLaunchKernelGGL((k_rk_update<float, double>), dim3((unsigned)blocks), dim3(256), 0, stream, n,
                       (const float*)F, PHI, (const double*)Uin, (double*)Uout, U64, dtp, mid, dt, A, B);
  else if (prec == PYR_RK_F64_U32)
    hipLaunchKernelGGL((k_rk_update<double, float>), dim3((unsigned)blocks), dim3(256), 0, stream, n,
                       (const double*)F, PHI, (const float*)Uin, (float*)Uout, U64, dtp, mid, dt, A, B);
  else
    return fail("bad precision");
  HIPCHECK(hipGetLastError());
  return 0;
}

int pyr_rk_update(pyr_ctx* c, int prec, long long n, const void* F, double* PHI, const void* Uin,
                  void* Uout, double* U64, double dt, double A, double B) {
  if (check_ctx(c)) return 1;
  return rk_update(c->stream, prec, n, F, PHI, Uin, Uout, U64, nullptr, nullptr, dt, A, B);
}

int pyr_rk_update_dt(pyr_ctx* c, int prec, long long n, const void* F, double* PHI, const void* Uin,
                     void* Uout, double* U64, const double* dt, double A, double B) {
  if (check_ctx(c)) return 1;
  if (!dt) return fail("null dt pointer to pyr_rk_update_dt");
  return rk_update(c->stream, prec, n, F, PHI, Uin, Uout, U64, dt, nullptr, 0.0, A, B);
}

int pyr_rk_update_member_dt(void* stream, int prec, long long n, const void* F, double* PHI,
                            const void* Uin, void* Uout, double* U64, const double* dt_member,
                            const int* member_of_point, double A, double B) {
  if (!dt_member || !member_of_point) return fail("null pointer to pyr_rk_update_member_dt");
  return rk_update((hipStream_t)stream, prec, n, F, PHI, Uin, Uout, U64, dt_member, member_of_point, 0.0, A, B);
}

int pyr_group_create(pyr_group** out, int M, pyr_ctx* const* members, const long long* offsets) {
  if (!out || M < 1 || !members || !offsets) return fail("bad arguments to pyr_group_create");
  std::unique_ptr<pyr_group> g(new pyr_group());
  for (int m = 0; m < M; ++m) {
    if (check_ctx(members[m])) return 1;
    if (offsets[m] < 0) return fail("negative member offset");
    g->mem.push_back(members[m]);
    g->off.push_back(offsets[m]);
  }
  *out = g.release();
  return 0;
}

int pyr_group_destroy(pyr_group* g) {
  delete g;
  return 0;
}

int pyr_group_set_stream(pyr_group* g, void* stream) {
  if (!g) return fail("null group");
  g->stream = (hipStream_t)stream;
  return 0;
}

int pyr_group_apply(pyr_group* g, int op, int dir, int prec, const void* in, void* out, int epi, double ec,
                    int mesh_power) {
  if (!g) return fail("null group");
  if (op < 0 || op >= NOPS || dir < 0 || dir > 2) return fail("bad operator or direction");
  if (!in || !out) return fail("null field pointer");
  if (in == out) return fail("input and output must not alias");
  if (mesh_power < 0 || mesh_power > 2) return fail("mesh_power must be 0, 1 or 2");
  GrpKey key;
  key.op = op;
  key.dir = dir;
  key.prec = prec;
  key.epi = epi;
  key.power = mesh_power;
  key.aligned = ((uintptr_t)in % 16 == 0) && ((uintptr_t)out % 16 == 0);
  key.ec = ec;
  key.psolve = g->mem[0]->periodic_solver;
  key.xlds = g->mem[0]->x_lds_kb;
  for (pyr_ctx* c : g->mem)
    if (c->periodic_solver != key.psolve || c->x_lds_kb != key.xlds)
      return fail("group members differ in their periodic solver or x-line LDS budget");
  if (prec == PYR_F64) {
    key.mode = MODE_F64;
    return grp_apply<double, MODE_F64>(g, key, (const double*)in, (double*)out);
  }
  if (prec == PYR_R32) {
    // the solve arithmetic of this operator and direction (per-operator
    // tuning included), the same for every member
    const int r = r32_mode(g->mem[0], op, dir);
    for (pyr_ctx* c : g->mem)
      if (r32_mode(c, op, dir) != r) return fail("group members differ in their FP32 solve mode");
    key.mode = r == PYR_R32_SOLVE_FP64 ? MODE_S32 : MODE_R32;
    if (r == PYR_R32_SOLVE_FP64) return grp_apply<float, MODE_S32>(g, key, (const float*)in, (float*)out);
    return grp_apply<float, MODE_R32>(g, key, (const float*)in, (float*)out);
  }
  return fail("bad precision");
}

int pyr_host_weights(int op, double* ali, double* ari, double* alb1, double* alb2, double* arb1,
                     double* arb2, int* dims6) {
  if (op < 0 || op >= NOPS) return fail("bad operator");
  const Scheme s = make_scheme(op);
  std::memcpy(ali, s.ali, sizeof s.ali);
  std::memcpy(ari, s.ari, sizeof s.ari);
  std::memcpy(alb1, s.alb1, sizeof s.alb1);
  std::memcpy(alb2, s.alb2, sizeof s.alb2);
  std::memcpy(arb1, s.arb1, sizeof s.arb1);
  std::memcpy(arb2, s.arb2, sizeof s.arb2);
  const int dims[6] = {s.nol, s.nor, s.ncl, s.ncr, s.null_option, s.implicit};
  std::memcpy(dims6, dims, sizeof dims);
  return 0;
}

int pyr_host_emulate(pyr_ctx* c, int op, int dir, int prec, const void* in, void* out) {
  if (!c) return fail("null context");
  if (op < 0 || op >= NOPS || dir < 0 || dir > 2) return fail("bad operator or direction");
  if (prec == PYR_F64) return emulate<double, MODE_F64>(c, op, dir, (const double*)in, (double*)out);
  if (prec == PYR_R32) {
    if (r32_mode(c, op, dir) == PYR_R32_SOLVE_FP64) return emulate<float, MODE_S32>(c, op, dir, (const float*)in, (float*)out);
    return emulate<float, MODE_R32>(c, op, dir, (const float*)in, (float*)out);
  }
  return fail("bad precision");
}

}  // extern "C"
#endif  // PYR_PART == 0
