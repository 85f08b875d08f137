// Learned stand-in for the LAD artificial-property chain (gfx950, BF16 MFMA).
//
// A pointwise MLP maps the 3x3 neighbourhoods of three feature fields of a
// 2-D grid (dilatation, vorticity, |grad rho|: 27 inputs, padded to 32) to two
// outputs (bulk viscosity beta and species diffusivity adiff):
//     h1 = relu(W1^T x + b1), h2 = relu(W2^T h1 + b2), o = W3^T h2 + b3
// with 64 hidden units per layer, BF16 weights and activations and FP32
// accumulation on the matrix cores.
//
// Inference (k_mlp_infer32) runs on v_mfma_f32_16x16x32_bf16 with the weights
// and the activations in registers (see the comment above it); it also
// flags, per point, inputs beyond a sensor threshold. Training (k_mlp_train)
// runs the forward pass, the backward pass (dH2 = W3 dO, dH1 = W2 dZ2) and the
// weight gradients (dW = A^T dZ summed over points, points as the MFMA K
// dimension) in one kernel on v_mfma_f32_16x16x16_bf16, tiles passing through
// a per-wave LDS slice; each workgroup writes one partial gradient and a
// second kernel sums the partials in a fixed order (deterministic).
//
// No counterpart in the reference: pyranda evaluates the LAD chain with
// compact operators only (examples/RM3D.py:95-104, TaylorGreen.py:77-81).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/pyranda_amd.h"

namespace pyr {
int report_error(const std::string& m);
}

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NIN = 32;   // 27 features padded to two K blocks of 16
constexpr int NH = 64;    // hidden units
constexpr int NO = 16;    // outputs padded to one 16-row tile (2 used)
constexpr int NF = 27;
constexpr int WAVES = 4;  // waves per workgroup, one 16-point tile each at a time
constexpr int GRAD = PYR_MLP_GRAD_FLOATS;
constexpr int OFF_W2 = NIN * NH, OFF_W3 = OFF_W2 + NH * NH, NBF = OFF_W3 + NH * NO;
constexpr int G_W1 = 0, G_W2 = NIN * NH, G_W3 = G_W2 + NH * NH, G_B1 = G_W3 + NH * NO, G_B2 = G_B1 + NH,
              G_B3 = G_B2 + NH, G_LOSS = G_B3 + NO;
static_assert(G_LOSS + 1 == GRAD, "gradient layout");
static_assert(2 * NBF == PYR_MLP_PARAM_FLOAT_OFFSET, "parameter layout");

// round to nearest even, as torch's float -> bfloat16
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ f32x4 mma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// 4 BF16 values at base + i*step
__device__ __forceinline__ bf16x4 frag(const unsigned short* base, int step) {
  bf16x4 r;
  r[0] = (short)base[0];
  r[1] = (short)base[step];
  r[2] = (short)base[2 * step];
  r[3] = (short)base[3 * step];
  return r;
}

__device__ __forceinline__ int wrap_clamp(int i, int n, int periodic) {
  if (periodic) return i < 0 ? i + n : (i >= n ? i - n : i);
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

template <typename TS>
struct MlpArgs {
  int nx, ny, px, py;
  const TS* f[3];
  const TS* t[2];    // training targets
  TS* o[2];          // inference outputs
  const unsigned char* params;
  float* partial;    // training: one GRAD-float block per workgroup
  float inv_n;       // training: 1 / number of points
};

// Training: forward and backward of 16-point tiles on 16x16x16 BF16 tiles
// (the weight gradients take the 16 points of a tile as the K dimension),
// activations in a per-wave LDS slice, one partial gradient per workgroup.
template <typename TS>
__global__ void __launch_bounds__(64 * WAVES) k_mlp_train(MlpArgs<TS> a) {
  constexpr int ACT = 16 * (NIN + 4 * NH + NO);
  static_assert(WAVES * ACT * 2 >= GRAD * 4, "reduction buffer");
  __shared__ unsigned short W1s[NIN * NH], W2s[NH * NH], W3s[NH * NO];
  __shared__ float B1s[NH], B2s[NH], B3s[NO], NRM[12];
  __shared__ __align__(16) unsigned short act[WAVES * ACT];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, li = l & 15, g = l >> 4;

  const unsigned short* pw = reinterpret_cast<const unsigned short*>(a.params);
  for (int i = tid; i < NIN * NH; i += 64 * WAVES) W1s[i] = pw[i];
  for (int i = tid; i < NH * NH; i += 64 * WAVES) W2s[i] = pw[OFF_W2 + i];
  for (int i = tid; i < NH * NO; i += 64 * WAVES) W3s[i] = pw[OFF_W3 + i];
  const float* pf = reinterpret_cast<const float*>(a.params + PYR_MLP_PARAM_FLOAT_OFFSET);
  if (tid < NH) {
    B1s[tid] = pf[tid];
    B2s[tid] = pf[NH + tid];
  }
  if (tid < NO) B3s[tid] = pf[2 * NH + tid];
  if (tid < 12) NRM[tid] = pf[2 * NH + NO + tid];  // fshift[4] fscale[4] oscale[4]
  unsigned short* Xs = act + w * ACT;             // [16 points][NIN]
  unsigned short* H1s = Xs + 16 * NIN;            // [16][NH]
  unsigned short* H2s = H1s + 16 * NH;            // [16][NH]
  unsigned short* dOs = H2s + 16 * NH;            // [16][NO]
  unsigned short* dZ2s = dOs + 16 * NO;           // [16][NH]
  unsigned short* dZ1s = dZ2s + 16 * NH;          // [16][NH]
  for (int i = l; i < 16 * NO; i += 64) dOs[i] = 0;
  __syncthreads();

  f32x4 acc1[2][4], acc2[4][4], acc3[4];
  float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, loss = 0.f;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 4; ++j) acc1[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int i = 0; i < 4; ++i) {
    acc3[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int j = 0; j < 4; ++j) acc2[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  }

  const long long npts = (long long)a.nx * a.ny;
  const long long ntiles = (npts + 15) / 16;
  const long long twaves = (long long)gridDim.x * WAVES;
  const long long iters = (ntiles + twaves - 1) / twaves;  // the same for every wave
  const long long wg = (long long)blockIdx.x * WAVES + w;
  for (long long it = 0; it < iters; ++it) {
    const long long t = it * twaves + wg;
    const long long p = t * 16 + li;
    const bool valid = t < ntiles && p < npts;
    const int pi = valid ? (int)(p % a.nx) : 0, pj = valid ? (int)(p / a.nx) : 0;
    // features: x[f*9 + 3*(dy+1) + (dx+1)] = (f_f(i+dx, j+dy) - shift_f) * scale_f
#pragma unroll
    for (int q = 0; q < NIN / 4; ++q) {
      const int k = g + 4 * q;
      float v = 0.f;
      if (valid && k < NF) {
        const int f = k / 9, r = k % 9;
        const int ii = wrap_clamp(pi + r % 3 - 1, a.nx, a.px);
        const int jj = wrap_clamp(pj + r / 3 - 1, a.ny, a.py);
        v = ((float)a.f[f][(long long)jj * a.nx + ii] - NRM[f]) * NRM[4 + f];
      }
      Xs[li * NIN + k] = f2bf(v);
    }
    __syncthreads();
    // layer 1: Z1^T[m][pt] = sum_k W1[k][m] X[pt][k]
#pragma unroll
    for (int mt = 0; mt < NH / 16; ++mt) {
      f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < NIN / 16; ++kb)
        d = mma(frag(&W1s[(16 * kb + 4 * g) * NH + 16 * mt + li], NH), frag(&Xs[li * NIN + 16 * kb + 4 * g], 1), d);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = 16 * mt + 4 * g + i;
        H1s[li * NH + m] = f2bf(fmaxf(d[i] + B1s[m], 0.f));
      }
    }
    __syncthreads();
    // layer 2
#pragma unroll
    for (int mt = 0; mt < NH / 16; ++mt) {
      f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < NH / 16; ++kb)
        d = mma(frag(&W2s[(16 * kb + 4 * g) * NH + 16 * mt + li], NH), frag(&H1s[li * NH + 16 * kb + 4 * g], 1), d);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = 16 * mt + 4 * g + i;
        H2s[li * NH + m] = f2bf(fmaxf(d[i] + B2s[m], 0.f));
      }
    }
    __syncthreads();
    // layer 3: O^T[j][pt]; lanes 0..15 hold outputs 0..3 of point li
    f32x4 o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < NH / 16; ++kb)
      o = mma(frag(&W3s[(16 * kb + 4 * g) * NO + li], NO), frag(&H2s[li * NH + 16 * kb + 4 * g], 1), o);
    {
      if (g == 0) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const float tj = valid ? (float)a.t[j][p] / NRM[8 + j] : 0.f;
          const float e = valid ? (o[j] + B3s[j]) - tj : 0.f;
          loss += e * e;
          dOs[li * NO + j] = f2bf(2.f * e * a.inv_n);
        }
      }
      __syncthreads();
      // dZ2^T[m][pt] = relu'(Z2) * sum_j W3[m][j] dO[pt][j]
#pragma unroll
      for (int mt = 0; mt < NH / 16; ++mt) {
        const f32x4 z = {0.f, 0.f, 0.f, 0.f};
        const f32x4 d = mma(frag(&W3s[(16 * mt + li) * NO + 4 * g], 1), frag(&dOs[li * NO + 4 * g], 1), z);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int m = 16 * mt + 4 * g + i;
          dZ2s[li * NH + m] = f2bf(bf2f(H2s[li * NH + m]) > 0.f ? d[i] : 0.f);
        }
      }
      __syncthreads();
      // dZ1^T[k][pt] = relu'(Z1) * sum_m W2[k][m] dZ2[pt][m]
#pragma unroll
      for (int kt = 0; kt < NH / 16; ++kt) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int mb = 0; mb < NH / 16; ++mb)
          d = mma(frag(&W2s[(16 * kt + li) * NH + 16 * mb + 4 * g], 1), frag(&dZ2s[li * NH + 16 * mb + 4 * g], 1), d);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int k = 16 * kt + 4 * g + i;
          dZ1s[li * NH + k] = f2bf(bf2f(H1s[li * NH + k]) > 0.f ? d[i] : 0.f);
        }
      }
      __syncthreads();
      // weight gradients, the 16 points of the tile as the K dimension
#pragma unroll
      for (int mt = 0; mt < NH / 16; ++mt)
        acc3[mt] = mma(frag(&H2s[(4 * g) * NH + 16 * mt + li], NH), frag(&dOs[(4 * g) * NO + li], NO), acc3[mt]);
#pragma unroll
      for (int kt = 0; kt < NH / 16; ++kt)
#pragma unroll
        for (int mt = 0; mt < NH / 16; ++mt)
          acc2[kt][mt] = mma(frag(&H1s[(4 * g) * NH + 16 * kt + li], NH), frag(&dZ2s[(4 * g) * NH + 16 * mt + li], NH),
                             acc2[kt][mt]);
#pragma unroll
      for (int ft = 0; ft < NIN / 16; ++ft)
#pragma unroll
        for (int kt = 0; kt < NH / 16; ++kt)
          acc1[ft][kt] = mma(frag(&Xs[(4 * g) * NIN + 16 * ft + li], NIN), frag(&dZ1s[(4 * g) * NH + 16 * kt + li], NH),
                             acc1[ft][kt]);
      for (int q = 0; q < 16; ++q) {
        gb1 += bf2f(dZ1s[q * NH + l]);
        gb2 += bf2f(dZ2s[q * NH + l]);
        if (l < NO) gb3 += bf2f(dOs[q * NO + l]);
      }
    }
    __syncthreads();
  }

  {
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    float* red = reinterpret_cast<float*>(act);  // the activation slices are free now
    for (int turn = 0; turn < WAVES; ++turn) {
      if (w == turn) {
        auto put = [&](int idx, float v) { red[idx] = turn == 0 ? v : red[idx] + v; };
#pragma unroll
        for (int ft = 0; ft < NIN / 16; ++ft)
#pragma unroll
          for (int kt = 0; kt < NH / 16; ++kt)
#pragma unroll
            for (int i = 0; i < 4; ++i) put(G_W1 + (16 * ft + 4 * g + i) * NH + 16 * kt + li, acc1[ft][kt][i]);
#pragma unroll
        for (int kt = 0; kt < NH / 16; ++kt)
#pragma unroll
          for (int mt = 0; mt < NH / 16; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) put(G_W2 + (16 * kt + 4 * g + i) * NH + 16 * mt + li, acc2[kt][mt][i]);
#pragma unroll
        for (int mt = 0; mt < NH / 16; ++mt)
#pragma unroll
          for (int i = 0; i < 4; ++i) put(G_W3 + (16 * mt + 4 * g + i) * NO + li, acc3[mt][i]);
        put(G_B1 + l, gb1);
        put(G_B2 + l, gb2);
        if (l < NO) put(G_B3 + l, gb3);
        if (l == 0) put(G_LOSS, loss);
      }
      __syncthreads();
    }
    float* dst = a.partial + (size_t)blockIdx.x * GRAD;
    for (int i = tid; i < GRAD; i += 64 * WAVES) dst[i] = red[i];
  }
}

// --- inference on gfx950's 16x16x32 BF16 matrix cores ----------------------
// v_mfma_f32_16x16x32_bf16: A is 16 x 32 (lane l holds row l%16, eight K
// values of K group l/16), B is 32 x 16 (lane l holds column l%16, eight K
// values of K group l/16), D is 16 x 16 (lane l holds rows 4(l/16)..+3 of
// column l%16). A and B take their K values from the same (lane group,
// element) slots, so the order of the 32 K values inside a block is free as
// long as both operands use it. Every layer is computed transposed (rows =
// units, columns = the 16 points of a tile); the units of a 32-wide K block
// are ordered so that lane group g's eight B values are rows 4g..4g+3 of two
// D tiles of the layer below: slot (g, e) of block kb is unit
//     32 kb + 16 (e / 4) + 4 g + e % 4.
// A layer's output therefore feeds the next layer straight from registers
// (no LDS, no barrier), and each lane keeps its A fragments of all three
// weight matrices in registers for the whole persistent loop.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x4 mma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int kunit(int kb, int g, int e) { return 32 * kb + 16 * (e >> 2) + 4 * g + (e & 3); }

// relu(d + b) of two D tiles (rows 4g..4g+3 each), as one B operand
__device__ __forceinline__ bf16x8 pack_relu(f32x4 d0, const float* b0, f32x4 d1, const float* b1) {
  s16x8 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    r[i] = (short)f2bf(fmaxf(d0[i] + b0[i], 0.f));
    r[4 + i] = (short)f2bf(fmaxf(d1[i] + b1[i], 0.f));
  }
  return __builtin_bit_cast(bf16x8, r);
}

template <typename TS>
struct InferArgs {
  int nx, ny, px, py;
  const TS* f[3];
  TS* o[2];
  const unsigned char* params;
  float sensor;            // > 0: flag points whose largest normalised |input| exceeds it
  unsigned char* mask;     // per point 1 (flagged) or 0, when not null
  int* count;              // flagged points are added here, when not null
};

template <typename TS>
__global__ void __launch_bounds__(64 * WAVES) k_mlp_infer32(InferArgs<TS> a) {
  const int l = threadIdx.x & 63, li = l & 15, g = l >> 4;
  const unsigned short* pw = reinterpret_cast<const unsigned short*>(a.params);
  const float* pf = reinterpret_cast<const float*>(a.params + PYR_MLP_PARAM_FLOAT_OFFSET);
  // A fragments: W^T rows (output units) of this lane, K slots of its group
  bf16x8 A1[NH / 16], A2[NH / 16][2], A3[2];
#pragma unroll
  for (int mt = 0; mt < NH / 16; ++mt) {
    s16x8 r;
#pragma unroll
    for (int e = 0; e < 8; ++e) r[e] = (short)pw[(8 * g + e) * NH + 16 * mt + li];  // input k = 8g + e
    A1[mt] = __builtin_bit_cast(bf16x8, r);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      s16x8 q;
#pragma unroll
      for (int e = 0; e < 8; ++e) q[e] = (short)pw[OFF_W2 + kunit(kb, g, e) * NH + 16 * mt + li];
      A2[mt][kb] = __builtin_bit_cast(bf16x8, q);
    }
  }
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    s16x8 q;
#pragma unroll
    for (int e = 0; e < 8; ++e) q[e] = (short)pw[OFF_W3 + kunit(kb, g, e) * NO + li];
    A3[kb] = __builtin_bit_cast(bf16x8, q);
  }
  // biases of this lane's D rows (units 16 mt + 4 g + i)
  float b1[NH / 16][4], b2[NH / 16][4], b3[4];
#pragma unroll
  for (int mt = 0; mt < NH / 16; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      b1[mt][i] = pf[16 * mt + 4 * g + i];
      b2[mt][i] = pf[NH + 16 * mt + 4 * g + i];
    }
#pragma unroll
  for (int i = 0; i < 4; ++i) b3[i] = pf[2 * NH + 4 * g + i];
  const float* nrm = pf + 2 * NH + NO;  // fshift[4] fscale[4] oscale[4]
  const float sh[3] = {nrm[0], nrm[1], nrm[2]}, sc[3] = {nrm[4], nrm[5], nrm[6]};
  const float os0 = nrm[8], os1 = nrm[9];

  const long long npts = (long long)a.nx * a.ny;
  const long long ntiles = (npts + 15) / 16;
  const long long twaves = (long long)gridDim.x * WAVES;
  for (long long t = (long long)blockIdx.x * WAVES + (threadIdx.x >> 6); t < ntiles; t += twaves) {
    const long long p = t * 16 + li;
    const bool valid = p < npts;
    const int pi = valid ? (int)(p % a.nx) : 0, pj = valid ? (int)(p / a.nx) : 0;
    // this lane's eight inputs k = 8g + e of point p
    float big = 0.f;
    s16x8 xr;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = 8 * g + e;
      float v = 0.f;
      if (valid && k < NF) {
        const int f = k / 9, r = k % 9;
        const int ii = wrap_clamp(pi + r % 3 - 1, a.nx, a.px);
        const int jj = wrap_clamp(pj + r / 3 - 1, a.ny, a.py);
        v = ((float)a.f[f][(long long)jj * a.nx + ii] - sh[f]) * sc[f];
        big = fmaxf(big, fabsf(v));
      }
      xr[e] = (short)f2bf(v);
    }
    const bf16x8 X = __builtin_bit_cast(bf16x8, xr);
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    f32x4 d1[NH / 16];
#pragma unroll
    for (int mt = 0; mt < NH / 16; ++mt) d1[mt] = mma32(A1[mt], X, z);
    const bf16x8 H1a = pack_relu(d1[0], b1[0], d1[1], b1[1]);
    const bf16x8 H1b = pack_relu(d1[2], b1[2], d1[3], b1[3]);
    f32x4 d2[NH / 16];
#pragma unroll
    for (int mt = 0; mt < NH / 16; ++mt) d2[mt] = mma32(A2[mt][1], H1b, mma32(A2[mt][0], H1a, z));
    const bf16x8 H2a = pack_relu(d2[0], b2[0], d2[1], b2[1]);
    const bf16x8 H2b = pack_relu(d2[2], b2[2], d2[3], b2[3]);
    const f32x4 o = mma32(A3[1], H2b, mma32(A3[0], H2a, z));
    // the point's largest |input| over the four lane groups
    big = fmaxf(big, __shfl_xor(big, 16, 64));
    big = fmaxf(big, __shfl_xor(big, 32, 64));
    const bool flag = valid && a.sensor > 0.f && big > a.sensor;
    if (g == 0 && valid) {
      a.o[0][p] = (TS)(fmaxf(o[0] + b3[0], 0.f) * os0);
      a.o[1][p] = (TS)(fmaxf(o[1] + b3[1], 0.f) * os1);
      if (a.mask) a.mask[p] = flag ? 1 : 0;
    }
    if (a.count) {
      const unsigned long long bal = __ballot(g == 0 && flag);
      if (bal && l == 0) atomicAdd(a.count, (int)__popcll(bal));
    }
  }
}

// grad[i] = sum over workgroups b of partial[b][i], in a fixed order: thread
// (i, q) of a 16 x 16 workgroup sums b = q, q + 16, ... and the 16 sums are
// added in order of q (many short independent load chains instead of one
// chain of nb loads per output)
constexpr int SUM_I = 16, SUM_Q = 16;
__global__ void __launch_bounds__(SUM_I * SUM_Q) k_mlp_sum(const float* __restrict__ partial, int nb,
                                                          float* __restrict__ grad) {
  __shared__ float red[SUM_Q][SUM_I];
  const int il = threadIdx.x % SUM_I, q = threadIdx.x / SUM_I;
  const int i = blockIdx.x * SUM_I + il;
  float s = 0.f;
  if (i < GRAD)
    for (int b = q; b < nb; b += SUM_Q) s += partial[(size_t)b * GRAD + i];
  red[q][il] = s;
  __syncthreads();
  if (q == 0 && i < GRAD) {
    float t = 0.f;
    for (int k = 0; k < SUM_Q; ++k) t += red[k][il];
    grad[i] = t;
  }
}

template <typename TS>
int mlp_launch(void* stream, int nx, int ny, int px, int py, const void* const* f, const void* const* t,
               void* const* o, const void* params, float* partial, int npartial, float* grad) {
  MlpArgs<TS> a;
  a.nx = nx;
  a.ny = ny;
  a.px = px ? 1 : 0;
  a.py = py ? 1 : 0;
  for (int i = 0; i < 3; ++i) a.f[i] = static_cast<const TS*>(f[i]);
  for (int i = 0; i < 2; ++i) {
    a.t[i] = t ? static_cast<const TS*>(t[i]) : nullptr;
    a.o[i] = o ? static_cast<TS*>(o[i]) : nullptr;
  }
  a.params = static_cast<const unsigned char*>(params);
  a.partial = partial;
  const long long npts = (long long)nx * ny;
  a.inv_n = (float)(1.0 / (double)npts);
  const long long ntiles = (npts + 15) / 16;
  long long blocks = (ntiles + WAVES - 1) / WAVES;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (!grad) {
    return pyr::report_error("mlp_launch: inference goes through infer_launch");
  } else {
    if (blocks > npartial) blocks = npartial;
    hipLaunchKernelGGL((k_mlp_train<TS>), dim3((unsigned)blocks), dim3(64 * WAVES), 0, s, a);
    hipLaunchKernelGGL(k_mlp_sum, dim3((GRAD + SUM_I - 1) / SUM_I), dim3(SUM_I * SUM_Q), 0, s,
                       (const float*)partial, (int)blocks, grad);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return pyr::report_error(std::string("MLP launch: ") + hipGetErrorString(e));
  return 0;
}

template <typename TS>
int infer_launch(void* stream, int nx, int ny, int px, int py, const void* const* f, void* const* o,
                 const void* params, float sensor, unsigned char* mask, int* count) {
  InferArgs<TS> a;
  a.nx = nx;
  a.ny = ny;
  a.px = px ? 1 : 0;
  a.py = py ? 1 : 0;
  for (int i = 0; i < 3; ++i) a.f[i] = static_cast<const TS*>(f[i]);
  for (int i = 0; i < 2; ++i) a.o[i] = static_cast<TS*>(o[i]);
  a.params = static_cast<const unsigned char*>(params);
  a.sensor = sensor;
  a.mask = mask;
  a.count = count;
  const long long ntiles = ((long long)nx * ny + 15) / 16;
  long long blocks = (ntiles + WAVES - 1) / WAVES;
  if (blocks > 2048) blocks = 2048;  // persistent: every wave loops over tiles
  hipLaunchKernelGGL((k_mlp_infer32<TS>), dim3((unsigned)blocks), dim3(64 * WAVES), 0,
                     static_cast<hipStream_t>(stream), a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return pyr::report_error(std::string("MLP launch: ") + hipGetErrorString(e));
  return 0;
}

int check_args(int prec, int nx, int ny, const void* const* f, const void* params) {
  if (prec != PYR_F64 && prec != PYR_R32) return pyr::report_error("bad precision");
  if (nx < 1 || ny < 1) return pyr::report_error("empty grid");
  if ((long long)nx * ny > (1LL << 31)) return pyr::report_error("grid too large for the MLP kernels");
  if (!f || !f[0] || !f[1] || !f[2] || !params) return pyr::report_error("null pointer to the MLP");
  return 0;
}

}  // namespace

extern "C" {

int pyr_mlp_infer_sensed(void* stream, int prec, int nx, int ny, int periodic_x, int periodic_y,
                         const void* const* features, const void* params, void* const* out, float sensor,
                         unsigned char* mask, int* count) {
  if (check_args(prec, nx, ny, features, params)) return 1;
  if (!out || !out[0] || !out[1]) return pyr::report_error("null output pointer to pyr_mlp_infer");
  if (prec == PYR_F64)
    return infer_launch<double>(stream, nx, ny, periodic_x, periodic_y, features, out, params, sensor, mask, count);
  return infer_launch<float>(stream, nx, ny, periodic_x, periodic_y, features, out, params, sensor, mask, count);
}

int pyr_mlp_infer(void* stream, int prec, int nx, int ny, int periodic_x, int periodic_y,
                  const void* const* features, const void* params, void* const* out) {
  return pyr_mlp_infer_sensed(stream, prec, nx, ny, periodic_x, periodic_y, features, params, out, 0.f, nullptr,
                              nullptr);
}

int pyr_mlp_train_grad(void* stream, int prec, int nx, int ny, int periodic_x, int periodic_y,
                       const void* const* features, const void* const* targets, const void* params,
                       float* partial, int npartial, float* grad) {
  if (check_args(prec, nx, ny, features, params)) return 1;
  if (!targets || !targets[0] || !targets[1] || !partial || !grad)
    return pyr::report_error("null pointer to pyr_mlp_train_grad");
  if (npartial < 1) return pyr::report_error("pyr_mlp_train_grad needs at least one partial block");
  if (prec == PYR_F64)
    return mlp_launch<double>(stream, nx, ny, periodic_x, periodic_y, features, targets, nullptr, params, partial,
                              npartial, grad);
  return mlp_launch<float>(stream, nx, ny, periodic_x, periodic_y, features, targets, nullptr, params, partial,
                           npartial, grad);
}

}  // extern "C"
