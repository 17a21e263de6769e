// Learned tangent surrogate for the hyperelastic quadrature-point kernel.
//
// A 9 -> 64 -> 64 -> 21 MLP with GELU activations maps the deformation gradient
// F to the material tangent CC = dS/dE (symmetric 6x6 Voigt, 21 entries):
//     x = (F - I) / s_in,        CC = CC0 + s_out * y(x)
// where CC0 is the small-strain tangent of the law and s_out = lambda + 2 mu.
// It has no counterpart in the reference: FEDDLib evaluates the AceGen tangent.
// The Newton solver always uses the exact residual; the surrogate only replaces
// the tangent modulus and the exact tangent is the fallback on a stall.
//
// Compute: BF16 MFMA with FP32 accumulation, FP32 master weights and Adam.
// Layers are evaluated transposed, Y^T = W X^T, so a wave holds 16 samples and the
// accumulator tile of one layer (features x samples) is, lane for lane, the B
// operand of the next layer.  Products over 32 or more features use gfx950's
// v_mfma_f32_16x16x32_bf16: two consecutive 16-feature tiles form one K = 32 step,
// each lane's 8 operand values being its 4 rows of the first tile and its 4 rows
// of the second (a permutation of K, applied alike to A and B).  Products over 16
// (the input layer, and the weight gradients, whose K is the 16 samples of a tile)
// use v_mfma_f32_16x16x16_bf16.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fedd_internal.h"
#include "fedd_law.h"

#ifndef FEDD_MLP_BPC
#define FEDD_MLP_BPC 3  // inference workgroups per CU (three waves per SIMD fit in 168 VGPRs)
#endif
#ifndef FEDD_MLP_TPW
#define FEDD_MLP_TPW 2  // 16-sample tiles per wave in a training step
#endif

namespace fedd {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Parameter layout (floats): W1[64][16] b1[64] W2[64][64] b2[64] W3[32][64] b3[32].
// Input rows are padded 9 -> 16 and output rows 21 -> 32 with zeros.
enum : int {
  P_W1 = 0,
  P_B1 = P_W1 + 64 * 16,
  P_W2 = P_B1 + 64,
  P_B2 = P_W2 + 64 * 64,
  P_W3 = P_B2 + 64,
  P_B3 = P_W3 + 32 * 64,
  P_TOTAL = P_B3 + 32,
  // BF16 copies of the three weight matrices (shorts)
  Q_W1 = 0,
  Q_W2 = Q_W1 + 64 * 16,
  Q_W3 = Q_W2 + 64 * 64,
  Q_TOTAL = Q_W3 + 32 * 64,
  // the three bias vectors packed: b1[64] b2[64] b3[32]
  B_B1 = 0,
  B_B2 = 64,
  B_B3 = 128,
  B_TOTAL = 160
};

// Normalisation of the surrogate: x = (F - I) / s_in, CC = CC0 + s_out * y.
struct MlpNorm {
  float cc0[21];
  float s_out;
  double inv_s_in;
};

__device__ __forceinline__ short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u = u + 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((unsigned)(unsigned short)b) << 16); }

// Standard normal CDF Phi(z) = (1 + erf(z / sqrt 2)) / 2, with erf from Abramowitz &
// Stegun 7.1.26: erf(x) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-x^2),
// t = 1 / (1 + p x), x >= 0, |error| <= 1.5e-7 (7e-8 on Phi), well below the bfloat16
// rounding of the activations that follows.  e = exp(-z^2 / 2) is returned for the
// derivative.  About a third of the instructions of erff, which bounds these kernels.
__device__ __forceinline__ float norm_cdf(float z, float& e) {
  const float t = __builtin_amdgcn_rcpf(fmaf(0.23164188826636f, fabsf(z), 1.0f));  // p / sqrt 2
  const float poly =
      t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 1.061405429f, -1.453152027f), 1.421413741f), -0.284496736f), 0.254829592f);
  e = __builtin_amdgcn_exp2f(z * z * -0.72134752044448f);  // exp(-z^2 / 2)
  const float q = 0.5f * poly * e;
  return z >= 0.f ? 1.0f - q : q;
}
__device__ __forceinline__ float gelu(float z) {
  float e;
  return z * norm_cdf(z, e);
}
__device__ __forceinline__ float gelu_d(float z) {
  float e;
  const float c = norm_cdf(z, e);
  return fmaf(z * 0.3989422804014327f, e, c);
}

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// K = 32 step from two K = 16 operand pairs (lo: features 0..15, hi: 16..31).
__device__ __forceinline__ f4 mfma2(s4 a_lo, s4 a_hi, s4 b_lo, s4 b_hi, f4 c) {
  const s8 a = __builtin_shufflevector(a_lo, a_hi, 0, 1, 2, 3, 4, 5, 6, 7);
  const s8 b = __builtin_shufflevector(b_lo, b_hi, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a), __builtin_bit_cast(b8, b), c, 0, 0, 0);
}

// A operand from a row-major BF16 matrix with `ld` columns: rows r0 + (lane % 16),
// columns c0 + 4 (lane / 16) .. + 3.
__device__ __forceinline__ s4 load_a(const short* W, int ld, int r0, int c0, int lane) {
  const short* p = W + (r0 + (lane & 15)) * ld + c0 + 4 * (lane >> 4);
  return *reinterpret_cast<const s4*>(p);
}

struct Fwd {
  f4 z1[4], z2[4], y[2];
  s4 h1[4], h2[4];
};

// Weight operands of the forward pass read from the LDS copy at every use (training:
// its backward pass needs the registers).
struct LdsWeights {
  const short* wb;   // BF16 weights, Q_ layout
  const float* bias; // b1, b2, b3 packed, B_ layout
  int lane;
  __device__ s4 a1(int t) const { return load_a(wb + Q_W1, 16, 16 * t, 0, lane); }
  __device__ s4 a2(int t, int kk) const { return load_a(wb + Q_W2, 64, 16 * t, 16 * kk, lane); }
  __device__ s4 a3(int t, int kk) const { return load_a(wb + Q_W3, 64, 16 * t, 16 * kk, lane); }
  __device__ float b1(int t, int r) const { return bias[B_B1 + 16 * t + 4 * (lane >> 4) + r]; }
  __device__ float b2(int t, int r) const { return bias[B_B2 + 16 * t + 4 * (lane >> 4) + r]; }
  __device__ float b3(int t, int r) const { return bias[B_B3 + 16 * t + 4 * (lane >> 4) + r]; }
};

// The same operands held in registers for a whole inference kernel (56 VGPRs of A
// operands, 40 of biases), loaded once per lane from the global BF16 copy and the
// FP32 parameters.
struct RegWeights {
  s4 w1[4], w2[4][4], w3[2][4];
  float c1[4][4], c2[4][4], c3[2][4];
  __device__ void load(const short* __restrict__ wb, const float* __restrict__ p, int lane) {
    const int g = lane >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      w1[t] = load_a(wb + Q_W1, 16, 16 * t, 0, lane);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) w2[t][kk] = load_a(wb + Q_W2, 64, 16 * t, 16 * kk, lane);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        c1[t][r] = p[P_B1 + 16 * t + 4 * g + r];
        c2[t][r] = p[P_B2 + 16 * t + 4 * g + r];
      }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) w3[t][kk] = load_a(wb + Q_W3, 64, 16 * t, 16 * kk, lane);
#pragma unroll
      for (int r = 0; r < 4; ++r) c3[t][r] = p[P_B3 + 16 * t + 4 * g + r];
    }
  }
  __device__ s4 a1(int t) const { return w1[t]; }
  __device__ s4 a2(int t, int kk) const { return w2[t][kk]; }
  __device__ s4 a3(int t, int kk) const { return w3[t][kk]; }
  __device__ float b1(int t, int r) const { return c1[t][r]; }
  __device__ float b2(int t, int r) const { return c2[t][r]; }
  __device__ float b3(int t, int r) const { return c3[t][r]; }
};

// Forward pass of one 16-sample tile; xb is the B operand of the input
// (lane: sample lane % 16, features 4 (lane / 16) .. + 3).
template <class W>
__device__ __forceinline__ void forward(const W& w, s4 xb, Fwd& f) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(w.a1(t), xb, acc);
    s4 h;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float z = acc[r] + w.b1(t, r);
      f.z1[t][r] = z;
      h[r] = f2bf(gelu(z));
    }
    f.h1[t] = h;
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 4; kk += 2) acc = mfma2(w.a2(t, kk), w.a2(t, kk + 1), f.h1[kk], f.h1[kk + 1], acc);
    s4 h;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float z = acc[r] + w.b2(t, r);
      f.z2[t][r] = z;
      h[r] = f2bf(gelu(z));
    }
    f.h2[t] = h;
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 4; kk += 2) acc = mfma2(w.a3(t, kk), w.a3(t, kk + 1), f.h2[kk], f.h2[kk + 1], acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) f.y[t][r] = acc[r] + w.b3(t, r);
  }
}

// Stages the BF16 weights (16-byte copies) and the packed bias vectors in LDS.
__device__ __forceinline__ void stage_params(short* wb, float* bias, const short* __restrict__ wb_g,
                                             const float* __restrict__ p_g) {
  for (int k = threadIdx.x; k < Q_TOTAL / 8; k += blockDim.x)
    reinterpret_cast<uint4*>(wb)[k] = reinterpret_cast<const uint4*>(wb_g)[k];
  for (int k = threadIdx.x; k < B_TOTAL; k += blockDim.x)
    bias[k] = p_g[k < B_B2 ? P_B1 + k : k < B_B3 ? P_B2 + (k - B_B2) : P_B3 + (k - B_B3)];
  __syncthreads();
}

// Input operand of sample rows x[n][9] (float): lane holds features 4g..4g+3 of sample j.
__device__ __forceinline__ s4 input_b(const float* __restrict__ x, int64_t s, bool valid, int lane) {
  const int g = lane >> 4;
  s4 b;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int k = 4 * g + r;
    b[r] = (valid && k < 9) ? f2bf(x[9 * s + k]) : (short)0;
  }
  return b;
}

// Inference on the quadrature-point records of the elasticity kernel: reads F,
// writes CC into the tangent slot.  One wave per 16 records, grid-stride; the weights
// stay in registers for the whole walk.  (Writing the 21 outputs of a tile through LDS
// as whole 168-byte runs per record measured the same: the kernel is bound by the
// GELU evaluations and the latency of the load -> three layers -> store chain.)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_mlp_tangent(
    double* __restrict__ qp, int64_t nrec, const short* __restrict__ wb_g, const float* __restrict__ p_g, MlpNorm nrm) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  RegWeights w;
  w.load(wb_g, p_g, lane);
  const int64_t ntile = (nrec + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < ntile; tile += (int64_t)gridDim.x * 4) {
    const int64_t s = tile * 16 + j;
    const bool valid = s < nrec;
    s4 xb;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = 4 * g + r;
      float v = 0.f;
      if (valid && k < 9) v = (float)((qp[52 * s + k] - ((k % 4) == 0 ? 1.0 : 0.0)) * nrm.inv_s_in);
      xb[r] = f2bf(v);
    }
    Fwd f;
    forward(w, xb, f);
    if (valid) {
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int o = 16 * t + 4 * g + r;
          if (o < 21) qp[52 * s + 15 + o] = (double)nrm.cc0[o] + (double)nrm.s_out * (double)f.y[t][r];
        }
    }
  }
}

// Inference on plain rows (tests / checks): x[n][9] normalised input, y[n][21].
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_mlp_forward_rows(const float* __restrict__ x, int64_t n,
                                                          const short* __restrict__ wb_g, const float* __restrict__ p_g,
                                                          float* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  RegWeights w;
  w.load(wb_g, p_g, lane);
  const int64_t ntile = (n + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < ntile; tile += (int64_t)gridDim.x * 4) {
    const int64_t s = tile * 16 + j;
    const bool valid = s < n;
    Fwd f;
    forward(w, input_b(x, s, valid, lane), f);
    if (valid)
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int o = 16 * t + 4 * g + r;
          if (o < 21) y[21 * s + o] = f.y[t][r];
        }
  }
}

// Moves a D-layout tile (lane holds rows 4g..4g+3 of column lane % 16) through LDS
// into the A-operand layout of the same matrix (lane holds row lane % 16, columns
// 4g..4g+3); read as a B operand it is the transpose.
__device__ __forceinline__ s4 transpose_tile(float* __restrict__ buf, f4 d, int lane) {
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r) buf[(4 * g + r) * 17 + j] = d[r];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  s4 o;
#pragma unroll
  for (int r = 0; r < 4; ++r) o[r] = f2bf(buf[j * 17 + 4 * g + r]);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  return o;
}

// As transpose_tile for a BF16 D-layout tile already rounded (h operands).
__device__ __forceinline__ s4 transpose_tile_bf(float* __restrict__ buf, s4 d, int lane) {
  f4 v;
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = bf2f(d[r]);
  return transpose_tile(buf, v, lane);
}

// One training pass: forward, mean-squared-error loss on the 21 outputs, backward,
// gradients accumulated in MFMA tiles and added to grad[] with atomics.
// Each wave handles `tiles_per_wave` consecutive 16-sample tiles of the batch.
__global__ __launch_bounds__(256) void k_mlp_train(const float* __restrict__ x, const float* __restrict__ tgt,
                                                   const int64_t* __restrict__ idx, int64_t nbatch,
                                                   const short* __restrict__ wb_g, const float* __restrict__ p_g,
                                                   float* __restrict__ grad, float* __restrict__ loss,
                                                   int tiles_per_wave) {
  __shared__ __align__(16) short wb[Q_TOTAL];
  __shared__ float bias[B_TOTAL];
  __shared__ float tbuf[4][16 * 17];
  stage_params(wb, bias, wb_g, p_g);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  float* buf = tbuf[wave];
  f4 dW1[4], dW2[4][4], dW3[2][4];
  float db1[4][4], db2[4][4], db3[2][4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    dW1[a] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      dW2[a][b] = f4{0.f, 0.f, 0.f, 0.f};
      db1[a][b] = db2[a][b] = 0.f;
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      dW3[a][b] = f4{0.f, 0.f, 0.f, 0.f};
      db3[a][b] = 0.f;
    }
  float lsum = 0.f;
  const int64_t wglobal = (int64_t)blockIdx.x * 4 + wave;
  const float inv_b = 1.0f / (float)nbatch;
  for (int tt = 0; tt < tiles_per_wave; ++tt) {
    const int64_t b0 = (wglobal * tiles_per_wave + tt) * 16;
    if (b0 >= nbatch) break;
    const int64_t bs = b0 + j;
    const bool valid = bs < nbatch;
    const int64_t s = valid ? idx[bs] : 0;
    const s4 xb = input_b(x, s, valid, lane);
    Fwd f;
    forward(LdsWeights{wb, bias, lane}, xb, f);
    // dL/dy = 2 (y - t) / B on the 21 real outputs
    f4 d3[2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int o = 16 * t + 4 * g + r;
        float dv = 0.f;
        if (valid && o < 21) {
          const float e = f.y[t][r] - tgt[21 * s + o];
          lsum += e * e;
          dv = 2.f * e * inv_b;
        }
        d3[t][r] = dv;
        db3[t][r] += dv;
      }
    // dW3 (32 x 64) += d3^T (o x s) * H2 (s x k): A = d3 (o rows, s as K), B = H2 (s as K, k cols)
    s4 a3[2], bh2[4];
#pragma unroll
    for (int t = 0; t < 2; ++t) a3[t] = transpose_tile(buf, d3[t], lane);  // lane: o = j, s = 4g..
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) bh2[kk] = transpose_tile_bf(buf, f.h2[kk], lane);  // lane: k = j, s = 4g..
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) dW3[t][kk] = mfma(a3[t], bh2[kk], dW3[t][kk]);
    // d2 = (W3^T d3) * gelu'(z2): A = W3^T (k rows, o as K), B = d3 (o as K, s cols) = D layout of d3
    s4 bd3[2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) bd3[t][r] = f2bf(d3[t][r]);
    f4 d2[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      s4 a[2];
#pragma unroll
      for (int kk = 0; kk < 2; ++kk)
#pragma unroll
        for (int r = 0; r < 4; ++r) a[kk][r] = wb[Q_W3 + (16 * kk + 4 * g + r) * 64 + 16 * t + j];
      f4 acc = mfma2(a[0], a[1], bd3[0], bd3[1], f4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        acc[r] *= gelu_d(f.z2[t][r]);
        db2[t][r] += acc[r];
      }
      d2[t] = acc;
    }
    // dW2 (64 x 64) += d2^T * H1
    s4 a2[4], bh1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) a2[t] = transpose_tile(buf, d2[t], lane);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) bh1[kk] = transpose_tile_bf(buf, f.h1[kk], lane);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) dW2[t][kk] = mfma(a2[t], bh1[kk], dW2[t][kk]);
    // d1 = (W2^T d2) * gelu'(z1)
    s4 bd2[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) bd2[t][r] = f2bf(d2[t][r]);
    f4 d1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      s4 a[4];
#pragma unroll
      for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int r = 0; r < 4; ++r) a[kk][r] = wb[Q_W2 + (16 * kk + 4 * g + r) * 64 + 16 * t + j];
      f4 acc = mfma2(a[0], a[1], bd2[0], bd2[1], f4{0.f, 0.f, 0.f, 0.f});
      acc = mfma2(a[2], a[3], bd2[2], bd2[3], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        acc[r] *= gelu_d(f.z1[t][r]);
        db1[t][r] += acc[r];
      }
      d1[t] = acc;
    }
    // dW1 (64 x 16) += d1^T * X: B = X (s as K, input feature cols): lane k = j, s = 4g..
    s4 bx;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t bsr = b0 + 4 * g + r;
      const bool v2 = bsr < nbatch && j < 9;
      bx[r] = v2 ? f2bf(x[9 * idx[bsr] + j]) : (short)0;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) dW1[t] = mfma(transpose_tile(buf, d1[t], lane), bx, dW1[t]);
  }
  // flush: D layout of dW tiles: lane holds rows 4g+r, column j
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      atomicAdd(grad + P_W1 + (16 * t + 4 * g + r) * 16 + j, dW1[t][r]);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) atomicAdd(grad + P_W2 + (16 * t + 4 * g + r) * 64 + 16 * kk + j, dW2[t][kk][r]);
    }
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) atomicAdd(grad + P_W3 + (16 * t + 4 * g + r) * 64 + 16 * kk + j, dW3[t][kk][r]);
  // bias gradients: sum over the 16 sample lanes (lanes with the same g)
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float a1 = db1[t][r], a2 = db2[t][r];
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) {
        a1 += __shfl_xor(a1, o, 16);
        a2 += __shfl_xor(a2, o, 16);
      }
      if (j == 0) {
        atomicAdd(grad + P_B1 + 16 * t + 4 * g + r, a1);
        atomicAdd(grad + P_B2 + 16 * t + 4 * g + r, a2);
      }
    }
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float a3 = db3[t][r];
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) a3 += __shfl_xor(a3, o, 16);
      if (j == 0) atomicAdd(grad + P_B3 + 16 * t + 4 * g + r, a3);
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
  if (lane == 0) atomicAdd(loss, lsum * inv_b);
}

// Adam on the FP32 master weights (padding rows/columns stay zero), then BF16 copies.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m,
                                                  float* __restrict__ v, short* __restrict__ wb, float lr, float b1,
                                                  float b2, float bc1, float bc2, float eps, int zero_grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P_TOTAL) return;
  bool pad = false;
  if (k < P_B1) pad = (k % 16) >= 9;
  else if (k >= P_W3 && k < P_B3) pad = (k - P_W3) / 64 >= 21;
  else if (k >= P_B3) pad = (k - P_B3) >= 21;
  if (!pad) {
    const float gr = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gr;
    const float vk = b2 * v[k] + (1.f - b2) * gr * gr;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
  }
  if (zero_grad) g[k] = 0.f;
  int q = -1;
  if (k < P_B1) q = Q_W1 + k;
  else if (k >= P_W2 && k < P_B2) q = Q_W2 + (k - P_W2);
  else if (k >= P_W3 && k < P_B3) q = Q_W3 + (k - P_W3);
  if (q >= 0) wb[q] = f2bf(p[k]);
}


__global__ __launch_bounds__(256) void k_mlp_refresh(const float* __restrict__ p, short* __restrict__ wb) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P_TOTAL) return;
  int q = -1;
  if (k < P_B1) q = Q_W1 + k;
  else if (k >= P_W2 && k < P_B2) q = Q_W2 + (k - P_W2);
  else if (k >= P_W3 && k < P_B3) q = Q_W3 + (k - P_W3);
  if (q >= 0) wb[q] = f2bf(p[k]);
}

// ------------------------------------------------------------------ host side
struct MlpHost {
  DevBuf buf;        // p, g, m, v (P_TOTAL floats each) + wb (Q_TOTAL shorts)
  DevBuf loss;       // 1 float
  DevBuf d_x, d_t;   // training set on the device: x[cap][9], t[cap][21] (normalised)
  DevBuf d_idx;      // batch order
  DevBuf d_io;       // scratch for forward / gradient calls
  int64_t n = 0, cap = 0;
  MlpNorm nrm;
  int model = 1;
  double E = 1000, nu = 0.3;
  int64_t adam_t = 0;
  float* p() const { return buf.as<float>(); }
  float* g() const { return buf.as<float>() + P_TOTAL; }
  float* m() const { return buf.as<float>() + 2 * P_TOTAL; }
  float* v() const { return buf.as<float>() + 3 * P_TOTAL; }
  short* wb() const { return reinterpret_cast<short*>(buf.as<float>() + 4 * P_TOTAL); }
};

void mlp_free(Ctx* c) {
  delete c->mlp;
  c->mlp = nullptr;
  c->mlp_ready = false;
}

bool mlp_matches(const Ctx* c, int model, double Emod, double nu) {
  return c->mlp && c->mlp_ready && c->mlp->model == model && c->mlp->E == Emod && c->mlp->nu == nu;
}

int launch_mlp_tangent(Ctx* c, int nq) {
  MlpHost* h = c->mlp;
  if (!h || !c->mlp_ready) return fail(c, FEDD_ERR_STATE, "no surrogate network loaded");
  const int64_t nrec = c->mesh[0].nelem * nq;
  // FEDD_MLP_BPC resident workgroups per CU load the weights once and walk their tiles
  const unsigned grid = (unsigned)std::min<int64_t>((nrec + 63) / 64, (int64_t)c->num_cu * FEDD_MLP_BPC);
  ktimer_begin(c);
  k_mlp_tangent<<<grid, 256, 0, c->stream>>>(c->d_qp.as<double>(), nrec, h->wb(), h->p(), h->nrm);
  FEDD_HIP(c, hipGetLastError());
  ktimer_end(c, "mlp_tangent");
  return FEDD_OK;
}

// deterministic normal samples (splitmix64 + Box-Muller)
static uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double unif(uint64_t& s) { return ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
static double gauss(uint64_t& s) {
  const double u1 = unif(s), u2 = unif(s);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

static void add_sample(const MlpHost* h, const double F[3][3], std::vector<float>& xs, std::vector<float>& ts) {
  double S6[6], CC[21];
  exact_law(h->model, h->E, h->nu, F, S6, CC);
  const double inv_s_in = h->nrm.inv_s_in;
  for (int k = 0; k < 9; ++k) xs.push_back((float)((F[k / 3][k % 3] - (k % 4 == 0 ? 1.0 : 0.0)) * inv_s_in));
  for (int k = 0; k < 21; ++k) ts.push_back((float)((CC[k] - h->nrm.cc0[k]) / h->nrm.s_out));
}

// Grows the device training set to hold `need` samples (contents kept).
static int reserve(Ctx* c, MlpHost* h, int64_t need) {
  if (need <= h->cap) return FEDD_OK;
  const int64_t cap = std::max(need, std::max<int64_t>(2 * h->cap, 65536));
  DevBuf nx, nt;
  int rc;
  if ((rc = alloc(c, nx, cap * 9 * 4))) return rc;
  if ((rc = alloc(c, nt, cap * 21 * 4))) return rc;
  if (h->n) {
    FEDD_HIP(c, hipMemcpyAsync(nx.p, h->d_x.p, h->n * 9 * 4, hipMemcpyDeviceToDevice, c->stream));
    FEDD_HIP(c, hipMemcpyAsync(nt.p, h->d_t.p, h->n * 21 * 4, hipMemcpyDeviceToDevice, c->stream));
    FEDD_HIP(c, hipStreamSynchronize(c->stream));
  }
  std::swap(nx.p, h->d_x.p);
  std::swap(nx.bytes, h->d_x.bytes);
  std::swap(nt.p, h->d_t.p);
  std::swap(nt.bytes, h->d_t.bytes);
  h->cap = cap;
  return FEDD_OK;
}

// Appends the deformation gradients of quadrature-point records (and their
// exact-law tangents as targets) to the training set.
__global__ __launch_bounds__(256) void k_mlp_collect(const double* __restrict__ qp, int64_t nrec, int model,
                                                     double Emod, double nu, MlpNorm nrm, float* __restrict__ x,
                                                     float* __restrict__ t) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nrec) return;
  double F[3][3];
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k / 3][k % 3] = qp[52 * i + k];
  double S6[6], CC[21];
  exact_law(model, Emod, nu, F, S6, CC);
#pragma unroll
  for (int k = 0; k < 9; ++k) x[9 * i + k] = (float)((F[k / 3][k % 3] - (k % 4 == 0 ? 1.0 : 0.0)) * nrm.inv_s_in);
#pragma unroll
  for (int k = 0; k < 21; ++k) t[21 * i + k] = (float)((CC[k] - (double)nrm.cc0[k]) / (double)nrm.s_out);
}

int mlp_collect_qp(Ctx* c) {
  MlpHost* h = c->mlp;
  if (!h) return fail(c, FEDD_ERR_STATE, "no surrogate network");
  const int64_t nrec = c->mesh[0].nelem * c->qp_nq;
  if (c->qp_nq == 0 || nrec == 0) return fail(c, FEDD_ERR_STATE, "no quadrature-point records");
  int rc = reserve(c, h, h->n + nrec);
  if (rc) return rc;
  k_mlp_collect<<<(unsigned)((nrec + 255) / 256), 256, 0, c->stream>>>(
      c->d_qp.as<double>(), nrec, h->model, h->E, h->nu, h->nrm, h->d_x.as<float>() + 9 * h->n,
      h->d_t.as<float>() + 21 * h->n);
  FEDD_HIP(c, hipGetLastError());
  h->n += nrec;
  return FEDD_OK;
}

}  // namespace fedd

using namespace fedd;

extern "C" {

int fedd_mlp_init(fedd_ctx* c, int neo_hooke_model, double E, double nu, uint64_t seed) {
  if (!c || !(E > 0) || !(nu > -1 && nu < 0.5)) return FEDD_ERR_ARG;
  (void)hipSetDevice(c->device);
  mlp_free(c);
  MlpHost* h = new MlpHost();
  c->mlp = h;
  h->model = neo_hooke_model == 1 ? 1 : 2;
  h->E = E;
  h->nu = nu;
  const double mu = E / (2 * (1 + nu)), lam = 2 * mu * nu / (1 - 2 * nu);
  double I[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, S6[6], CC[21];
  exact_law(h->model, E, nu, I, S6, CC);
  for (int k = 0; k < 21; ++k) h->nrm.cc0[k] = (float)CC[k];
  h->nrm.s_out = (float)(lam + 2 * mu);
  h->nrm.inv_s_in = 1.0 / 0.1;
  int rc = alloc(c, h->buf, 4 * P_TOTAL * sizeof(float) + Q_TOTAL * sizeof(short));
  if (rc) return rc;
  if ((rc = alloc(c, h->loss, 64))) return rc;
  // Xavier-normal weights, zero biases; padding rows/columns zero
  std::vector<float> p(4 * P_TOTAL, 0.f);
  uint64_t s = seed * 0x2545F4914F6CDD1Dull + 1;
  auto fill = [&](int off, int rows, int cols, int ld, int fan_in, int fan_out) {
    const double sd = std::sqrt(2.0 / (fan_in + fan_out));
    for (int r = 0; r < rows; ++r)
      for (int k = 0; k < cols; ++k) p[off + r * ld + k] = (float)(sd * gauss(s));
  };
  fill(P_W1, 64, 9, 16, 9, 64);
  fill(P_W2, 64, 64, 64, 64, 64);
  fill(P_W3, 21, 64, 64, 64, 21);
  FEDD_HIP(c, hipMemcpyAsync(h->buf.p, p.data(), p.size() * 4, hipMemcpyHostToDevice, c->stream));
  k_mlp_refresh<<<(P_TOTAL + 255) / 256, 256, 0, c->stream>>>(h->p(), h->wb());
  FEDD_HIP(c, hipGetLastError());
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  h->adam_t = 0;
  c->mlp_ready = true;
  return FEDD_OK;
}

int fedd_mlp_add_samples_random(fedd_ctx* c, int64_t n, double sigma, uint64_t seed) {
  if (!c || !c->mlp || n < 0 || !(sigma >= 0)) return FEDD_ERR_ARG;
  MlpHost* h = c->mlp;
  (void)hipSetDevice(c->device);
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 7;
  std::vector<float> xs, ts;
  xs.reserve(n * 9);
  ts.reserve(n * 21);
  int64_t added = 0;
  while (added < n) {
    double F[3][3];
    for (int i = 0; i < 3; ++i)
      for (int A = 0; A < 3; ++A) F[i][A] = (i == A ? 1.0 : 0.0) + sigma * gauss(s);
    const double J = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
                     F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
    if (!(J > 0.2)) continue;  // the laws are defined for J > 0; keep physically admissible samples
    add_sample(h, F, xs, ts);
    ++added;
  }
  int rc = reserve(c, h, h->n + n);
  if (rc) return rc;
  FEDD_HIP(c, hipMemcpyAsync(h->d_x.as<float>() + 9 * h->n, xs.data(), xs.size() * 4, hipMemcpyHostToDevice,
                             c->stream));
  FEDD_HIP(c, hipMemcpyAsync(h->d_t.as<float>() + 21 * h->n, ts.data(), ts.size() * 4, hipMemcpyHostToDevice,
                             c->stream));
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  h->n += n;
  return FEDD_OK;
}

int fedd_mlp_add_samples_qp(fedd_ctx* c) {
  if (!c || !c->mlp) return FEDD_ERR_ARG;
  (void)hipSetDevice(c->device);
  return mlp_collect_qp(c);
}

int fedd_mlp_num_samples(fedd_ctx* c, int64_t* n) {
  if (!c || !c->mlp || !n) return FEDD_ERR_ARG;
  *n = c->mlp->n;
  return FEDD_OK;
}

static int train_step(Ctx* c, const int64_t* d_idx, int64_t nb, float lr, int64_t t) {
  MlpHost* h = c->mlp;
  const int tiles_per_wave = FEDD_MLP_TPW;
  const int64_t ntile = (nb + 15) / 16;
  const unsigned grid = (unsigned)((ntile + 4 * tiles_per_wave - 1) / (4 * tiles_per_wave));
  k_mlp_train<<<grid, 256, 0, c->stream>>>(h->d_x.as<float>(), h->d_t.as<float>(), d_idx, nb, h->wb(), h->p(),
                                          h->g(), h->loss.as<float>(), tiles_per_wave);
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.f - std::pow(b1, (float)t), bc2 = 1.f - std::pow(b2, (float)t);
  k_mlp_adam<<<(P_TOTAL + 255) / 256, 256, 0, c->stream>>>(h->p(), h->g(), h->m(), h->v(), h->wb(), lr, b1, b2, bc1,
                                                          bc2, 1e-8f, 1);
  FEDD_HIP(c, hipGetLastError());
  return FEDD_OK;
}

int fedd_mlp_train(fedd_ctx* c, int steps, int batch, float lr, uint64_t seed, float* loss_history) {
  if (!c || !c->mlp || steps < 0 || batch < 16) return FEDD_ERR_ARG;
  MlpHost* h = c->mlp;
  (void)hipSetDevice(c->device);
  const int64_t n = h->n;
  if (n < batch) return fail(c, FEDD_ERR_STATE, "fewer training samples than one batch");
  int rc;
  // one random permutation per epoch, consecutive batches from it
  std::vector<int64_t> perm(n);
  uint64_t s = seed * 0xD1B54A32D192ED03ull + 3;
  int64_t pos = n;
  if ((rc = alloc(c, h->d_idx, n * 8))) return rc;
  FEDD_HIP(c, hipMemsetAsync(h->g(), 0, P_TOTAL * 4, c->stream));
  for (int st = 0; st < steps; ++st) {
    if (pos + batch > n) {
      for (int64_t k = 0; k < n; ++k) perm[k] = k;
      for (int64_t k = n - 1; k > 0; --k) std::swap(perm[k], perm[splitmix(s) % (k + 1)]);
      FEDD_HIP(c, hipMemcpyAsync(h->d_idx.p, perm.data(), n * 8, hipMemcpyHostToDevice, c->stream));
      pos = 0;
    }
    FEDD_HIP(c, hipMemsetAsync(h->loss.p, 0, 4, c->stream));
    // cosine decay from lr to lr / 50 over the call
    const float lr_t = lr * (0.02f + 0.49f * (1.f + std::cos(3.14159265f * (float)st / (float)steps)));
    if ((rc = train_step(c, h->d_idx.as<int64_t>() + pos, batch, lr_t, ++h->adam_t))) return rc;
    pos += batch;
    if (loss_history) {
      float l = 0.f;
      FEDD_HIP(c, hipMemcpyAsync(&l, h->loss.p, 4, hipMemcpyDeviceToHost, c->stream));
      FEDD_HIP(c, hipStreamSynchronize(c->stream));
      loss_history[st] = l;
    }
  }
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_mlp_params(fedd_ctx* c, float* out) {
  if (!c || !c->mlp || !out) return FEDD_ERR_ARG;
  (void)hipSetDevice(c->device);
  FEDD_HIP(c, hipMemcpyAsync(out, c->mlp->p(), P_TOTAL * 4, hipMemcpyDeviceToHost, c->stream));
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_mlp_set_params(fedd_ctx* c, const float* in) {
  if (!c || !c->mlp || !in) return FEDD_ERR_ARG;
  (void)hipSetDevice(c->device);
  FEDD_HIP(c, hipMemcpyAsync(c->mlp->p(), in, P_TOTAL * 4, hipMemcpyHostToDevice, c->stream));
  k_mlp_refresh<<<(P_TOTAL + 255) / 256, 256, 0, c->stream>>>(c->mlp->p(), c->mlp->wb());
  FEDD_HIP(c, hipGetLastError());
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_mlp_forward(fedd_ctx* c, const float* x, int64_t n, float* y) {
  if (!c || !c->mlp || !x || !y || n < 1) return FEDD_ERR_ARG;
  MlpHost* h = c->mlp;
  (void)hipSetDevice(c->device);
  int rc = alloc(c, h->d_io, n * 30 * 4);
  if (rc) return rc;
  float* dx = h->d_io.as<float>();
  float* dy = dx + 9 * n;
  FEDD_HIP(c, hipMemcpyAsync(dx, x, n * 9 * 4, hipMemcpyHostToDevice, c->stream));
  const unsigned grid = (unsigned)std::min<int64_t>((n + 63) / 64, 4096);
  k_mlp_forward_rows<<<grid, 256, 0, c->stream>>>(dx, n, h->wb(), h->p(), dy);
  FEDD_HIP(c, hipGetLastError());
  FEDD_HIP(c, hipMemcpyAsync(y, dy, n * 21 * 4, hipMemcpyDeviceToHost, c->stream));
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_mlp_gradient(fedd_ctx* c, const float* x, const float* t, int64_t n, float* grad, float* loss) {
  if (!c || !c->mlp || !x || !t || !grad || n < 1) return FEDD_ERR_ARG;
  MlpHost* h = c->mlp;
  (void)hipSetDevice(c->device);
  int rc;
  if ((rc = alloc(c, h->d_io, n * 30 * 4 + n * 8))) return rc;
  float* dx = h->d_io.as<float>();
  float* dt = dx + 9 * n;
  int64_t* didx = reinterpret_cast<int64_t*>(dx + 30 * n);
  std::vector<int64_t> idx(n);
  for (int64_t k = 0; k < n; ++k) idx[k] = k;
  FEDD_HIP(c, hipMemcpyAsync(dx, x, n * 9 * 4, hipMemcpyHostToDevice, c->stream));
  FEDD_HIP(c, hipMemcpyAsync(dt, t, n * 21 * 4, hipMemcpyHostToDevice, c->stream));
  FEDD_HIP(c, hipMemcpyAsync(didx, idx.data(), n * 8, hipMemcpyHostToDevice, c->stream));
  FEDD_HIP(c, hipMemsetAsync(h->g(), 0, P_TOTAL * 4, c->stream));
  FEDD_HIP(c, hipMemsetAsync(h->loss.p, 0, 4, c->stream));
  const int tiles_per_wave = FEDD_MLP_TPW;
  const int64_t ntile = (n + 15) / 16;
  const unsigned grid = (unsigned)((ntile + 4 * tiles_per_wave - 1) / (4 * tiles_per_wave));
  k_mlp_train<<<grid, 256, 0, c->stream>>>(dx, dt, didx, n, h->wb(), h->p(), h->g(), h->loss.as<float>(),
                                          tiles_per_wave);
  FEDD_HIP(c, hipGetLastError());
  FEDD_HIP(c, hipMemcpyAsync(grad, h->g(), P_TOTAL * 4, hipMemcpyDeviceToHost, c->stream));
  if (loss) FEDD_HIP(c, hipMemcpyAsync(loss, h->loss.p, 4, hipMemcpyDeviceToHost, c->stream));
  FEDD_HIP(c, hipStreamSynchronize(c->stream));
  return FEDD_OK;
}

int fedd_mlp_norm(fedd_ctx* c, float* cc0, float* s_out, double* s_in) {
  if (!c || !c->mlp) return FEDD_ERR_ARG;
  if (cc0) std::copy(c->mlp->nrm.cc0, c->mlp->nrm.cc0 + 21, cc0);
  if (s_out) *s_out = c->mlp->nrm.s_out;
  if (s_in) *s_in = 1.0 / c->mlp->nrm.inv_s_in;
  return FEDD_OK;
}

}  // extern "C"
