// Learned closure: per-cell relaxation-time correction from a small network.
//
// An extension with no counterpart in ASL.  Inputs (7 per cell): the
// non-equilibrium second moments of the pre-collision populations,
//   Pi_ab = sum_i w_i c_ia c_ib f_i - (rho/3 delta_ab + m_a m_b / rho),
// for ab = xx, yy, zz, xy, xz, yz, and rho - 1.  Network 7-32-32-1, ReLU
// after both hidden layers, BF16-rounded parameters evaluated in FP32:
//   xn = (x - mu) * sc,  h1 = relu(W1 xn + b1),  h2 = relu(W2 h1 + b2),
//   y = w3 . h2 + b3,    tau = max(tau0 + y * out, 1/2 + (tau0 - 1/2) / 2),  w = 1 / tau
// (the correction may lower the viscosity by at most half).
// Every sum is an fma chain in index order (the CPU oracle evaluates the same
// chain, so results agree bit for bit).
#pragma once

#include "lattice_consts.hpp"

namespace aslhip {

constexpr int CL_NIN = 7, CL_H = 32, CL_NPARAM = 1345;
constexpr int CL_W1 = 0, CL_B1 = 224, CL_W2 = 256, CL_B2 = 1280, CL_W3 = 1312, CL_B3 = 1344;

// Closure parameters as the step kernel reads them (device memory).  The
// matrix-core path reads its operands pre-arranged in lane order (filled by
// pack_closure from p): one 8-byte load per BF16 A operand and 16-byte loads
// for the biases and output weights, instead of one load and one conversion
// per element.
struct ClosureParams {
  float p[CL_NPARAM];  // BF16-rounded weights and biases
  float mu[CL_NIN], sc[CL_NIN];
  float out;           // scale of the relaxation-time correction
  alignas(16) unsigned long long a1[64];  // lane l: W1'[l % 32][4 (l / 32) + j], j < 4 (BF16 x 4)
  unsigned long long a2[4][64];           // k-block kb, lane l: W2[l % 32][8 kb + 4 (l / 32) + j]
  alignas(16) float b2[2][16];            // half h, register r: b2[8 (r / 4) + 4 h + r % 4]
  alignas(16) float w3[2][16];            // half h, register r: w3[8 (r / 4) + 4 h + r % 4]
};

// BF16 bits of an already BF16-rounded float (the top half of its bits).
inline unsigned long long bf16x4_pack(const float *v) {
  unsigned long long r = 0;
  for (int j = 0; j < 4; ++j) {
    unsigned u;
    __builtin_memcpy(&u, &v[j], 4);
    r |= static_cast<unsigned long long>(u >> 16) << (16 * j);
  }
  return r;
}

// Arranges the matrix-core operands of c.p (already BF16-rounded).
inline void pack_closure(ClosureParams &c) {
  for (int l = 0; l < 64; ++l) {
    const int col = l & 31, half = l >> 5;
    float v[4];
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * half + j;
      v[j] = k < CL_NIN ? c.p[CL_W1 + col * CL_NIN + k] : c.p[CL_B1 + col];
    }
    c.a1[l] = bf16x4_pack(v);
    for (int kb = 0; kb < 4; ++kb) {
      for (int j = 0; j < 4; ++j) v[j] = c.p[CL_W2 + col * CL_H + 8 * kb + 4 * half + j];
      c.a2[kb][l] = bf16x4_pack(v);
    }
  }
  for (int h = 0; h < 2; ++h)
    for (int r = 0; r < 16; ++r) {
      c.b2[h][r] = c.p[CL_B2 + 8 * (r / 4) + 4 * h + (r % 4)];
      c.w3[h][r] = c.p[CL_W3 + 8 * (r / 4) + 4 * h + (r % 4)];
    }
}

template <int Q>
__device__ __forceinline__ void closure_features(const float (&f)[Q], float r, const float (&m)[3], float (&x)[7]) {
  constexpr int A[6] = {0, 1, 2, 0, 0, 1}, B[6] = {0, 1, 2, 1, 2, 2};
  float mr[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) mr[d] = m[d] / r;
  const float third = 1.f / 3.f;
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    float acc = 0.f;
    bool first = true;
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      const int cc = Lat<Q>::c[j][A[p]] * Lat<Q>::c[j][B[p]];
      if (cc == 0) continue;
      const float wcc = static_cast<float>(Lat<Q>::w(j) * cc);
      acc = first ? f[j] * wcc : __builtin_fmaf(f[j], wcc, acc);
      first = false;
    }
    const float t = A[p] == B[p] ? __builtin_fmaf(m[A[p]], mr[B[p]], r * third) : m[A[p]] * mr[B[p]];
    x[p] = acc - t;
  }
  x[6] = r - 1.f;
}

__device__ __forceinline__ float closure_relax(float y, const ClosureParams &c, float tau0) {
  const float tau = tau0 + y * c.out;
  const float lo = 0.5f + 0.5f * (tau0 - 0.5f);
  return 1.f / (tau < lo ? lo : tau);
}

// Round to the nearest BF16 (ties to even); the bits of the result.  gfx950
// converts in hardware (v_cvt_pk_bf16_f32, two values per instruction).
__device__ __forceinline__ unsigned short bf16_bits(float x) {
  const __bf16 b = static_cast<__bf16>(x);
  unsigned short u;
  __builtin_memcpy(&u, &b, 2);
  return u;
}

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// The network on the matrix cores: every lane of the wavefront brings its own
// cell's inputs, all 64 lanes must be active.  Two column blocks of 32 cells
// (cells 0-31, 32-63 of the wavefront) go through
//   H1^T = relu(W1' xn')   one 32x32x8 BF16 MFMA (xn' = (xn, 1), W1' = (W1, b1)),
//   H2^T = relu(W2 H1^T + b2)   four 32x32x8 BF16 MFMAs, FP32 accumulation,
//   y = w3 . h2 + b3       FP32: each lane sums the 16 hidden units it holds,
//                          the two halves of the wavefront are added.
// v_mfma_f32_32x32x8bf16_1k operand layout (lane l): A[l % 32][4 (l / 32) + j],
// B[4 (l / 32) + j][l % 32], j < 4; result register r holds
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32] — the layout the next layer's B
// operand needs for k-block r / 4, so layers chain without shuffles.
__device__ __forceinline__ float closure_y_mfma(const float (&x)[7], const ClosureParams &c) {
  const int l = static_cast<int>(threadIdx.x) & 63;
  const int half = l >> 5;
  // own inputs as BF16 pairs: (x0, x1), (x2, x3), (x4, x5), (x6, 1.0 = the bias column of W1')
  unsigned own[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const unsigned lo = bf16_bits((x[2 * p] - c.mu[2 * p]) * c.sc[2 * p]);
    const unsigned hi = p < 3 ? bf16_bits((x[2 * p + 1] - c.mu[2 * p + 1]) * c.sc[2 * p + 1]) : 0x3f80u;
    own[p] = lo | (hi << 16);
  }
  // inputs 4 half .. 4 half + 3 of the partner lane's cell (lane l ^ 32)
  const unsigned send0 = half ? own[0] : own[2], send1 = half ? own[1] : own[3];
  const unsigned par0 = static_cast<unsigned>(__shfl_xor(static_cast<int>(send0), 32));
  const unsigned par1 = static_cast<unsigned>(__shfl_xor(static_cast<int>(send1), 32));
  bf16x4_t a1;
  __builtin_memcpy(&a1, &c.a1[l], 8);  // weights in lane order (pack_closure)
  float y[2];
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
    // column block cb holds cells 32 cb + col; this lane supplies inputs 4 half .. 4 half + 3
    const bool mine = (half == cb);  // the cell 32 cb + col is this lane's own
    const unsigned b0 = mine ? own[2 * half] : par0, b1w = mine ? own[2 * half + 1] : par1;
    bf16x4_t b1;
    b1[0] = static_cast<short>(b0 & 0xffffu);
    b1[1] = static_cast<short>(b0 >> 16);
    b1[2] = static_cast<short>(b1w & 0xffffu);
    b1[3] = static_cast<short>(b1w >> 16);
    f32x16_t d1 = {};
    d1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1, b1, d1, 0, 0, 0);
    // relu and BF16 rounding of the hidden layer first: d1 is dead before the second layer starts
    bf16x4_t hb[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float h = d1[4 * kb + j];
        hb[kb][j] = static_cast<short>(bf16_bits(h > 0.f ? h : 0.f));
      }
    f32x16_t d2;
    __builtin_memcpy(&d2, c.b2[half], 64);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      bf16x4_t a2;
      __builtin_memcpy(&a2, &c.a2[kb][l], 8);
      d2 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a2, hb[kb], d2, 0, 0, 0);
    }
    float part = 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const float h = d2[r] > 0.f ? d2[r] : 0.f;
      part = __builtin_fmaf(c.w3[half][r], h, part);
    }
    const float other = __shfl_xor(part, 32);
    y[cb] = (half == 0 ? part + other : other + part) + c.p[CL_B3];
  }
  return y[half];
}

// tau0: the plain relaxation time 3 nu + 1/2 (float).
__device__ __forceinline__ float closure_w(const float (&x)[7], const ClosureParams &c, float tau0) {
  float xn[7], h1[CL_H];
#pragma unroll
  for (int j = 0; j < 7; ++j) xn[j] = (x[j] - c.mu[j]) * c.sc[j];
#pragma unroll
  for (int o = 0; o < CL_H; ++o) {
    float a = c.p[CL_B1 + o];
#pragma unroll
    for (int j = 0; j < 7; ++j) a = __builtin_fmaf(c.p[CL_W1 + o * 7 + j], xn[j], a);
    h1[o] = a > 0.f ? a : 0.f;
  }
  float y = c.p[CL_B3];
#pragma unroll
  for (int o = 0; o < CL_H; ++o) {
    float a = c.p[CL_B2 + o];
#pragma unroll
    for (int j = 0; j < CL_H; ++j) a = __builtin_fmaf(c.p[CL_W2 + o * CL_H + j], h1[j], a);
    const float h2 = a > 0.f ? a : 0.f;
    y = __builtin_fmaf(c.p[CL_W3 + o], h2, y);
  }
  return closure_relax(y, c, tau0);
}

}  // namespace aslhip
