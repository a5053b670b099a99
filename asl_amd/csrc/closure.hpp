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

// Closure parameters as the step kernel reads them (device memory).
struct ClosureParams {
  float p[CL_NPARAM];  // BF16-rounded weights and biases
  float mu[CL_NIN], sc[CL_NIN];
  float out;           // scale of the relaxation-time correction
};

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
  const float tau = tau0 + y * c.out;
  const float lo = 0.5f + 0.5f * (tau0 - 0.5f);
  return 1.f / (tau < lo ? lo : tau);
}

}  // namespace aslhip
