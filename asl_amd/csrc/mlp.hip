// Training kernels of the closure network (7-32-32-1, ReLU; parameter
// layout and model in closure.hpp): forward + backward of a batch with
// deterministic per-block gradient sums, the reduction over blocks, and Adam.
//
// Loss: half the mean squared error over the batch.  Adam: Kingma & Ba,
// "Adam: A Method for Stochastic Optimization" (ICLR 2015), Algorithm 1.
// All FP32 except the cross-block gradient and loss sums (FP64).
#include <hip/hip_runtime.h>

#include "closure.hpp"
#include "mlp.hpp"

namespace aslhip {
namespace {

constexpr int GB = MLP_GRAD_BLOCK;  // samples per workgroup
constexpr int ROW = 7 + 4 * CL_H + 2;  // per sample in LDS: xn, h1, dh1, h2, dh2, dy, r^2 / 2
constexpr int RDY = ROW - 2, RLOSS = ROW - 1;

// splitmix64 finaliser (S. Vigna, public domain): sample indices of a batch.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(GB) void k_mlp_grad(const MlpGradArgs a) {
  __shared__ float sh[ROW][GB];
  const int k = blockIdx.x * GB + threadIdx.x;  // sample within the batch
  const bool live = k < a.count;
  long long s = 0;
  if (live)
    s = a.seed ? static_cast<long long>(mix64(a.seed + (static_cast<unsigned long long>(k) + 1ull) *
                                                          0x9E3779B97F4A7C15ull) %
                                        static_cast<unsigned long long>(a.n))
               : a.first + k;
  const float *p = a.p;
  float xn[7], h1[CL_H], a1m[CL_H], h2[CL_H], dh2[CL_H];
#pragma unroll
  for (int j = 0; j < 7; ++j) xn[j] = live ? a.x[s * 7 + j] : 0.f;
#pragma unroll
  for (int o = 0; o < CL_H; ++o) {
    float v = p[CL_B1 + o];
#pragma unroll
    for (int j = 0; j < 7; ++j) v = __builtin_fmaf(p[CL_W1 + o * 7 + j], xn[j], v);
    a1m[o] = v > 0.f ? 1.f : 0.f;
    h1[o] = v > 0.f ? v : 0.f;
  }
  float y = p[CL_B3];
#pragma unroll
  for (int o = 0; o < CL_H; ++o) {
    float v = p[CL_B2 + o];
#pragma unroll
    for (int j = 0; j < CL_H; ++j) v = __builtin_fmaf(p[CL_W2 + o * CL_H + j], h1[j], v);
    h2[o] = v > 0.f ? v : 0.f;
    dh2[o] = v > 0.f ? 1.f : 0.f;  // mask for now
    y = __builtin_fmaf(p[CL_W3 + o], h2[o], y);
  }
  const float r = live ? y - a.t[s] : 0.f;
  const float dy = r * a.inv_batch;
#pragma unroll
  for (int o = 0; o < CL_H; ++o) dh2[o] *= dy * p[CL_W3 + o];
  float dh1[CL_H];
#pragma unroll
  for (int j = 0; j < CL_H; ++j) {
    float v = 0.f;
#pragma unroll
    for (int o = 0; o < CL_H; ++o) v = __builtin_fmaf(p[CL_W2 + o * CL_H + j], dh2[o], v);
    dh1[j] = v * a1m[j];
  }
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < 7; ++j) sh[j][t] = xn[j];
#pragma unroll
  for (int o = 0; o < CL_H; ++o) {
    sh[7 + o][t] = h1[o];
    sh[7 + CL_H + o][t] = dh1[o];
    sh[7 + 2 * CL_H + o][t] = h2[o];
    sh[7 + 3 * CL_H + o][t] = dh2[o];
  }
  sh[RDY][t] = dy;
  sh[RLOSS][t] = 0.5f * r * r;
  __syncthreads();
  // parameter q of this block: sum over the block's samples of u * v
  float *out = a.partial + static_cast<long long>(blockIdx.x) * MLP_PSTRIDE;
  for (int q = t; q < CL_NPARAM; q += GB) {
    int ru, rv;  // LDS rows of the two factors (rv < 0: factor 1)
    if (q < CL_B1) {
      ru = 7 + CL_H + q / 7;  // dh1[o]
      rv = q % 7;             // xn[j]
    } else if (q < CL_W2) {
      ru = 7 + CL_H + (q - CL_B1);
      rv = -1;
    } else if (q < CL_B2) {
      ru = 7 + 3 * CL_H + (q - CL_W2) / CL_H;  // dh2[o]
      rv = 7 + (q - CL_W2) % CL_H;             // h1[j]
    } else if (q < CL_W3) {
      ru = 7 + 3 * CL_H + (q - CL_B2);
      rv = -1;
    } else if (q < CL_B3) {
      ru = RDY;
      rv = 7 + 2 * CL_H + (q - CL_W3);  // h2[j]
    } else {
      ru = RDY;
      rv = -1;
    }
    float acc = 0.f;
    for (int i = 0; i < GB; ++i) acc = __builtin_fmaf(sh[ru][i], rv >= 0 ? sh[rv][i] : 1.f, acc);
    out[q] = acc;
  }
  if (t == 0) {
    float l = 0.f;
    for (int i = 0; i < GB; ++i) l += sh[RLOSS][i];
    out[CL_NPARAM] = l;
  }
}

// grad[q] = sum over blocks of partial[b][q] (FP64, in block order); the last
// entry is the loss (0.5 * sum r^2 / count).
__global__ __launch_bounds__(256) void k_mlp_reduce(const float *partial, int nblocks, double inv_count, float *grad,
                                                    double *loss) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q > CL_NPARAM) return;
  double acc = 0.;
  for (int b = 0; b < nblocks; ++b) acc += static_cast<double>(partial[static_cast<long long>(b) * MLP_PSTRIDE + q]);
  if (q < CL_NPARAM)
    grad[q] = static_cast<float>(acc);
  else
    *loss = acc * inv_count;
}

__global__ __launch_bounds__(256) void k_adam(float *p, const float *g, float *m, float *v, AdamArgs a) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= CL_NPARAM) return;
  const float gq = g[q];
  const float mq = __builtin_fmaf(a.beta1, m[q], (1.f - a.beta1) * gq);
  const float vq = __builtin_fmaf(a.beta2, v[q], (1.f - a.beta2) * gq * gq);
  m[q] = mq;
  v[q] = vq;
  const float mhat = mq / a.bc1;
  const float vhat = vq / a.bc2;
  p[q] = p[q] - a.lr * mhat / (__builtin_sqrtf(vhat) + a.eps);
}

}  // namespace

hipError_t launch_mlp_grad(const MlpGradArgs &a, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(GB), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_mlp_reduce(const float *partial, int nblocks, double inv_count, float *grad, double *loss,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_reduce, dim3((CL_NPARAM + 1 + 255) / 256), dim3(256), 0, s, partial, nblocks, inv_count,
                     grad, loss);
  return hipGetLastError();
}

hipError_t launch_adam(float *p, const float *g, float *m, float *v, const AdamArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3((CL_NPARAM + 255) / 256), dim3(256), 0, s, p, g, m, v, a);
  return hipGetLastError();
}

}  // namespace aslhip
