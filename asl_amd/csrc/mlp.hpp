// Launchers of the closure-network training kernels (mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace aslhip {

constexpr int MLP_GRAD_BLOCK = 64;    // samples per gradient workgroup
constexpr int MLP_PSTRIDE = 1346;     // floats per block partial: 1345 gradients + loss

struct MlpGradArgs {
  const float *x;        // n x 7 normalised inputs
  const float *t;        // n targets
  long long n;           // samples in the data set
  long long first;       // seed == 0: the batch is samples first .. first+count-1
  unsigned long long seed;  // else: sample k of the batch is mix64-hashed from (seed, k)
  int count;             // samples in the batch
  float inv_batch;       // 1 / count
  const float *p;        // parameters (FP32)
  float *partial;        // nblocks x MLP_PSTRIDE
};

struct AdamArgs {
  float lr, beta1, beta2, eps;
  float bc1, bc2;  // 1 - beta1^t, 1 - beta2^t
};

hipError_t launch_mlp_grad(const MlpGradArgs &a, int nblocks, hipStream_t s);
hipError_t launch_mlp_reduce(const float *partial, int nblocks, double inv_count, float *grad, double *loss,
                             hipStream_t s);
hipError_t launch_adam(float *p, const float *g, float *m, float *v, const AdamArgs &a, hipStream_t s);

}  // namespace aslhip
