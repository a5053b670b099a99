// C ABI of the closure-network trainer (include/aslhip.h, "learned closure").
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/aslhip.h"
#include "closure.hpp"
#include "mlp.hpp"
#include "status.hpp"

using namespace aslhip;

struct aslhip_mlp {
  hipStream_t stream = nullptr;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  float *x = nullptr, *t = nullptr;
  long long n = 0;
  float *partial = nullptr;
  int partial_blocks = 0;
  double *loss = nullptr;
  long long adam_t = 0;  // Adam steps taken
};

namespace {

void mlp_free(aslhip_mlp *m) {
  void *ptrs[] = {m->p, m->g, m->m, m->v, m->x, m->t, m->partial, m->loss};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (m->stream) (void)hipStreamDestroy(m->stream);
}

int ensure_partial(aslhip_mlp *m, int nblocks) {
  if (nblocks <= m->partial_blocks) return 0;
  HIPCHK(hipStreamSynchronize(m->stream));
  if (m->partial) HIPCHK(hipFree(m->partial));
  m->partial = nullptr;
  HIPCHK(hipMalloc(&m->partial, (size_t)nblocks * MLP_PSTRIDE * sizeof(float)));
  m->partial_blocks = nblocks;
  return 0;
}

unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Gradient of the batch into m->g and its loss into m->loss (device).
int batch_gradient(aslhip_mlp *m, long long first, int count, unsigned long long seed) {
  const int nb = (count + MLP_GRAD_BLOCK - 1) / MLP_GRAD_BLOCK;
  if (int r = ensure_partial(m, nb)) return r;
  MlpGradArgs a;
  a.x = m->x;
  a.t = m->t;
  a.n = m->n;
  a.first = first;
  a.seed = seed;
  a.count = count;
  a.inv_batch = 1.f / static_cast<float>(count);
  a.p = m->p;
  a.partial = m->partial;
  HIPCHK(launch_mlp_grad(a, nb, m->stream));
  HIPCHK(launch_mlp_reduce(m->partial, nb, 1.0 / count, m->g, m->loss, m->stream));
  return 0;
}

}  // namespace

extern "C" {

int aslhip_mlp_create(aslhip_mlp **out, const float *params) {
  if (!out || !params) return fail(ASLHIP_EINVAL, "ASL ERROR: mlp_create needs an output handle and parameters");
  *out = nullptr;
  for (int i = 0; i < CL_NPARAM; ++i)
    if (!std::isfinite(params[i])) return fail(ASLHIP_EINVAL, "ASL ERROR: mlp_create: non-finite parameter %d", i);
  aslhip_mlp *m = new aslhip_mlp();
  const size_t pb = CL_NPARAM * sizeof(float);
  hipError_t e = hipStreamCreate(&m->stream);
  if (e == hipSuccess) e = hipMalloc(&m->p, pb);
  if (e == hipSuccess) e = hipMalloc(&m->g, pb);
  if (e == hipSuccess) e = hipMalloc(&m->m, pb);
  if (e == hipSuccess) e = hipMalloc(&m->v, pb);
  if (e == hipSuccess) e = hipMalloc(&m->loss, sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(m->p, params, pb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(m->m, 0, pb);
  if (e == hipSuccess) e = hipMemset(m->v, 0, pb);
  if (e != hipSuccess) {
    mlp_free(m);
    delete m;
    return fail(ASLHIP_EHIP, "ASL ERROR: mlp_create: %s", hipGetErrorString(e));
  }
  *out = m;
  return 0;
}

int aslhip_mlp_destroy(aslhip_mlp *m) {
  if (!m) return 0;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  mlp_free(m);
  delete m;
  return 0;
}

int aslhip_mlp_set_data(aslhip_mlp *m, const float *x, const float *t, long long n) {
  if (!m || !x || !t || n < 1) return fail(ASLHIP_EINVAL, "ASL ERROR: mlp_set_data needs data");
  HIPCHK(hipStreamSynchronize(m->stream));
  if (m->x) HIPCHK(hipFree(m->x));
  if (m->t) HIPCHK(hipFree(m->t));
  m->x = m->t = nullptr;
  m->n = 0;
  HIPCHK(hipMalloc(&m->x, (size_t)n * CL_NIN * sizeof(float)));
  HIPCHK(hipMalloc(&m->t, (size_t)n * sizeof(float)));
  HIPCHK(hipMemcpy(m->x, x, (size_t)n * CL_NIN * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->t, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  m->n = n;
  return 0;
}

int aslhip_mlp_gradient(aslhip_mlp *m, long long first, int count, float *grad, double *loss) {
  if (!m || !m->n) return fail(ASLHIP_ESTATE, "ASL ERROR: mlp_gradient before mlp_set_data");
  if (first < 0 || count < 1 || first + count > m->n)
    return fail(ASLHIP_EINVAL, "ASL ERROR: mlp_gradient: samples %lld..%lld outside the data set of %lld", first,
                first + count - 1, m->n);
  if (int r = batch_gradient(m, first, count, 0ull)) return r;
  if (grad) HIPCHK(hipMemcpyAsync(grad, m->g, CL_NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  if (loss) HIPCHK(hipMemcpyAsync(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int aslhip_mlp_train(aslhip_mlp *m, int steps, int batch, double lr, unsigned long long seed, double *losses) {
  if (!m || !m->n) return fail(ASLHIP_ESTATE, "ASL ERROR: mlp_train before mlp_set_data");
  if (steps < 0 || batch < 1 || !(lr > 0.)) return fail(ASLHIP_EINVAL, "ASL ERROR: mlp_train: bad arguments");
  std::vector<double> lv(losses ? steps : 0);
  const double b1 = 0.9, b2 = 0.999;
  for (int s = 0; s < steps; ++s) {
    unsigned long long sd = mix64(seed + 0x9E3779B97F4A7C15ull * (static_cast<unsigned long long>(m->adam_t) + 1ull));
    if (sd == 0) sd = 1;
    if (int r = batch_gradient(m, 0, batch, sd)) return r;
    ++m->adam_t;
    AdamArgs a;
    a.lr = static_cast<float>(lr);
    a.beta1 = static_cast<float>(b1);
    a.beta2 = static_cast<float>(b2);
    a.eps = 1e-8f;
    a.bc1 = static_cast<float>(1. - std::pow(b1, static_cast<double>(m->adam_t)));
    a.bc2 = static_cast<float>(1. - std::pow(b2, static_cast<double>(m->adam_t)));
    HIPCHK(launch_adam(m->p, m->g, m->m, m->v, a, m->stream));
    if (losses) HIPCHK(hipMemcpyAsync(&lv[s], m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  }
  HIPCHK(hipStreamSynchronize(m->stream));
  for (int s = 0; s < (losses ? steps : 0); ++s) losses[s] = lv[s];
  return 0;
}

int aslhip_mlp_get_params(aslhip_mlp *m, float *out) {
  if (!m || !out) return fail(ASLHIP_EINVAL, "ASL ERROR: mlp_get_params needs a handle and a buffer");
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(out, m->p, CL_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
