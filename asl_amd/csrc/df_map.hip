// Level-set maps from distance-function expressions on the GPU: ASL's
// initData(map, DF) / generateDataContainer_SP(block, DF, ...)
// (src/aslGeomInc.cxx:187-221), which evaluate the DistanceFunction tree
// (src/math/aslDistanceFunction.cxx:55-164) at every node of the block in the
// map's element type.
//
// The tree arrives as a postfix program of doubles (op code, then the op's
// parameters; see include/aslhip.h); one thread evaluates it at one node with
// a small stack, constants rounded to the element type T where they are used,
// every operation in ASL's order and rounded once (no contraction), so that
// the map equals the host evaluation of the same tree bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/aslhip.h"
#include "status.hpp"

namespace aslhip {
namespace {

constexpr int DF_BLOCK = 256;
constexpr int DF_STACK = 16;

// Number of parameters of each op (index = op code).
constexpr int df_nparam(int op) {
  return op == ASLHIP_DF_SPHERE ? 4
         : (op == ASLHIP_DF_CYLINDER || op == ASLHIP_DF_CONE) ? 7
         : op == ASLHIP_DF_PLANE ? 4
         : op == ASLHIP_DF_NORMALIZE ? 1
         : 0;
}

template <class T>
__global__ __launch_bounds__(DF_BLOCK) void k_df_map(const double *prog, int nwords, int sy, int sz, long long n,
                                                    T p0, T p1, T p2, T dx, T *out) {
  const long long i = static_cast<long long>(blockIdx.x) * DF_BLOCK + threadIdx.x;
  if (i >= n) return;
  const int z = static_cast<int>(i % sz);
  const long long r = i / sz;
  const int y = static_cast<int>(r % sy), x = static_cast<int>(r / sy);
  // Index2PositionACL: position + dx * index (src/math/aslIndex2Position.cxx:64-108)
  const T p[3] = {p0 + dx * T(x), p1 + dx * T(y), p2 + dx * T(z)};
  T st[DF_STACK];
  int top = 0;
  for (int pc = 0; pc < nwords;) {
    const int op = static_cast<int>(prog[pc]);
    const double *a = prog + pc + 1;
    switch (op) {
      case ASLHIP_DF_SPHERE: {  // sqrt(l2(pos - c)) - r
        const T d0 = p[0] - T(a[0]), d1 = p[1] - T(a[1]), d2 = p[2] - T(a[2]);
        st[top++] = sqrt((d0 * d0 + d1 * d1) + d2 * d2) - T(a[3]);
        break;
      }
      case ASLHIP_DF_CYLINDER:    // sqrt(l2(d - (d.o) o)) - r, d = pos - c
      case ASLHIP_DF_CONE: {      // sqrt(l2(d - (d.o) o)) - tan * (d.o), d = pos - apex
        const T d0 = p[0] - T(a[0]), d1 = p[1] - T(a[1]), d2 = p[2] - T(a[2]);
        const T o0 = T(a[3]), o1 = T(a[4]), o2 = T(a[5]);
        const T proj = (d0 * o0 + d1 * o1) + d2 * o2;
        const T e0 = d0 - proj * o0, e1 = d1 - proj * o1, e2 = d2 - proj * o2;
        const T rad = sqrt((e0 * e0 + e1 * e1) + e2 * e2);
        st[top++] = op == ASLHIP_DF_CONE ? rad - T(a[6]) * proj : rad - T(a[6]);
        break;
      }
      case ASLHIP_DF_PLANE:  // pos . n - b
        st[top++] = ((p[0] * T(a[0]) + p[1] * T(a[1])) + p[2] * T(a[2])) - T(a[3]);
        break;
      case ASLHIP_DF_AND: {  // max(x, y)
        const T y2 = st[--top], x2 = st[top - 1];
        st[top - 1] = x2 < y2 ? y2 : x2;
        break;
      }
      case ASLHIP_DF_OR: {  // min(x, y)
        const T y2 = st[--top], x2 = st[top - 1];
        st[top - 1] = y2 < x2 ? y2 : x2;
        break;
      }
      case ASLHIP_DF_NEG:
        st[top - 1] = -st[top - 1];
        break;
      case ASLHIP_DF_NORMALIZE: {  // clamp(f d, -1, 1)
        const T dist = T(a[0]) * st[top - 1];
        const T t = dist > T(1.) ? T(1.) : dist;
        st[top - 1] = dist < T(-1.) ? T(-1.) : t;
        break;
      }
      default:
        break;
    }
    pc += 1 + df_nparam(op);
  }
  out[i] = st[0];
}

// Checks the program: known ops, complete parameters, stack within bounds,
// one value left.
bool df_check(const double *prog, int nwords) {
  int depth = 0;
  for (int pc = 0; pc < nwords;) {
    const double w = prog[pc];
    if (!(w >= 1. && w <= 8.) || w != std::floor(w)) return false;
    const int op = static_cast<int>(w);
    if (pc + 1 + df_nparam(op) > nwords) return false;
    for (int k = 1; k <= df_nparam(op); ++k)
      if (!std::isfinite(prog[pc + k])) return false;
    if (op <= ASLHIP_DF_PLANE) {
      if (++depth > DF_STACK) return false;
    } else if (op == ASLHIP_DF_AND || op == ASLHIP_DF_OR) {
      if (depth < 2) return false;
      --depth;
    } else if (depth < 1) {
      return false;
    }
    pc += 1 + df_nparam(op);
  }
  return depth == 1;
}

struct Buf {
  void *p = nullptr;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

template <class T>
int df_run(const double *prog, int nwords, const int size[3], const double pos[3], double dx, T *out) {
  const long long n = static_cast<long long>(size[0]) * size[1] * size[2];
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreate(&s));
  Buf dprog, dout;
  hipError_t e = hipMalloc(&dprog.p, nwords * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dout.p, n * sizeof(T));
  if (e == hipSuccess) e = hipMemcpyAsync(dprog.p, prog, nwords * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    const unsigned nb = static_cast<unsigned>((n + DF_BLOCK - 1) / DF_BLOCK);
    hipLaunchKernelGGL(k_df_map<T>, dim3(nb), dim3(DF_BLOCK), 0, s, static_cast<const double *>(dprog.p), nwords,
                       size[1], size[2], n, T(pos[0]), T(pos[1]), T(pos[2]), T(dx), static_cast<T *>(dout.p));
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout.p, n * sizeof(T), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipStreamDestroy(s);
  if (e != hipSuccess) return fail(ASLHIP_EHIP, "ASL ERROR: distance map: %s", hipGetErrorString(e));
  return 0;
}

}  // namespace
}  // namespace aslhip

using namespace aslhip;

extern "C" int aslhip_df_map(const double *prog, int nwords, const int size[3], const double pos[3], double dx,
                             int fp64, void *out) {
  if (!prog || !size || !pos || !out || nwords < 1) return fail(ASLHIP_EINVAL, "ASL ERROR: distance map: null argument");
  for (int d = 0; d < 3; ++d)
    if (size[d] < 1 || !std::isfinite(pos[d])) return fail(ASLHIP_EINVAL, "ASL ERROR: distance map: bad grid");
  if (!(dx > 0.) || !std::isfinite(dx)) return fail(ASLHIP_EINVAL, "ASL ERROR: distance map: bad dx");
  if (!df_check(prog, nwords)) return fail(ASLHIP_EINVAL, "ASL ERROR: distance map: malformed program");
  return fp64 ? df_run<double>(prog, nwords, size, pos, dx, static_cast<double *>(out))
              : df_run<float>(prog, nwords, size, pos, dx, static_cast<float *>(out));
}
