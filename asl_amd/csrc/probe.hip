// probe.hip — memory-system probes behind the step kernel's design choices
// (libaslprobe.so; used by bench.py --pattern-probe, not by the library).
//
// copy: a plain device copy, one or four floats per lane: the bandwidth a
//   read+write stream reaches on the device.
// aa:   the D3Q19 AA access pattern of the step kernel on its own (19 reads
//   and 19 writes of FP32 populations per interior cell, even and odd steps,
//   a stand-in relaxation so every store depends on every load, no boundary
//   logic), on two population layouts:
//     LAYOUT_ASL  ASL's ghost-inclusive block, rows of nz + 2 cells;
//     LAYOUT_PAD  rows padded to a multiple of 32 cells with z = 1 of every
//                 row on a 128-byte line boundary;
//   with 1, 2 or 4 consecutive cells along z per lane.  Lanes cover the
//   interior only.  Workgroups take chunks of a plane, dealt to the XCDs in
//   runs and visited in reverse on odd steps as the step kernel does.
#include <hip/hip_runtime.h>

#include <vector>

#include "lattice_consts.hpp"

namespace {

constexpr int Q = 19;
using aslhip::inv_dir;
using aslhip::Lat;

enum { LAYOUT_ASL = 0, LAYOUT_PAD = 1 };

struct AAArgs {
  float *f;             // Q direction arrays, qs elements apart
  long long qs;         // elements per direction array
  int sx, sy, base;     // index of (x, y, z) = base + x sx + y sy + z
  int nx, ny, nz;
  int cpl;              // cells per lane
  int per_plane;        // lanes per plane = ny nz / cpl
  int bpp, nblocks, remap;
};

template <int N>
struct alignas(4) VecT {
  float v[N];
};

__device__ __forceinline__ int xcd_remap(int L, int n, int g) {
  const int span = 8 * g, grp = L / span;
  if ((grp + 1) * span > n) return L;
  const int r = L - grp * span;
  return grp * span + (r & 7) * g + (r >> 3);
}

__device__ __forceinline__ int xcd_reverse(int b, int n) {
  const int r = b & 7, nr = (n - r + 7) >> 3;
  return r + 8 * (nr - 1 - (b >> 3));
}

template <int N>
__device__ __forceinline__ VecT<N> ld(const float *p) {
  return *reinterpret_cast<const VecT<N> *>(p);
}
template <int N>
__device__ __forceinline__ void st(float *p, const VecT<N> &v) {
  *reinterpret_cast<VecT<N> *>(p) = v;
}

// Addresses are 32-bit byte offsets from the first direction array, as in
// the step kernel (the host keeps the populations below 4 GiB).
template <int CPL, bool ODD>
__global__ __launch_bounds__(1024) void k_aa(const AAArgs a) {
  int b = static_cast<int>(blockIdx.x);
  if (ODD) b = xcd_reverse(b, a.nblocks);
  if (a.remap > 1) b = xcd_remap(b, a.nblocks, a.remap);
  const int plane = b / a.bpp, chunk = b - plane * a.bpp;
  const int j = chunk * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
  if (j >= a.per_plane) return;
  const int c0 = j * CPL;
  const int row = c0 / a.nz, z = c0 - row * a.nz + 1;
  const unsigned ib = 4u * static_cast<unsigned>(a.base + (plane + 1) * a.sx + (row + 1) * a.sy + z);
  const unsigned qb = 4u * static_cast<unsigned>(a.qs);
  const char *f = reinterpret_cast<const char *>(a.f);
  VecT<CPL> v[Q];
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    const int kb = inv_dir<Q>(k);
    const unsigned o = 4u * static_cast<unsigned>(Lat<Q>::c[k][0] * a.sx + Lat<Q>::c[k][1] * a.sy + Lat<Q>::c[k][2]);
    // even: slot (x, k); odd: slot (x - c_k, inv k)
    const unsigned p = ODD ? static_cast<unsigned>(kb) * qb + ib - o : static_cast<unsigned>(k) * qb + ib;
    v[k] = ld<CPL>(reinterpret_cast<const float *>(f + p));
  }
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < Q; ++k) s += v[k].v[c];
    s *= 1.f / Q;
#pragma unroll
    for (int k = 0; k < Q; ++k) v[k].v[c] = __builtin_fmaf(v[k].v[c], .5f, .5f * s);
  }
  char *fw = reinterpret_cast<char *>(a.f);
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    const int kb = inv_dir<Q>(k);
    const unsigned o = 4u * static_cast<unsigned>(Lat<Q>::c[k][0] * a.sx + Lat<Q>::c[k][1] * a.sy + Lat<Q>::c[k][2]);
    // even: slot (x, inv k); odd: slot (x + c_k, k)
    const unsigned p = ODD ? static_cast<unsigned>(k) * qb + ib + o : static_cast<unsigned>(kb) * qb + ib;
    st<CPL>(reinterpret_cast<float *>(fw + p), v[k]);
  }
}

template <int N>
__global__ __launch_bounds__(256) void k_copy(const float *__restrict__ src, float *__restrict__ dst, long long n) {
  const long long i = (static_cast<long long>(blockIdx.x) * 256 + threadIdx.x) * N;
  if (i + N <= n) st<N>(dst + i, ld<N>(src + i));
}

__global__ void k_fill(float *p, long long n, float v) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

template <class F>
int time_launches(int reps, double *ms, F launch) {
  std::vector<hipEvent_t> ev(2 * static_cast<size_t>(reps));
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) return -3;
  for (int r = 0; r < reps; ++r) {
    (void)hipEventRecord(ev[2 * r], nullptr);
    launch(r);
    (void)hipEventRecord(ev[2 * r + 1], nullptr);
  }
  int rc = hipEventSynchronize(ev.back()) == hipSuccess && hipGetLastError() == hipSuccess ? 0 : -3;
  for (int r = 0; r < reps && rc == 0; ++r) {
    float t = 0.f;
    (void)hipEventElapsedTime(&t, ev[2 * r], ev[2 * r + 1]);
    ms[r] = t;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  return rc;
}

}  // namespace

extern "C" {

// One device copy of nbytes (vec = 1 or 4 floats per lane), reps times;
// ms[reps]: duration of each.  Returns 0, or a negative code.
int aslprobe_copy(long long nbytes, int vec, int reps, double *ms) {
  if ((vec != 1 && vec != 4) || reps < 1 || nbytes < 1024 || !ms) return -1;
  const long long n = nbytes / 16 * 4;  // a multiple of 4 floats
  float *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, n * 4) != hipSuccess) return -4;
  if (hipMalloc(&dst, n * 4) != hipSuccess) {
    (void)hipFree(src);
    return -4;
  }
  hipLaunchKernelGGL(k_fill, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, nullptr, src, n, 1.f);
  const unsigned grid = static_cast<unsigned>((n / vec + 255) / 256);
  int rc = time_launches(reps, ms, [&](int) {
    if (vec == 4)
      hipLaunchKernelGGL(k_copy<4>, dim3(grid), dim3(256), 0, nullptr, src, dst, n);
    else
      hipLaunchKernelGGL(k_copy<1>, dim3(grid), dim3(256), 0, nullptr, src, dst, n);
  });
  (void)hipFree(src);
  (void)hipFree(dst);
  return rc;
}

// The bare D3Q19 AA pattern on an nx x ny x nz interior: reps launches
// alternating even (first) and odd steps; ms[reps]: duration of each.
int aslprobe_aa(int nx, int ny, int nz, int layout, int cpl, int block, int remap, int reps, double *ms) {
  if (nx < 1 || ny < 1 || nz < 1 || (cpl != 1 && cpl != 2 && cpl != 4) || nz % cpl || block < 64 ||
      block > 1024 || block % 64 || reps < 1 || !ms || (layout != LAYOUT_ASL && layout != LAYOUT_PAD))
    return -1;
  AAArgs a;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.cpl = cpl;
  if (layout == LAYOUT_ASL) {
    a.sy = nz + 2;
    a.base = 0;
  } else {
    a.sy = (nz + 2 + 31) / 32 * 32;
    a.base = 31;  // z = 1 at a multiple of 32 elements
  }
  a.sx = (ny + 2) * a.sy;
  const long long len = static_cast<long long>(a.base) + static_cast<long long>(nx + 2) * a.sx;
  a.qs = (len + 63) / 64 * 64 + 64;  // 256-byte aligned arrays
  if (a.qs * Q * 4 >= (1LL << 32)) return -1;  // 32-bit byte offsets
  a.per_plane = ny * nz / cpl;
  a.bpp = (a.per_plane + block - 1) / block;
  a.nblocks = a.bpp * nx;
  a.remap = remap;
  if (hipMalloc(&a.f, a.qs * Q * 4) != hipSuccess) return -4;
  const long long nf = a.qs * Q;
  hipLaunchKernelGGL(k_fill, dim3(static_cast<unsigned>((nf + 255) / 256)), dim3(256), 0, nullptr, a.f, nf, 1.f);
  int rc = time_launches(reps, ms, [&](int r) {
    const bool odd = r & 1;
#define PROBE_L(C)                                                                                \
  if (odd)                                                                                        \
    hipLaunchKernelGGL((k_aa<C, true>), dim3(a.nblocks), dim3(block), 0, nullptr, a);            \
  else                                                                                            \
    hipLaunchKernelGGL((k_aa<C, false>), dim3(a.nblocks), dim3(block), 0, nullptr, a);
    if (cpl == 1) {
      PROBE_L(1)
    } else if (cpl == 2) {
      PROBE_L(2)
    } else {
      PROBE_L(4)
    }
#undef PROBE_L
  });
  (void)hipFree(a.f);
  return rc;
}

}  // extern "C"
