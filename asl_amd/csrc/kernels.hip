// kernels.hip — hand-written gfx950 kernels of the LBGK hot path.
//
// One fused collide+stream pass per time step in the AA pattern: the
// populations live in one set of Q structure-of-arrays fields (ASL's
// ghost-inclusive, z-fastest layout per direction).  An even step reads and
// writes each cell's own slots (outgoing population i goes to slot inv(i));
// an odd step reads slot (x - c_i, inv i) and writes slot (x + c_i, i).  Both
// move Q reads + Q writes per cell and need no second copy of the lattice.
//
// Boundary conditions fold into the same pass.  ASL fills ghost links after
// the collision (src/num/aslLBGKBC.cxx); here every computation node resolves
// its incoming links when it reads them, through a packed one-byte cell-type
// map and, for the few cells next to a boundary, a 32-byte link pattern:
//   MODE_BOUNCE  f_k(x) = f*_kbar(x) [+ addend]: the cell's own outgoing kbar
//                population of the previous step, which the AA pattern left in
//                the other slot of the same link pair (the read simply uses
//                the other parity's address);
//   MODE_CONST   a constant population (pressure BCs);
//   MODE_LINK    written between steps by the boundary-link kernel
//                (velocity map BCs, whose value needs rho at the ghost node).
// Wavefronts test the cell types with a ballot: a wavefront without
// boundary cells runs the plain streaming path.
//
// The collision is ASL's LBGK::init0 (src/num/aslLBGK.cxx:187-235) in ASL's
// operation order; the library is compiled with -ffp-contract=off so each
// operation rounds once, as in the CPU oracle.
#include <hip/hip_runtime.h>

#include "../../include/aslhip.h"
#include "kernels.hpp"

namespace aslhip {
namespace {

constexpr int BLOCK = 256;

// ---------------------------------------------------------------------------
// Storage codecs: FP32, FP64, and BF16 / FP16 deviations f - 1 (ASL's
// weight-normalised populations are 1 at rest), all computed in FP32/FP64.
// ---------------------------------------------------------------------------
template <class S>
struct Codec;

template <>
struct Codec<float> {
  using C = float;
  using M = float;
  static __device__ __forceinline__ C ld(const float *p, int i) { return p[i]; }
  static __device__ __forceinline__ void st(float *p, int i, C v) { p[i] = v; }
};

template <>
struct Codec<double> {
  using C = double;
  using M = double;
  static __device__ __forceinline__ C ld(const double *p, int i) { return p[i]; }
  static __device__ __forceinline__ void st(double *p, int i, C v) { p[i] = v; }
};

// BF16 bit pattern of f - 1.
template <>
struct Codec<unsigned short> {
  using C = float;
  using M = float;
  static __device__ __forceinline__ C ld(const unsigned short *p, int i) {
    return 1.0f + __uint_as_float(static_cast<unsigned>(p[i]) << 16);
  }
  static __device__ __forceinline__ void st(unsigned short *p, int i, C v) {
    unsigned u = __float_as_uint(v - 1.0f);
    if ((u & 0x7fffffffu) > 0x7f800000u)
      u |= 0x00400000u;  // keep NaN quiet
    else
      u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    p[i] = static_cast<unsigned short>(u >> 16);
  }
};

template <>
struct Codec<_Float16> {
  using C = float;
  using M = float;
  static __device__ __forceinline__ C ld(const _Float16 *p, int i) { return 1.0f + static_cast<float>(p[i]); }
  static __device__ __forceinline__ void st(_Float16 *p, int i, C v) { p[i] = static_cast<_Float16>(v - 1.0f); }
};

template <int Q>
__device__ __forceinline__ constexpr int offc(int k, int sx, int sy) {
  return Lat<Q>::c[k][0] * sx + Lat<Q>::c[k][1] * sy + Lat<Q>::c[k][2];
}

// v . c_i with the zero terms left out (x + (+-0) == x for finite x).
template <int Q, class C>
__device__ __forceinline__ C dot_c(const C *u, int i) {
  C a = C(0);
  bool first = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int cd = Lat<Q>::c[i][d];
    if (cd == 0) continue;
    const C t = cd > 0 ? u[d] : -u[d];
    a = first ? t : a + t;
    first = false;
  }
  return a;
}

// The BGK collision of one cell: ASL's LBGK::init0 (src/num/aslLBGK.cxx:
// 187-235) with computeRho / computeMomentum / a_i_dot_v (:46-72).
template <int Q, class C>
__device__ __forceinline__ void collide(C (&v)[Q], C w, bool compressible, C &r_o, C (&vel)[3]) {
  using LT = Lat<Q>;
  C m[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    C acc = C(0);
    bool first = true;
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      if (LT::c[j][d] == 0) continue;
      const C t = v[j] * C(double(LT::c[j][d]) * LT::w(j));
      acc = first ? t : acc + t;
      first = false;
    }
    m[d] = acc;
  }
  C r = v[0] * C(LT::w(0));
#pragma unroll
  for (int j = 1; j < Q; ++j) r = r + v[j] * C(LT::w(j));
  r_o = r;
  C mm = (C(1.5) * m[0]) * m[0];
  mm = mm + (C(1.5) * m[1]) * m[1];
  mm = mm + (C(1.5) * m[2]) * m[2];
  C mmr;
  if (compressible) {
#pragma unroll
    for (int d = 0; d < 3; ++d) vel[d] = m[d] / r;
    mmr = mm / r;
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) vel[d] = m[d];
    mmr = mm;
  }
  const C one_w = C(1.) - w;
  v[0] = v[0] * one_w + w * (r - mmr);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    const C a = dot_c<Q>(m, i);
    C feq;
    if (compressible)
      feq = ((r + C(3.) * a) + (C(4.5) * (a * a)) / r) - mmr;
    else
      feq = ((r + C(3.) * a) + C(4.5) * (a * a)) - mmr;
    v[i] = v[i] * one_w + w * feq;
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs (id % 8).  Give each
// XCD one contiguous range of chunks so that neighbouring chunks, which share
// cache lines at their edges, meet in the same L2.
__device__ __forceinline__ int xcd_remap(int L, int n) {
  const int q = n >> 3, r = n & 7, j = L & 7, t = L >> 3;
  return (j < r) ? j * (q + 1) + t : r * (q + 1) + (j - r) * q + t;
}

__device__ __forceinline__ int wrap1(int v, int n) { return v < 1 ? n : (v > n ? 1 : v); }

template <int Q, class S, bool ODD, bool PER>
__global__ __launch_bounds__(BLOCK) void k_step(const StepArgs a) {
  using CD = Codec<S>;
  using C = typename CD::C;
  using M = typename CD::M;
  const int L = xcd_remap(blockIdx.x, a.nblocks);
  const int plane = L / a.bpp;
  const int chunk = L - plane * a.bpp;
  const int off = chunk * BLOCK + static_cast<int>(threadIdx.x);
  if (off >= a.ny * a.sy) return;
  const int idx = (plane + 1) * a.sx + a.sy + off;
  const unsigned char code = a.code[idx];
  if (code == CODE_SKIP) return;

  auto fk = [&](int k) -> S * { return reinterpret_cast<S *>(a.f + k * a.qs_bytes); };

  // Neighbour indices (periodic lattices wrap over the interior).
  int nbm[Q], nbp[Q];
  if constexpr (PER) {
    const int x = plane + 1, y = off / a.sy + 1, z = off - (y - 1) * a.sy;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      nbm[k] = wrap1(x - Lat<Q>::c[k][0], a.nx) * a.sx + wrap1(y - Lat<Q>::c[k][1], a.ny) * a.sy +
               wrap1(z - Lat<Q>::c[k][2], a.nz);
      nbp[k] = wrap1(x + Lat<Q>::c[k][0], a.nx) * a.sx + wrap1(y + Lat<Q>::c[k][1], a.ny) * a.sy +
               wrap1(z + Lat<Q>::c[k][2], a.nz);
    }
  } else {
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      nbm[k] = idx - offc<Q>(k, a.sx, a.sy);
      nbp[k] = idx + offc<Q>(k, a.sx, a.sy);
    }
  }

  C v[Q];
  unsigned skipm = 0;
  const bool bnd = code != CODE_FLUID;
  if (!PER && __any(bnd)) {
    // General path: some lane of this wavefront has BC-resolved links.
    LinkRec rec;
    if (bnd) {
      rec = a.rec[a.pat[idx]];
      skipm = rec.skip;
    } else {
#pragma unroll
      for (int k = 0; k < 28; ++k) rec.slot[k] = 0;
    }
    const bool raw = (a.flags & STEP_F_RAW) != 0;
    const C *val = static_cast<const C *>(a.val);
    v[0] = CD::ld(fk(0), idx);
#pragma unroll
    for (int k = 1; k < Q; ++k) {
      const int kb = inv_dir<Q>(k);
      const int slot = raw ? 0 : rec.slot[k];
      const unsigned char md = slot ? a.mode[slot] : static_cast<unsigned char>(MODE_NONE);
      C x;
      if (md == MODE_BOUNCE)
        x = CD::ld(fk(ODD ? k : kb), ODD ? idx : nbm[k]) + val[slot * Q + k];
      else if (md == MODE_CONST)
        x = val[slot * Q + k];
      else
        x = CD::ld(fk(ODD ? kb : k), ODD ? nbm[k] : idx);
      v[k] = x;
    }
  } else {
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const int kb = inv_dir<Q>(k);
      v[k] = CD::ld(fk(ODD ? kb : k), ODD ? nbm[k] : idx);
    }
  }

  C r, vel[3];
  collide<Q, C>(v, C(a.w), (a.flags & STEP_F_COMPRESSIBLE) != 0, r, vel);

#pragma unroll
  for (int k = 0; k < Q; ++k) {
    if ((skipm >> k) & 1u) continue;
    const int kb = inv_dir<Q>(k);
    CD::st(fk(ODD ? k : kb), ODD ? nbp[k] : idx, v[k]);
  }
  if (a.flags & STEP_F_RHO) static_cast<M *>(a.rho)[idx] = M(r);
  if (a.flags & STEP_F_VEL) {
    M *u = static_cast<M *>(a.u);
    u[idx] = M(vel[0]);
    u[a.len + idx] = M(vel[1]);
    u[2 * a.len + idx] = M(vel[2]);
  }
}

// Boundary-link kernel: ASL's BCVelocityMap (src/num/aslLBGKBC.cxx:290-348).
// After a step, rho at ghost node g is estimated from the populations its
// computation-node neighbours sent towards it; the links g -> g+c_i are then
// written where the receivers read them in the next step.
template <int Q, class S>
__global__ __launch_bounds__(BLOCK) void k_vlinks(const VLinkArgs a) {
  using CD = Codec<S>;
  using C = typename CD::C;
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= a.nlinks) return;
  const VLink l = a.links[e];
  const VLinkBC bc = a.bc[l.slot];
  auto fk = [&](int k) -> S * { return reinterpret_cast<S *>(a.f + k * a.qs_bytes); };
  C coef = C(0), r = C(0);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    if (!((l.comp >> i) & 1u)) continue;
    const int kb = inv_dir<Q>(i);
    const int p = l.g + offc<Q>(i, a.sx, a.sy);
    const C fi = a.after_odd ? CD::ld(fk(kb), l.g) : CD::ld(fk(i), p);
    coef = coef + C(Lat<Q>::w(i));
    r = r + fi * C(Lat<Q>::w(i));
  }
  r = r / coef;
  // Index2PositionACL over the ghost-inclusive block (src/math/aslIndex2Position.cxx:64-108)
  const int I0 = l.g / a.sx, rem = l.g - I0 * a.sx, I1 = rem / a.sy, I2 = rem - I1 * a.sy;
  const int I[3] = {I0, I1, I2};
  C pos[3], uw[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) pos[d] = C(a.pos0[d] - a.dx) + C(a.dx) * C(I[d]);
  if (bc.has_rot) {
    // PFRotationField: crossProduct(axis, pos - c)
    const C d0 = pos[0] - C(bc.centre[0]), d1 = pos[1] - C(bc.centre[1]), d2 = pos[2] - C(bc.centre[2]);
    uw[0] = C(bc.omega[1]) * d2 - C(bc.omega[2]) * d1;
    uw[1] = C(bc.omega[2]) * d0 - C(bc.omega[0]) * d2;
    uw[2] = C(bc.omega[0]) * d1 - C(bc.omega[1]) * d0;
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) uw[d] = C(bc.u[d]);
  }
  const C mom[3] = {uw[0] * r, uw[1] * r, uw[2] * r};
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    if (!((l.own >> i) & 1u)) continue;
    const int kb = inv_dir<Q>(i);
    const int p = l.g + offc<Q>(i, a.sx, a.sy);
    const C val = r + dot_c<Q>(mom, i) * C(3.);
    if (a.after_odd)
      CD::st(fk(i), p, val);  // next step is even: p reads slot (p, i)
    else
      CD::st(fk(kb), l.g, val);  // next step is odd: p reads slot (g, inv i)
  }
}

// LBGKUtilities::initF(rho, vel) (src/num/aslLBGK.cxx:262-275) into the state
// "after an even step": the outgoing population i of cell x in slot (x, inv i).
template <int Q, class S>
__global__ __launch_bounds__(BLOCK) void k_init(const InitArgs a) {
  using CD = Codec<S>;
  using C = typename CD::C;
  using M = typename CD::M;
  const int idx = blockIdx.x * BLOCK + threadIdx.x;
  if (idx >= a.len) return;
  C r, u[3];
  r = a.rho_in ? C(static_cast<const M *>(a.rho_in)[idx]) : C(a.rho0);
#pragma unroll
  for (int d = 0; d < 3; ++d) u[d] = a.u_in ? C(static_cast<const M *>(a.u_in)[d * a.len + idx]) : C(a.u0[d]);
  C uu = (C(1.5) * u[0]) * u[0];
  uu = uu + (C(1.5) * u[1]) * u[1];
  uu = uu + (C(1.5) * u[2]) * u[2];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const C ai = i == 0 ? C(0) : dot_c<Q>(u, i);
    const C fi = ((C(1.) + C(3.) * ai) + C(4.5) * (ai * ai)) - uu;
    CD::st(reinterpret_cast<S *>(a.f + inv_dir<Q>(i) * a.qs_bytes), idx, r * fi);
  }
  static_cast<M *>(a.rho)[idx] = M(r);
#pragma unroll
  for (int d = 0; d < 3; ++d) static_cast<M *>(a.u)[d * a.len + idx] = M(u[d]);
}

// The post-collision populations in ASL's view (f_i of cell x), from the AA
// state.  Slots outside the array (outer ghost layer) read as NaN.
template <int Q, class S>
__global__ __launch_bounds__(BLOCK) void k_unpermute(const UnpermuteArgs a) {
  using CD = Codec<S>;
  using M = typename CD::M;
  const int idx = blockIdx.x * BLOCK + threadIdx.x;
  if (idx >= a.len) return;
  M *out = static_cast<M *>(a.out);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    M v;
    if (a.after_odd) {
      const int src = idx + offc<Q>(i, a.sx, a.sy);
      v = (src >= 0 && src < a.len)
              ? M(CD::ld(reinterpret_cast<const S *>(a.f + i * a.qs_bytes), src))
              : M(__builtin_nan(""));
    } else {
      v = M(CD::ld(reinterpret_cast<const S *>(a.f + inv_dir<Q>(i) * a.qs_bytes), idx));
    }
    out[static_cast<long long>(i) * a.len + idx] = v;
  }
}

constexpr int RED_MAX_BLOCKS = 1024;

// FP64 mass / momentum / kinetic energy over computation nodes.
template <int Q, class S>
__global__ __launch_bounds__(BLOCK) void k_reduce(const ReduceArgs a) {
  using CD = Codec<S>;
  __shared__ double sh[5][BLOCK];
  double s[5] = {0., 0., 0., 0., 0.};
  for (int idx = blockIdx.x * BLOCK + threadIdx.x; idx < a.len; idx += gridDim.x * BLOCK) {
    const unsigned char c = a.code[idx];
    if (c != CODE_FLUID && c != CODE_FLUID_BC) continue;
    double r = 0., m[3] = {0., 0., 0.};
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      double fi;
      if (a.after_odd)
        fi = double(CD::ld(reinterpret_cast<const S *>(a.f + i * a.qs_bytes), idx + offc<Q>(i, a.sx, a.sy)));
      else
        fi = double(CD::ld(reinterpret_cast<const S *>(a.f + inv_dir<Q>(i) * a.qs_bytes), idx));
      fi *= Lat<Q>::w(i);
      r += fi;
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (Lat<Q>::c[i][d]) m[d] += Lat<Q>::c[i][d] > 0 ? fi : -fi;
    }
    s[0] += r;
    s[1] += m[0];
    s[2] += m[1];
    s[3] += m[2];
    s[4] += (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) / (2. * r);
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w)
#pragma unroll
      for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 5) a.partial[blockIdx.x * 5 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(BLOCK) void k_reduce_final(const double *partial, int nblocks, double *out) {
  __shared__ double sh[5][BLOCK];
  double s[5] = {0., 0., 0., 0., 0.};
  for (int b = threadIdx.x; b < nblocks; b += BLOCK)
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] += partial[b * 5 + k];
#pragma unroll
  for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w)
#pragma unroll
      for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 5) out[threadIdx.x] = sh[threadIdx.x][0];
}

template <int Q, class S>
hipError_t step_q_s(bool odd, bool per, const StepArgs &a, hipStream_t s) {
  const dim3 grid(a.nblocks), block(BLOCK);
  if (per) {
    if (odd)
      hipLaunchKernelGGL((k_step<Q, S, true, true>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((k_step<Q, S, false, true>), grid, block, 0, s, a);
  } else {
    if (odd)
      hipLaunchKernelGGL((k_step<Q, S, true, false>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((k_step<Q, S, false, false>), grid, block, 0, s, a);
  }
  return hipGetLastError();
}

template <template <int, class> class F, class... Args>
hipError_t dispatch(int q, int dtype, Args &&...args) {
#define ASLHIP_DT(QQ)                                                   \
  switch (dtype) {                                                      \
    case ASLHIP_F32: return F<QQ, float>::run(args...);                 \
    case ASLHIP_F64: return F<QQ, double>::run(args...);                \
    case ASLHIP_BF16: return F<QQ, unsigned short>::run(args...);       \
    case ASLHIP_F16: return F<QQ, _Float16>::run(args...);              \
    default: return hipErrorInvalidValue;                               \
  }
  switch (q) {
    case 15: ASLHIP_DT(15)
    case 19: ASLHIP_DT(19)
    case 27: ASLHIP_DT(27)
    default: return hipErrorInvalidValue;
  }
#undef ASLHIP_DT
}

template <int Q, class S>
struct StepF {
  static hipError_t run(bool odd, bool per, const StepArgs &a, hipStream_t s) { return step_q_s<Q, S>(odd, per, a, s); }
};
template <int Q, class S>
struct VLinkF {
  static hipError_t run(const VLinkArgs &a, hipStream_t s) {
    if (a.nlinks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_vlinks<Q, S>), dim3((a.nlinks + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
  }
};
template <int Q, class S>
struct InitF {
  static hipError_t run(const InitArgs &a, hipStream_t s) {
    hipLaunchKernelGGL((k_init<Q, S>), dim3((a.len + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
  }
};
template <int Q, class S>
struct UnpermuteF {
  static hipError_t run(const UnpermuteArgs &a, hipStream_t s) {
    hipLaunchKernelGGL((k_unpermute<Q, S>), dim3((a.len + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
  }
};
template <int Q, class S>
struct ReduceF {
  static hipError_t run(const ReduceArgs &a, double *out, hipStream_t s) {
    const int nb = reduce_blocks(a.len);
    hipLaunchKernelGGL((k_reduce<Q, S>), dim3(nb), dim3(BLOCK), 0, s, a);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(BLOCK), 0, s, a.partial, nb, out);
    return hipGetLastError();
  }
};

}  // namespace

int reduce_blocks(int len) {
  const int nb = (len + BLOCK - 1) / BLOCK;
  return nb < RED_MAX_BLOCKS ? (nb < 1 ? 1 : nb) : RED_MAX_BLOCKS;
}

hipError_t launch_step(int q, int dtype, bool odd, bool periodic, const StepArgs &a, hipStream_t s) {
  return dispatch<StepF>(q, dtype, odd, periodic, a, s);
}
hipError_t launch_vlinks(int q, int dtype, const VLinkArgs &a, hipStream_t s) {
  return dispatch<VLinkF>(q, dtype, a, s);
}
hipError_t launch_init(int q, int dtype, const InitArgs &a, hipStream_t s) { return dispatch<InitF>(q, dtype, a, s); }
hipError_t launch_unpermute(int q, int dtype, const UnpermuteArgs &a, hipStream_t s) {
  return dispatch<UnpermuteF>(q, dtype, a, s);
}
hipError_t launch_reduce(int q, int dtype, const ReduceArgs &a, double *out, hipStream_t s) {
  return dispatch<ReduceF>(q, dtype, a, out, s);
}

}  // namespace aslhip
