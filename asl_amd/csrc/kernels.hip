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

#include <mutex>

#include "../../include/aslhip.h"
#include "kernels.hpp"

namespace aslhip {
namespace {

constexpr int BLOCK = 256;

// The populations live in global memory; their pointers come from the
// argument block in device memory, so the compiler cannot tell the address
// space and would emit flat accesses.  These casts make them global ones.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gload(const char *p) {
  return (const __attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gstore(char *p) {
  return (__attribute__((address_space(1))) T *)p;
}

// ---------------------------------------------------------------------------
// Storage codecs: FP32, FP64, and BF16 / FP16 deviations f - 1 (ASL's
// weight-normalised populations are 1 at rest), all computed in FP32/FP64.
// ---------------------------------------------------------------------------
template <class S>
struct Codec;

// ld/st take a direction array base and a 32-bit byte offset, so loads use
// the scalar-base + 32-bit vector-offset addressing form.
template <>
struct Codec<float> {
  using C = float;
  using M = float;
  static __device__ __forceinline__ C ld(const char *b, unsigned o) { return *gload<float>(b + o); }
  static __device__ __forceinline__ void st(char *b, unsigned o, C v) { *gstore<float>(b + o) = v; }
  static __device__ __forceinline__ C dec(float s) { return s; }
  static __device__ __forceinline__ float enc(C v) { return v; }
};

template <>
struct Codec<double> {
  using C = double;
  using M = double;
  static __device__ __forceinline__ C ld(const char *b, unsigned o) { return *gload<double>(b + o); }
  static __device__ __forceinline__ void st(char *b, unsigned o, C v) { *gstore<double>(b + o) = v; }
  static __device__ __forceinline__ C dec(double s) { return s; }
  static __device__ __forceinline__ double enc(C v) { return v; }
};

// BF16 bit pattern of f - 1.
template <>
struct Codec<unsigned short> {
  using C = float;
  using M = float;
  static __device__ __forceinline__ C ld(const char *b, unsigned o) {
    return 1.0f + __uint_as_float(static_cast<unsigned>(*gload<unsigned short>(b + o)) << 16);
  }
  // round to nearest even in one instruction (v_cvt_pk_bf16_f32 on gfx950)
  static __device__ __forceinline__ void st(char *b, unsigned o, C v) {
    *gstore<unsigned short>(b + o) = __builtin_bit_cast(unsigned short, static_cast<__bf16>(v - 1.0f));
  }
  static __device__ __forceinline__ C dec(unsigned short s) { return 1.0f + __uint_as_float(static_cast<unsigned>(s) << 16); }
  static __device__ __forceinline__ unsigned short enc(C v) {
    return __builtin_bit_cast(unsigned short, static_cast<__bf16>(v - 1.0f));
  }
};

template <>
struct Codec<_Float16> {
  using C = float;
  using M = float;
  static __device__ __forceinline__ C ld(const char *b, unsigned o) {
    return 1.0f + static_cast<float>(*gload<_Float16>(b + o));
  }
  static __device__ __forceinline__ void st(char *b, unsigned o, C v) {
    *gstore<_Float16>(b + o) = static_cast<_Float16>(v - 1.0f);
  }
  static __device__ __forceinline__ C dec(_Float16 s) { return 1.0f + static_cast<float>(s); }
  static __device__ __forceinline__ _Float16 enc(C v) { return static_cast<_Float16>(v - 1.0f); }
};

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

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

// computeMomentum / computeRho (src/num/aslLBGK.cxx:46-63): m = sum w_i c_i f_i
// and rho = sum w_i f_i, left-to-right fma chains without the zero terms.
template <int Q, class C>
__device__ __forceinline__ void moments(const C (&v)[Q], C &r, C (&m)[3]) {
  using LT = Lat<Q>;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    C acc = C(0);
    bool first = true;
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      if (LT::c[j][d] == 0) continue;
      const C cw = C(double(LT::c[j][d]) * LT::w(j));
      acc = first ? v[j] * cw : fmaT(v[j], cw, acc);
      first = false;
    }
    m[d] = acc;
  }
  r = v[0] * C(LT::w(0));
#pragma unroll
  for (int j = 1; j < Q; ++j) r = fmaT(v[j], C(LT::w(j)), r);
}

// LBGKTurbulence::init0 (src/num/aslLBGK.cxx:298-355): the relaxation of a
// cell from how far its populations are from equilibrium,
//   r = 4 |f - feq|^2 / |f|^2,   w = r > 1 ? 1 / (3 min(nu r, 1/2) + 1/2) : 1 / (3 nu + 1/2).
// ASL reads feq there before assigning it (undefined in its generated code);
// here feq is the equilibrium of the cell's own rho and momentum (before the
// rotating-frame term), the value that statement is evidently meant to use.
template <int Q, class C>
__device__ __forceinline__ C turbulence_w(const C (&v)[Q], C r, const C (&m)[3], bool compressible, C nu, C w0) {
  C mm = (C(1.5) * m[0]) * m[0];
  mm = fmaT(C(1.5) * m[1], m[1], mm);
  mm = fmaT(C(1.5) * m[2], m[2], mm);
  const C mmr = compressible ? mm / r : mm;
  C fl2 = v[0] * v[0];
  C d = v[0] - (r - mmr);
  C dl2 = d * d;
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    const C a = dot_c<Q>(m, i);
    const C fe = compressible ? (fmaT(C(3.), a, r) + (C(4.5) * (a * a)) / r) - mmr
                              : fmaT(C(4.5), a * a, fmaT(C(3.), a, r)) - mmr;
    fl2 = fmaT(v[i], v[i], fl2);
    const C di = v[i] - fe;
    dl2 = fmaT(di, di, dl2);
  }
  const C rt = C(4.) * (dl2 / fl2);
  const C nr = nu * rt;
  const C lim = nr < C(.5) ? nr : C(.5);
  return rt > C(1.) ? C(1.) / (C(3.) * lim + C(.5)) : w0;
}

// The BGK collision of one cell: ASL's LBGK::init0 (src/num/aslLBGK.cxx:
// 187-235) with computeRho / computeMomentum / a_i_dot_v (:46-72).  Each
// x*y + z of ASL's expression tree is one fma (OpenCL's default FP_CONTRACT
// ON); the CPU oracle contracts the same sites.  Opposite directions share
// (4.5 a^2)/rho: a_inv(i) = -a_i exactly in IEEE arithmetic, so this is the
// same value ASL computes twice.  On return v holds the post-collision
// populations, r the density and m the momentum.  CL: a per-cell relaxation
// replaces w: 1 LBGKTurbulence's (turb) or the closure network's in FP32 on
// the vector ALUs, 2 the closure network's on the matrix cores (closure:
// D3Q19, FP32 collision).  Each is its own kernel, so that the plain and the
// matrix-core variants do not hold the registers of the FP32 network.
template <int Q, class C, int CL = 0>
__device__ __forceinline__ void collide(C (&v)[Q], C w, bool compressible, bool omega, const C (&om)[3], C &r_o,
                                        C (&m_o)[3], const ClosureParams *cl = nullptr, float tau0 = 0.f,
                                        bool turb = false, C nu = C(0)) {
  C m[3], r;
  moments<Q, C>(v, r, m);
  if constexpr (CL == 1) {  // LBGKTurbulence or the FP32 closure
    if (turb) {
      w = turbulence_w<Q, C>(v, r, m, compressible, nu, w);
    } else if constexpr (Q == 19 && sizeof(C) == 4) {
      float x[7];
      closure_features<Q>(v, r, m, x);
      w = closure_w(x, *cl, tau0);
    }
  } else if constexpr (CL == 2 && Q == 19 && sizeof(C) == 4) {  // the closure on the matrix cores
    float x[7];
    closure_features<Q>(v, r, m, x);
    w = closure_relax(closure_y_mfma(x, *cl), *cl, tau0);
  }
  // LBGK::setOmega (src/num/aslLBGK.cxx:207-208): momentum += (2/w) omega x
  // momentum, one component after the other, each seeing the ones before it
  // (ASL emits one statement per component).
  if (omega) {
    const C t = C(2.) / w;
    m[0] = fmaT(t, fmaT(om[1], m[2], -(om[2] * m[1])), m[0]);
    m[1] = fmaT(t, fmaT(om[2], m[0], -(om[0] * m[2])), m[1]);
    m[2] = fmaT(t, fmaT(om[0], m[1], -(om[1] * m[0])), m[2]);
  }
  C mm = (C(1.5) * m[0]) * m[0];
  mm = fmaT(C(1.5) * m[1], m[1], mm);
  mm = fmaT(C(1.5) * m[2], m[2], mm);
  const C mmr = compressible ? mm / r : mm;
  const C one_w = C(1.) - w;
  v[0] = fmaT(v[0], one_w, w * (r - mmr));
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    const int j = inv_dir<Q>(i);
    if (j < i) continue;
    const C a = dot_c<Q>(m, i);
    C fi, fj;
    if (compressible) {
      const C t = (C(4.5) * (a * a)) / r;
      fi = (fmaT(C(3.), a, r) + t) - mmr;
      fj = (fmaT(C(3.), -a, r) + t) - mmr;
    } else {
      const C aa = a * a;
      fi = fmaT(C(4.5), aa, fmaT(C(3.), a, r)) - mmr;
      fj = fmaT(C(4.5), aa, fmaT(C(3.), -a, r)) - mmr;
    }
    v[i] = fmaT(v[i], one_w, w * fi);
    v[j] = fmaT(v[j], one_w, w * fj);
  }
  r_o = r;
#pragma unroll
  for (int d = 0; d < 3; ++d) m_o[d] = m[d];
}

// Workgroups are dispatched round-robin over the 8 XCDs (id % 8).  With a
// group size g > 1, each XCD takes runs of g consecutive chunks (workgroup b
// of a group of 8 g works on chunk  group * 8g + (b % 8) * g + (b % 8g) / 8),
// so that neighbouring chunks, which share cache lines at their edges, meet
// in the same L2; the last, partial group keeps the hardware order.
__device__ __forceinline__ int xcd_remap(int L, int n, int g) {
  const int span = 8 * g, grp = L / span;
  if ((grp + 1) * span > n) return L;
  const int r = L - grp * span;
  return grp * span + (r & 7) * g + (r >> 3);
}

// Workgroup b of n taking the chunks in reverse order within its XCD's share:
// the chunks b' with b' % 8 == b % 8 are visited from the last to the first,
// so every chunk stays on the XCD (and in the L2) that processed it in the
// previous step.
__device__ __forceinline__ int xcd_reverse(int b, int n) {
  const int r = b & 7, nr = (n - r + 7) >> 3;
  return r + 8 * (nr - 1 - (b >> 3));
}

__device__ __forceinline__ int wrap1(int v, int n) { return v < 1 ? n : (v > n ? 1 : v); }

template <class C>
__device__ __forceinline__ C face_val(const StepArgs &a, int f, int k);
template <>
__device__ __forceinline__ float face_val<float>(const StepArgs &a, int f, int k) {
  return a.fval32[f][k];
}
template <>
__device__ __forceinline__ double face_val<double>(const StepArgs &a, int f, int k) {
  return a.fval64[f][k];
}

// One cell of one step: chunk L of the lattice described by a.
// PEEK: instead of colliding and storing, write the populations the cell
// receives (after every BC link is resolved) to peek[k * len + cell].
template <int Q, class S, bool ODD, bool PER, bool WIDE, int CL = 0, bool PEEK = false>
__device__ __forceinline__ void step_body(const StepArgs &a, const int L, void *peek = nullptr) {
  using CD = Codec<S>;
  using C = typename CD::C;
  using M = typename CD::M;
  constexpr unsigned ES = sizeof(S);
  const int plane = L / a.bpp;
  const int chunk = L - plane * a.bpp;
  unsigned off = static_cast<unsigned>(chunk * a.block) + threadIdx.x;
  // The closure variant keeps every lane to the end (its network may run on
  // matrix instructions across the wavefront): a lane without a cell of its
  // own works on a stand-in interior cell of its plane and stores nothing.
  bool act = off < static_cast<unsigned>(a.ny * a.sy);
  if constexpr (!CL) {
    if (!act) return;
  } else {
    if (!act) off = 1u;
  }
  // Row and column inside the plane (exact: off * sy < 2^40, see StepArgs::zmagic).
  const unsigned row = static_cast<unsigned>((static_cast<unsigned long long>(off) * a.zmagic) >> 40);
  unsigned z = off - row * static_cast<unsigned>(a.sy);
  if (z == 0 || z == static_cast<unsigned>(a.sy - 1)) {  // ghost columns
    if constexpr (!CL) return;
    act = false;
    off = row * static_cast<unsigned>(a.sy) + 1u;
    z = 1u;
  }
  const unsigned idx = static_cast<unsigned>((plane + 1) * a.sx + a.sy) + off;
  const int x = plane + 1, y = static_cast<int>(row) + 1, zi = static_cast<int>(z);
  const unsigned ib = idx * ES;

  // Byte offsets of x - c_k (rd) and x + c_k (wr).
  auto nb = [&](int k, int sgn) -> unsigned {
    if constexpr (PER) {
      return ES * static_cast<unsigned>(wrap1(x + sgn * Lat<Q>::c[k][0], a.nx) * a.sx +
                                        wrap1(y + sgn * Lat<Q>::c[k][1], a.ny) * a.sy +
                                        wrap1(zi + sgn * Lat<Q>::c[k][2], a.nz));
    } else {
      const unsigned o = ES * static_cast<unsigned>(offc<Q>(k, a.sx, a.sy));
      return sgn > 0 ? ib + o : ib - o;
    }
  };

  // Face rule (see StepArgs): bit f of fb is set if the cell lies on face f.
  const bool raw = (a.flags & STEP_F_RAW) != 0;
  unsigned fb = 0;
  if (!PER && !raw)
    fb = (x == 1 ? 1u : 0u) | (x == a.nx ? 2u : 0u) | (y == 1 ? 4u : 0u) | (y == a.ny ? 8u : 0u) |
         (zi == 1 ? 16u : 0u) | (zi == a.nz ? 32u : 0u);
  // Winning face BC of link k: slot * 4 + mode, and its value.
  auto winner = [&](int k, C &val) -> int {
    int wn = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int cd = Lat<Q>::c[k][d];
      if (cd == 0) continue;
      const int fc = 2 * d + (cd > 0 ? 0 : 1);  // the face this link enters through
      const int cand = ((fb >> fc) & 1u) ? a.fface[fc] : 0;
      if (cand > wn) {
        wn = cand;
        val = face_val<C>(a, fc, k);
      }
    }
    return wn;
  };
  // The face rule only touches wavefronts with a face cell: a wavefront-
  // uniform branch, so interior wavefronts skip it and load plainly.
  const bool faces = !PER && __any(fb != 0);
  unsigned bmask = 0, cmask = 0;  // bounce-back / constant links
  if (faces) {
#pragma unroll
    for (int k = 1; k < Q; ++k) {
      C unused;
      const int md = winner(k, unused) & 3;
      if (md == MODE_BOUNCE) bmask |= 1u << k;
      if (md == MODE_CONST) cmask |= 1u << k;
    }
  }

  // One round of loads: the cell type and every population (bounce-back
  // links read the other slot of their link pair, constant links nothing).
  // All addresses lie inside the lattice for interior cells.
  // Loads take no branch (divergent branches around loads into the same
  // register force the compiler to wait for memory inside the branch).
  const unsigned char code = *gload<unsigned char>(reinterpret_cast<const char *>(a.code) + idx);
  C v[Q];
  auto load_all = [&](unsigned bm) {
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const int kb = inv_dir<Q>(k);
      const bool bb = (bm >> k) & 1u;
      if constexpr (WIDE) {
        const char *pn = a.fd[ODD ? kb : k] + (ODD ? nb(k, -1) : ib);  // normal slot
        const char *pp = a.fd[ODD ? k : kb] + (ODD ? ib : nb(k, -1));  // the other slot of the pair
        v[k] = CD::ld(bb ? pp : pn, 0u);
      } else {
        // byte offsets from the first direction array
        const unsigned on = static_cast<unsigned>(ODD ? kb : k) * a.qs32 + (ODD ? nb(k, -1) : ib);
        if (bm == 0u) {
          v[k] = CD::ld(a.fd[0], on);
        } else {
          const unsigned op = static_cast<unsigned>(ODD ? k : kb) * a.qs32 + (ODD ? ib : nb(k, -1));
          v[k] = CD::ld(a.fd[0], bb ? op : on);
        }
      }
    }
  };
  if (faces)
    load_all(bmask);
  else
    load_all(0u);
  // Keep the scheduler from hoisting the wait on the cell type above the
  // population loads.
  __builtin_amdgcn_sched_barrier(0);
  if (faces) {
#pragma unroll
    for (int k = 1; k < Q; ++k) {
      if (!(((bmask | cmask) >> k) & 1u)) continue;
      C val = C(0);
      (void)winner(k, val);
      v[k] = ((bmask >> k) & 1u) ? v[k] + val : val;
    }
  }

  // Cells that are not computed run the collision on whatever they read and
  // store nothing; cells with a link pattern re-read their links from it.
  const bool live = code != CODE_SKIP;
  const bool bnd = live && code != CODE_FLUID;
  const bool general = !PER && __any(bnd);
  unsigned skipm = 0;
  if (general && bnd) {
    const LinkRec rec = a.rec[*gload<unsigned short>(reinterpret_cast<const char *>(a.pat) + 2u * idx)];
    skipm = rec.skip;
    const C *val = static_cast<const C *>(a.val);
#pragma unroll
    for (int k = 1; k < Q; ++k) {
      const int kb = inv_dir<Q>(k);
      const int slot = raw ? 0 : rec.slot[k];
      const unsigned char md = slot ? a.mode[slot] : static_cast<unsigned char>(MODE_NONE);
      C x;
      if constexpr (WIDE) {
        const char *pn = a.fd[ODD ? kb : k] + (ODD ? nb(k, -1) : ib);
        const char *pp = a.fd[ODD ? k : kb] + (ODD ? ib : nb(k, -1));
        x = CD::ld(md == MODE_BOUNCE ? pp : pn, 0u);
      } else {
        const unsigned on = static_cast<unsigned>(ODD ? kb : k) * a.qs32 + (ODD ? nb(k, -1) : ib);
        const unsigned op = static_cast<unsigned>(ODD ? k : kb) * a.qs32 + (ODD ? ib : nb(k, -1));
        x = CD::ld(a.fd[0], md == MODE_BOUNCE ? op : on);
      }
      const C sv = val[slot * Q + k];
      v[k] = md == MODE_BOUNCE ? x + sv : (md == MODE_CONST ? sv : x);
    }
  }

  if constexpr (PEEK) {
    if (!live) return;
    M *o = static_cast<M *>(peek);
#pragma unroll
    for (int k = 0; k < Q; ++k) o[static_cast<long long>(k) * a.len + idx] = M(v[k]);
    return;
  }

  C r, m[3];
  const bool omega = (a.flags & STEP_F_OMEGA) != 0;
  C om[3] = {C(0), C(0), C(0)};
  if (omega) {
#pragma unroll
    for (int d = 0; d < 3; ++d) om[d] = sizeof(C) == 4 ? C(a.om32[d]) : C(a.om64[d]);
  }
  collide<Q, C, CL>(v, C(a.w), (a.flags & STEP_F_COMPRESSIBLE) != 0, omega, om, r, m, a.cl, a.cl_tau0,
                    (a.flags & STEP_F_TURB) != 0, sizeof(C) == 4 ? C(a.nu32) : C(a.nu64));

  if (!live || !act) return;
  // Stores address every direction from fd[0] too (one scalar base pair
  // instead of Q live pointers) unless the lattice is too large for it.
  auto st = [&](int k) {
    const int dir = ODD ? k : inv_dir<Q>(k);
    const unsigned o = ODD ? nb(k, 1) : ib;
    if constexpr (WIDE)
      CD::st(a.fd[dir], o, v[k]);
    else
      CD::st(a.fd[0], static_cast<unsigned>(dir) * a.qs32 + o, v[k]);
  };
  if (general) {
#pragma unroll
    for (int k = 0; k < Q; ++k)
      if (!((skipm >> k) & 1u)) st(k);
  } else {
#pragma unroll
    for (int k = 0; k < Q; ++k) st(k);
  }
  if (a.flags & STEP_F_RHO) *gstore<M>(static_cast<char *>(a.rho) + sizeof(M) * idx) = M(r);
  if (a.flags & STEP_F_VEL) {
    const bool comp = (a.flags & STEP_F_COMPRESSIBLE) != 0;
    char *u = static_cast<char *>(a.u);
#pragma unroll
    for (int d = 0; d < 3; ++d)
      *gstore<M>(u + sizeof(M) * (static_cast<size_t>(d) * a.len + idx)) = M(comp ? m[d] / r : m[d]);
  }
}

// CPL consecutive cells along z in one lane (a vector of raw storage
// elements; only element alignment is assumed, gfx950 loads and stores such
// vectors as one access).
template <class S, int CPL>
struct RunVec {
  typedef S type __attribute__((ext_vector_type(CPL), aligned(sizeof(S))));
};

// One step of CPL consecutive cells z0 .. z0 + CPL - 1 of a row (lanes cover
// the interior only; nz is a multiple of CPL), non-periodic lattices under
// 4 GiB, plain or rotating-frame LBGK.  Same addresses, same collision and
// same link resolution per cell as step_body, so every cell ends bit for bit
// where step_body leaves it; each direction moves CPL elements per load and
// store.  Cells on a face layer or with a link pattern are fixed up after the
// vector loads (wavefront-uniform branches), and a run with a cell that is
// not computed, or that must not write some links, stores cell by cell.
template <int Q, class S, bool ODD, int CPL>
__device__ __forceinline__ void step_body_run(const StepArgs &a, const int L) {
  using CD = Codec<S>;
  using C = typename CD::C;
  using M = typename CD::M;
  using V = typename RunVec<S, CPL>::type;
  constexpr unsigned ES = sizeof(S);
  const int plane = L / a.bpp;
  const int chunk = L - plane * a.bpp;
  const unsigned j = static_cast<unsigned>(chunk * a.block) + threadIdx.x;
  if (j >= static_cast<unsigned>(a.ny * a.nzv)) return;
  // row and run inside the plane (exact: j * nzv < 2^40, see StepArgs::vmagic)
  const unsigned row = static_cast<unsigned>((static_cast<unsigned long long>(j) * a.vmagic) >> 40);
  const int z0 = 1 + CPL * static_cast<int>(j - row * static_cast<unsigned>(a.nzv));
  const int x = plane + 1, y = static_cast<int>(row) + 1;
  const unsigned idx = static_cast<unsigned>(x * a.sx + y * a.sy + z0);
  const unsigned ib = idx * ES;
  const char *f0 = a.fd[0];

  // Face rule (as step_body), in one round of loads.  A run lies on an x or y
  // face with all its cells, so a bounce-back link there is a vector load of
  // the other slot of the link pair; only the first (z = 1) and the last
  // (z = nz) cell of a row lie on a z face, and their links through it come
  // from side loads of one element, issued with the vector loads.
  const bool raw_links = (a.flags & STEP_F_RAW) != 0;
  const unsigned fxy = raw_links ? 0u
                                 : (x == 1 ? 1u : 0u) | (x == a.nx ? 2u : 0u) | (y == 1 ? 4u : 0u) |
                                       (y == a.ny ? 8u : 0u);
  const bool zlo = !raw_links && z0 == 1, zhi = !raw_links && z0 + CPL - 1 == a.nz;
  const bool xyf = __any(fxy != 0u), zf = __any(zlo || zhi);  // wavefront-uniform
  // winning face BC of link k among the faces in fb: slot * 4 + mode, and its value
  auto winner = [&](int k, unsigned fb, C &val) -> int {
    int wn = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int cd = Lat<Q>::c[k][d];
      if (cd == 0) continue;
      const int fc = 2 * d + (cd > 0 ? 0 : 1);  // the face this link enters through
      const int cand = ((fb >> fc) & 1u) ? a.fface[fc] : 0;
      if (cand > wn) {
        wn = cand;
        val = face_val<C>(a, fc, k);
      }
    }
    return wn;
  };
  unsigned bxy = 0, cxy = 0;  // bounce-back / constant links of the x and y faces
  if (xyf) {
#pragma unroll
    for (int k = 1; k < Q; ++k) {
      C unused;
      const int md = winner(k, fxy, unused) & 3;
      if (md == MODE_BOUNCE) bxy |= 1u << k;
      if (md == MODE_CONST) cxy |= 1u << k;
    }
  }
  // the other slot of link k's pair for the cell at byte offset ibc
  auto other = [&](int k, unsigned ibc) -> unsigned {
    const unsigned o = ES * static_cast<unsigned>(offc<Q>(k, a.sx, a.sy));
    return static_cast<unsigned>(ODD ? k : inv_dir<Q>(k)) * a.qs32 + (ODD ? ibc : ibc - o);
  };
  V raw[Q];
  C side[Q];  // links through z = 1 (first cell) or z = nz (last cell): the other slot
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    // normal slot: (x, k) (even) or (x - c_k, inv k) (odd)
    const unsigned o = ES * static_cast<unsigned>(offc<Q>(k, a.sx, a.sy));
    const unsigned on = static_cast<unsigned>(ODD ? inv_dir<Q>(k) : k) * a.qs32 + (ODD ? ib - o : ib);
    if (xyf) {
      const unsigned op = other(k, ib);
      raw[k] = *gload<V>(f0 + (((bxy >> k) & 1u) ? op : on));
    } else {
      raw[k] = *gload<V>(f0 + on);
    }
    if (zf && Lat<Q>::c[k][2] != 0) side[k] = CD::ld(f0, other(k, Lat<Q>::c[k][2] > 0 ? ib : ib + ES * (CPL - 1)));
  }
  unsigned codes = 0;  // CPL cell-type bytes
  const char *cb = reinterpret_cast<const char *>(a.code) + idx;
#pragma unroll
  for (int c = 0; c < CPL; ++c) codes |= static_cast<unsigned>(*gload<unsigned char>(cb + c)) << (8 * c);
  __builtin_amdgcn_sched_barrier(0);
  C v[CPL][Q];
#pragma unroll
  for (int c = 0; c < CPL; ++c)
#pragma unroll
    for (int k = 0; k < Q; ++k) v[c][k] = CD::dec(raw[k][c]);
  if (xyf) {
#pragma unroll
    for (int k = 1; k < Q; ++k) {
      if (!(((bxy | cxy) >> k) & 1u)) continue;
      C val = C(0);
      (void)winner(k, fxy, val);
#pragma unroll
      for (int c = 0; c < CPL; ++c) v[c][k] = ((bxy >> k) & 1u) ? v[c][k] + val : val;
    }
  }
  if (zf) {
    // a link through the z face of an end cell: the z face's BC where it wins
#pragma unroll
    for (int k = 1; k < Q; ++k) {
      const int cz = Lat<Q>::c[k][2];
      if (cz == 0) continue;
      const int c = cz > 0 ? 0 : CPL - 1;
      const bool on_z = cz > 0 ? zlo : zhi;
      const unsigned zbit = cz > 0 ? 16u : 32u;
      C vz = C(0), vxy = C(0);
      const int wz = winner(k, fxy | zbit, vz);
      const int wxy = winner(k, fxy, vxy);
      if (on_z && wz != wxy) {  // the z face wins the link
        v[c][k] = (wz & 3) == MODE_BOUNCE ? side[k] + vz : vz;
      }
    }
  }

  // Cells with a link pattern re-read their links from it.
  bool live[CPL];
  bool bnd_any = false;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const unsigned char code = static_cast<unsigned char>(codes >> (8 * c));
    live[c] = code != CODE_SKIP;
    bnd_any |= live[c] && code != CODE_FLUID;
  }
  unsigned skipm[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) skipm[c] = 0u;
  const bool general = __any(bnd_any);
  if (general && bnd_any) {
    const C *val = static_cast<const C *>(a.val);
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const unsigned char code = static_cast<unsigned char>(codes >> (8 * c));
      if (!live[c] || code == CODE_FLUID) continue;
      const unsigned ibc = ib + ES * c;
      const LinkRec rec = a.rec[a.pat[idx + c]];
      skipm[c] = rec.skip;
#pragma unroll
      for (int k = 1; k < Q; ++k) {
        const int slot = raw_links ? 0 : rec.slot[k];
        const unsigned char md = slot ? a.mode[slot] : static_cast<unsigned char>(MODE_NONE);
        const unsigned o = ES * static_cast<unsigned>(offc<Q>(k, a.sx, a.sy));
        const unsigned on = static_cast<unsigned>(ODD ? inv_dir<Q>(k) : k) * a.qs32 + (ODD ? ibc - o : ibc);
        const unsigned op = static_cast<unsigned>(ODD ? k : inv_dir<Q>(k)) * a.qs32 + (ODD ? ibc : ibc - o);
        const C xk = CD::ld(f0, md == MODE_BOUNCE ? op : on);
        const C sv = val[slot * Q + k];
        v[c][k] = md == MODE_BOUNCE ? xk + sv : (md == MODE_CONST ? sv : xk);
      }
    }
  }

  const bool omega = (a.flags & STEP_F_OMEGA) != 0;
  const bool comp = (a.flags & STEP_F_COMPRESSIBLE) != 0;
  C om[3] = {C(0), C(0), C(0)};
  if (omega) {
#pragma unroll
    for (int d = 0; d < 3; ++d) om[d] = sizeof(C) == 4 ? C(a.om32[d]) : C(a.om64[d]);
  }
  C r[CPL], m[CPL][3];
#pragma unroll
  for (int c = 0; c < CPL; ++c) collide<Q, C>(v[c], C(a.w), comp, omega, om, r[c], m[c]);

  bool plain = true;
#pragma unroll
  for (int c = 0; c < CPL; ++c) plain = plain && live[c] && skipm[c] == 0u;
  char *fw = a.fd[0];
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    // even: slot (x, inv k); odd: slot (x + c_k, k)
    const unsigned o = ES * static_cast<unsigned>(offc<Q>(k, a.sx, a.sy));
    const unsigned ow = static_cast<unsigned>(ODD ? k : inv_dir<Q>(k)) * a.qs32 + (ODD ? ib + o : ib);
    if (plain) {
      V out;
#pragma unroll
      for (int c = 0; c < CPL; ++c) out[c] = CD::enc(v[c][k]);
      *gstore<V>(fw + ow) = out;
    } else {
#pragma unroll
      for (int c = 0; c < CPL; ++c)
        if (live[c] && !((skipm[c] >> k) & 1u)) CD::st(fw, ow + ES * c, v[c][k]);
    }
  }
  if (a.flags & (STEP_F_RHO | STEP_F_VEL)) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      if (!live[c]) continue;
      if (a.flags & STEP_F_RHO) static_cast<M *>(a.rho)[idx + c] = M(r[c]);
      if (a.flags & STEP_F_VEL) {
        M *u = static_cast<M *>(a.u);
#pragma unroll
        for (int d = 0; d < 3; ++d) u[d * a.len + idx + c] = M(comp ? m[c][d] / r[c] : m[c][d]);
      }
    }
  }
}

// LDS slot of a raw storage element: the element itself, or its bits in a
// 32-bit word for 16-bit storage.
template <class S>
struct LdsSlot {
  using type = S;
  static __device__ __forceinline__ S put(S x) { return x; }
  static __device__ __forceinline__ S get(S x) { return x; }
};
template <>
struct LdsSlot<unsigned short> {
  using type = unsigned;
  static __device__ __forceinline__ unsigned put(unsigned short x) { return x; }
  static __device__ __forceinline__ unsigned short get(unsigned x) { return static_cast<unsigned short>(x); }
};
template <>
struct LdsSlot<_Float16> {
  using type = unsigned;
  static __device__ __forceinline__ unsigned put(_Float16 x) {
    unsigned short b;
    __builtin_memcpy(&b, &x, 2);
    return b;
  }
  static __device__ __forceinline__ _Float16 get(unsigned x) {
    const unsigned short b = static_cast<unsigned short>(x);
    _Float16 r;
    __builtin_memcpy(&r, &b, 2);
    return r;
  }
};

// Direction components packed two bits each ((c + 1) << 2k), so that the
// offset of a direction chosen at run time is a few scalar instructions.
template <int Q>
constexpr unsigned long long packed_dir(int d) {
  unsigned long long m = 0;
  for (int k = 0; k < Q; ++k) m |= static_cast<unsigned long long>(Lat<Q>::c[k][d] + 1) << (2 * k);
  return m;
}

// Offset of c_j in the layout for a direction j known at run time.
template <int Q>
__device__ __forceinline__ int offc_rt(int j, int sx, int sy) {
  constexpr unsigned long long PX = packed_dir<Q>(0), PY = packed_dir<Q>(1), PZ = packed_dir<Q>(2);
  const int cx = static_cast<int>((PX >> (2 * j)) & 3ull) - 1;
  const int cy = static_cast<int>((PY >> (2 * j)) & 3ull) - 1;
  const int cz = static_cast<int>((PZ >> (2 * j)) & 3ull) - 1;
  return cx * sx + cy * sy + cz;
}

// LDS-staged step: a workgroup of STEP_TILE lanes takes STEP_TILE consecutive
// cells of a plane (chunks as in step_body, ghost columns included; the last
// chunk of a plane is shorter).  In the AA pattern the slots the tile's cells
// use in direction array j form one contiguous range ("row" j): cells
// idx0 .. idx0 + n - 1 on even steps, shifted by the offset of c_j on odd
// steps.  So the kernel
//   1. loads every row into LDS with 16-byte loads,
//   2. lets each lane collide its own cell from LDS (one cell per lane, the
//      same link resolution and collision as step_body) and write its
//      outgoing populations into the rows in place, in its own column,
//   3. stores the rows back to the ranges they came from with 16-byte stores.
// Column t of the rows holds exactly the slots cell idx0 + t reads and
// writes: incoming k is row k (even) or row inv k (odd), outgoing k goes to
// row inv k (even) or row k (odd).  Each slot belongs to exactly one cell,
// so tiles never touch each other's slots, and cells that are not computed
// keep their slots as loaded.  A bounce-back link k reads the other slot of
// its pair (even: array inv k at x - c_k; odd: array k at x), which is the
// outgoing row of k at column t - (offset of c_k): from LDS if that column
// lies in the tile, else from global memory (read before the first barrier
// for face cells, after it for link-pattern cells; the slot belongs to a
// cell that is not computed, so its value does not change during the step).
template <int Q, class S, bool ODD>
__device__ __forceinline__ void step_body_tile(const StepArgs &a, const int L) {
  using CD = Codec<S>;
  using C = typename CD::C;
  using M = typename CD::M;
  constexpr int TB = STEP_TILE;
  constexpr int NW = TB / 64;             // wavefronts per workgroup
  constexpr int RPW = (Q + NW - 1) / NW;  // rows per wavefront, at most
  constexpr int EPL = TB / 64;            // row elements per lane
  constexpr unsigned ES = sizeof(S);
  static_assert(TB % 64 == 0 && (EPL & (EPL - 1)) == 0, "tile: a power-of-two multiple of 64 cells");
  // LDS slots hold the raw storage bits; 16-bit storage gets 32-bit slots
  // (sub-dword LDS writes from neighbouring lanes collide in a bank)
  using LS = typename LdsSlot<S>::type;
  typedef S V4 __attribute__((ext_vector_type(EPL), aligned(sizeof(S))));  // global: element alignment
  typedef LS L4 __attribute__((ext_vector_type(EPL)));                      // LDS: EPL slots aligned
  __shared__ __attribute__((aligned(16 * sizeof(LS)))) LS row[Q][TB];
  auto rd = [&](int j, int col) -> C { return CD::dec(LdsSlot<S>::get(row[j][col])); };
  const int t = static_cast<int>(threadIdx.x);
  const int plane = L / a.bpp;
  const int chunk = L - plane * a.bpp;
  const int c0 = chunk * TB;                // first cell of the tile in the plane's rows
  const int n = min(TB, a.ny * a.sy - c0);  // cells in the tile
  const long long idx0 = static_cast<long long>(plane + 1) * a.sx + a.sy + c0;
  const long long len = a.len;
  const char *f0 = a.fd[0];
  // first cell of row j's range (j at run time: scalar arithmetic)
  auto row_base = [&](int j) -> long long { return ODD ? idx0 + offc_rt<Q>(j, a.sx, a.sy) : idx0; };
  auto in_row = [](int k) -> int { return ODD ? inv_dir<Q>(k) : k; };   // incoming k, column t
  auto out_row = [](int k) -> int { return ODD ? k : inv_dir<Q>(k); };  // outgoing k, column t

  // this lane's cell
  const bool in_tile = t < n;
  const unsigned off = static_cast<unsigned>(c0 + (in_tile ? t : 0));
  const unsigned rw = static_cast<unsigned>((static_cast<unsigned long long>(off) * a.zmagic) >> 40);
  const int zi = static_cast<int>(off - rw * static_cast<unsigned>(a.sy));
  const int x = plane + 1, y = static_cast<int>(rw) + 1;
  const unsigned idx = static_cast<unsigned>(idx0 + (in_tile ? t : 0));
  const bool column = in_tile && zi != 0 && zi != a.sy - 1;  // not a ghost column
  const unsigned char code = column ? *gload<unsigned char>(reinterpret_cast<const char *>(a.code) + idx) : CODE_SKIP;

  // 1. rows into LDS: wavefront w moves rows w, w + NW, w + 2 NW, ..., lane
  // t elements e .. e + EPL - 1 (e = EPL (t % 64)) of each; every load is issued
  // before the first wait.  A vector whose range leaves the lattice's cells
  // [0, len) (only at the two ends of the lattice) is loaded from cell 0 and
  // then replaced element by element; elements past the tile's n cells are
  // loaded but never used.
  const int e = EPL * (t & 63);
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  {
    V4 xr[RPW];
    bool vok[RPW];
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const int j = w + i * NW;
      const long long g = j < Q ? row_base(j) + e : 0;
      vok[i] = g >= 0 && g + EPL <= len;
      const unsigned gi = vok[i] ? static_cast<unsigned>(g) : 0u;
      if (j < Q) xr[i] = *gload<V4>(f0 + static_cast<unsigned>(j) * a.qs32 + gi * ES);
    }
    bool fix = false;
#pragma unroll
    for (int i = 0; i < RPW; ++i) fix = fix || (w + i * NW < Q && !vok[i]);
    if (fix) {
#pragma unroll
      for (int i = 0; i < RPW; ++i) {
        const int j = w + i * NW;
        if (j >= Q || vok[i]) continue;
        const long long g = row_base(j) + e;
#pragma unroll
        for (int q = 0; q < EPL; ++q)
          if (g + q >= 0 && g + q < len)
            xr[i][q] = *gload<S>(f0 + static_cast<unsigned>(j) * a.qs32 + static_cast<unsigned>(g + q) * ES);
      }
    }
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const int j = w + i * NW;
      if (j < Q) {
        L4 v4;
#pragma unroll
        for (int q = 0; q < EPL; ++q) v4[q] = LdsSlot<S>::put(xr[i][q]);
        *reinterpret_cast<L4 *>(&row[j][e]) = v4;
      }
    }
  }

  // face rule (as step_body)
  const bool raw = (a.flags & STEP_F_RAW) != 0;
  const unsigned fb = (!column || raw) ? 0u
                                       : (x == 1 ? 1u : 0u) | (x == a.nx ? 2u : 0u) | (y == 1 ? 4u : 0u) |
                                             (y == a.ny ? 8u : 0u) | (zi == 1 ? 16u : 0u) | (zi == a.nz ? 32u : 0u);
  auto winner = [&](int k, C &val) -> int {
    int wn = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int cd = Lat<Q>::c[k][d];
      if (cd == 0) continue;
      const int fc = 2 * d + (cd > 0 ? 0 : 1);
      const int cand = ((fb >> fc) & 1u) ? a.fface[fc] : 0;
      if (cand > wn) {
        wn = cand;
        val = face_val<C>(a, fc, k);
      }
    }
    return wn;
  };
  const bool faces = __any(fb != 0u);
  unsigned bmask = 0, cmask = 0;
  if (faces) {
#pragma unroll
    for (int k = 1; k < Q; ++k) {
      C unused;
      const int md = winner(k, unused) & 3;
      if (md == MODE_BOUNCE) bmask |= 1u << k;
      if (md == MODE_CONST) cmask |= 1u << k;
    }
  }
  // the other slot of link k's pair: the outgoing row of k at column t - (offset of c_k)
  auto other_col = [&](int k) -> int { return t - offc<Q>(k, a.sx, a.sy); };
  auto other_in_tile = [&](int k) -> bool {
    const int oc = other_col(k);
    return oc >= 0 && oc < n;
  };
  auto other_global = [&](int k) -> C {
    const int j = out_row(k);
    const long long g = (ODD ? idx0 + offc<Q>(j, a.sx, a.sy) : idx0) + other_col(k);
    return CD::ld(f0, static_cast<unsigned>(j) * a.qs32 + static_cast<unsigned>(g) * ES);
  };
  C side[Q];  // bounce-back values of face cells from outside the tile
  if (faces) {
#pragma unroll
    for (int k = 1; k < Q; ++k)
      if (((bmask >> k) & 1u) && !other_in_tile(k)) side[k] = other_global(k);
  }
  __syncthreads();

  // 2. collide this lane's cell from the rows
  const int tc = in_tile ? t : 0;
  C v[Q];
#pragma unroll
  for (int k = 0; k < Q; ++k) v[k] = rd(in_row(k), tc);
  if (faces) {
#pragma unroll
    for (int k = 1; k < Q; ++k) {
      if (!(((bmask | cmask) >> k) & 1u)) continue;
      C val = C(0);
      (void)winner(k, val);
      if ((bmask >> k) & 1u) {
        const C o = other_in_tile(k) ? rd(out_row(k), other_col(k)) : side[k];
        v[k] = o + val;
      } else {
        v[k] = val;
      }
    }
  }
  const bool live = code != CODE_SKIP;
  const bool bnd = live && code != CODE_FLUID;
  const bool general = __any(bnd);
  unsigned skipm = 0;
  if (general && bnd) {
    const LinkRec rec = a.rec[*gload<unsigned short>(reinterpret_cast<const char *>(a.pat) + 2u * idx)];
    skipm = rec.skip;
    const C *val = static_cast<const C *>(a.val);
#pragma unroll
    for (int k = 1; k < Q; ++k) {
      const int slot = raw ? 0 : rec.slot[k];
      const unsigned char md = slot ? a.mode[slot] : static_cast<unsigned char>(MODE_NONE);
      C xk = rd(in_row(k), t);
      if (md == MODE_BOUNCE)
        xk = other_in_tile(k) ? rd(out_row(k), other_col(k)) : other_global(k);
      const C sv = val[slot * Q + k];
      v[k] = md == MODE_BOUNCE ? xk + sv : (md == MODE_CONST ? sv : xk);
    }
  }
  // every lane has read what it needs from the rows before any lane writes them
  __syncthreads();
  C r, m[3];
  const bool omega = (a.flags & STEP_F_OMEGA) != 0;
  const bool comp = (a.flags & STEP_F_COMPRESSIBLE) != 0;
  C om[3] = {C(0), C(0), C(0)};
  if (omega) {
#pragma unroll
    for (int d = 0; d < 3; ++d) om[d] = sizeof(C) == 4 ? C(a.om32[d]) : C(a.om64[d]);
  }
  collide<Q, C>(v, C(a.w), comp, omega, om, r, m);
  if (live) {
#pragma unroll
    for (int k = 0; k < Q; ++k)
      if (!((skipm >> k) & 1u)) row[out_row(k)][t] = LdsSlot<S>::put(CD::enc(v[k]));
    if (a.flags & STEP_F_RHO) *gstore<M>(static_cast<char *>(a.rho) + sizeof(M) * idx) = M(r);
    if (a.flags & STEP_F_VEL) {
      char *u = static_cast<char *>(a.u);
#pragma unroll
      for (int d = 0; d < 3; ++d)
        *gstore<M>(u + sizeof(M) * (static_cast<size_t>(d) * a.len + idx)) = M(comp ? m[d] / r : m[d]);
    }
  }
  __syncthreads();

  // 3. rows back to their ranges
  {
    char *fw = a.fd[0];
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const int j = w + i * NW;
      if (j >= Q || e >= n) continue;
      const long long g = row_base(j) + e;
      if (e + EPL <= n && g >= 0 && g + EPL <= len) {
        const L4 l4 = *reinterpret_cast<const L4 *>(&row[j][e]);
        V4 x4;
#pragma unroll
        for (int q = 0; q < EPL; ++q) x4[q] = LdsSlot<S>::get(l4[q]);
        *gstore<V4>(fw + static_cast<unsigned>(j) * a.qs32 + static_cast<unsigned>(g) * ES) = x4;
      } else {
#pragma unroll
        for (int q = 0; q < EPL; ++q)
          if (e + q < n && g + q >= 0 && g + q < len)
            *gstore<S>(fw + static_cast<unsigned>(j) * a.qs32 + static_cast<unsigned>(g + q) * ES) =
                LdsSlot<S>::get(row[j][e + q]);
      }
    }
  }
}

template <int Q, class S, bool ODD>
__global__ __launch_bounds__(STEP_TILE) void k_step_tile(const StepArgs *args) {
  const StepArgs &a = *args;
  const int b = ODD ? xcd_reverse(static_cast<int>(blockIdx.x), a.nblocks) : static_cast<int>(blockIdx.x);
  step_body_tile<Q, S, ODD>(a, a.remap > 1 ? xcd_remap(b, a.nblocks, a.remap) : b);
}

// The arguments live in device memory, not in the kernel-argument segment:
// loads from the argument segment are invariant, and the compiler hoists all
// of them (27 direction pointers, face tables) into scalar registers, runs
// out of them and spills into vector registers (D3Q19 FP32: 103 VGPRs
// instead of 70).  Loads through a plain pointer stay where they are used.
//
// Odd steps take the chunks in reverse order (zig-zag over steps), each XCD
// over its own share (xcd_reverse): a step then starts where the previous
// one ended, on the populations written last, which its XCD's L2 still
// holds.  Every cell's reads and writes in one step are its own (AA
// pattern), so the order of the chunks does not change results.
template <int Q, class S, bool ODD, bool PER, bool WIDE, int CL>
__global__ __launch_bounds__(STEP_BLOCK) void k_step(const StepArgs *args) {
  const StepArgs &a = *args;
  const int b = ODD ? xcd_reverse(static_cast<int>(blockIdx.x), a.nblocks) : static_cast<int>(blockIdx.x);
  step_body<Q, S, ODD, PER, WIDE, CL>(a, a.remap > 1 ? xcd_remap(b, a.nblocks, a.remap) : b);
}

// The matrix-core closure variant, held to 5 waves per SIMD (96 VGPRs).
template <int Q, class S, bool ODD, bool PER, bool WIDE>
__global__ __launch_bounds__(STEP_BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void k_step_mfma(
    const StepArgs *args) {
  const StepArgs &a = *args;
  const int b = ODD ? xcd_reverse(static_cast<int>(blockIdx.x), a.nblocks) : static_cast<int>(blockIdx.x);
  step_body<Q, S, ODD, PER, WIDE, 2>(a, a.remap > 1 ? xcd_remap(b, a.nblocks, a.remap) : b);
}

// The step with CPL cells per lane (step_body_run); same chunk order.  At
// most STEP_RUN_BLOCK lanes per workgroup: four cells per lane hold about
// 130 VGPRs, more than a 1024-lane workgroup allows without spilling.
template <int Q, class S, bool ODD, int CPL>
__global__ __launch_bounds__(STEP_RUN_BLOCK) void k_step_run(const StepArgs *args) {
  const StepArgs &a = *args;
  const int b = ODD ? xcd_reverse(static_cast<int>(blockIdx.x), a.nblocks) : static_cast<int>(blockIdx.x);
  step_body_run<Q, S, ODD, CPL>(a, a.remap > 1 ? xcd_remap(b, a.nblocks, a.remap) : b);
}

// The incoming populations of every computed cell for the next step.
template <int Q, class S, bool ODD, bool PER>
__global__ __launch_bounds__(STEP_BLOCK) void k_peek(const StepArgs *args, void *out) {
  step_body<Q, S, ODD, PER, false, 0, true>(*args, static_cast<int>(blockIdx.x), out);
}

// asl::ComputeSurfaceForceMap / ComputeSurfaceFluxMap (src/num/aslLBGKBC.cxx:
// 623-736) from two host-layout views: post[i * len + x] = f_i of cell x after
// the last collision, in_[i * len + x] = f_i that cell x receives next (the
// value ASL's BCs leave at ghost x - c_i).  For a ghost cell x (map <= 0) next
// to a computation node y = x + c_i (map > 0), ASL's f_i(x) is in_[i][y];
// for a computation node x next to a ghost y, ASL's f_inv(i)(y) is
// in_[inv(i)][x].  Cells with |map| >= 0.9999 get 0 (ASL leaves them as they
// were); map reads outside the block are -0.9 (a ghost node), as in ASL.
template <int Q, class S>
__global__ __launch_bounds__(BLOCK) void k_surface(const SurfaceArgs a) {
  using C = typename Codec<S>::C;
  using M = typename Codec<S>::M;
  const int x = blockIdx.x * BLOCK + threadIdx.x;
  if (x >= a.len) return;
  const M *post = static_cast<const M *>(a.post), *in = static_cast<const M *>(a.in);
  M *out = static_cast<M *>(a.out);
  const float m0 = a.map[x];
  C acc[3] = {C(0), C(0), C(0)};
  if (__builtin_fabsf(m0) < .9999f) {
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      const int kb = inv_dir<Q>(i);
      const long long y = static_cast<long long>(x) + offc<Q>(i, a.sx, a.sy);
      const float my = (y < 0 || y > a.len - 1) ? -.9f : a.map[y];
      const bool b0 = m0 <= 0.f && my > 0.f;  // isGhostNode(0) && isComputationNode(i)
      const bool bi = m0 > 0.f && my <= 0.f;  // isGhostNode(i) && isComputationNode(0)
      const long long yy = b0 ? y : x;        // the computation node of the link
      const C f0 = b0 ? C(in[static_cast<long long>(i) * a.len + yy]) : C(post[static_cast<long long>(i) * a.len + x]);
      const C fI = b0 ? C(post[static_cast<long long>(kb) * a.len + yy]) : C(in[static_cast<long long>(kb) * a.len + x]);
      if (a.flux) {
        const C wi = C(Lat<Q>::w(i));
        acc[0] = acc[0] + (b0 ? (f0 - fI) * wi : C(0));
        acc[0] = acc[0] + (bi ? (fI - f0) * wi : C(0));
      } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const C wc = C(Lat<Q>::w(i) * double(Lat<Q>::c[i][d]));
          acc[d] = acc[d] + (b0 ? ((f0 + fI) - C(2.)) * wc : C(0));
          acc[d] = acc[d] + (bi ? ((-f0 - fI) + C(2.)) * wc : C(0));
        }
      }
    }
  }
  if (a.flux) {
    out[x] = M(acc[0]);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) out[static_cast<long long>(d) * a.len + x] = M(acc[d]);
  }
}

// Visualisation fills of rho / u (VisArgs).  Reads come only from cells with
// map > 0 and writes go only to cells with map <= 0, so the update in place
// is free of races.  The operations follow the CPU oracle's restatement of
// ASL's kernels (src/num/aslBasicBC.cxx:124-145, 167-227, 283-351) in its
// order, none contracted (the library is built with -ffp-contract=off):
//   NOSLIP_VEL: v += select(0 - v(x + c_i), b_i) * lc_i, n += select(lc_i, b_i),
//               v(x) = v / max(n, .01), over i = 1..Q-1, b_i = map(x + c_i) > 0;
//   NOSLIP_RHO: g_k = sum_i map(x + c_i) (gc_i c_i[k]) over c_i[k] != 0,
//               g /= |g|, p_i = g . c_i/|c_i|, d_i = (p_i + 1)^2,
//               n = max(sum_i select(d_i, b_i), .001),
//               r = sum_i select((rho(x + c_i) - p_i * 0) d_i, b_i),
//               rho(x) = r / n where n > 0.1;
//   VELOCITY:   u(x) = u0, or axis x (position - centre).
template <int Q, class S>
__global__ __launch_bounds__(BLOCK) void k_vis(const VisArgs a) {
  using M = typename Codec<S>::M;
  const int g = blockIdx.x * BLOCK + threadIdx.x;
  if (g >= a.len) return;
  const float mg = a.map[g];
  if (!(mg <= 0.f && mg > -.9999f)) return;
  M *rho = static_cast<M *>(a.rho), *u = static_cast<M *>(a.u);
  const long long len = a.len;
  auto mx = [&](long long y) { return (y < 0 || y > len - 1) ? -.9f : a.map[y]; };
  auto wrap = [&](long long y) { return (len + y) % len; };
  if (a.kind == ASLHIP_VIS_VELOCITY) {
    M uw[3] = {M(a.u0[0]), M(a.u0[1]), M(a.u0[2])};
    if (a.has_rot) {
      // Index2PositionACL over the ghost-inclusive block (src/math/aslIndex2Position.cxx:64-108)
      const int I0 = g / a.sx, rem = g - I0 * a.sx, I1 = rem / a.sy, I2 = rem - I1 * a.sy;
      const int I[3] = {I0, I1, I2};
      M pos[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) pos[d] = M(a.pos0[d] - a.dx) + M(a.dx) * M(I[d]);
      const M d0 = pos[0] - M(a.rot[3]), d1 = pos[1] - M(a.rot[4]), d2 = pos[2] - M(a.rot[5]);
      uw[0] = M(a.rot[1]) * d2 - M(a.rot[2]) * d1;
      uw[1] = M(a.rot[2]) * d0 - M(a.rot[0]) * d2;
      uw[2] = M(a.rot[0]) * d1 - M(a.rot[1]) * d0;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) u[c * len + g] = uw[c];
    return;
  }
  if constexpr (Q != 27) {
    using LT = Lat<Q>;
    if (a.kind == ASLHIP_VIS_NOSLIP_VEL) {
      M v[3] = {M(0), M(0), M(0)}, n = M(0);
#pragma unroll
      for (int i = 1; i < Q; ++i) {
        const long long y = g + offc<Q>(i, a.sx, a.sy);
        const bool b = mx(y) > 0.f;
        const M l = M(LT::lc(i));
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = v[c] + (b ? M(0) - u[c * len + wrap(y)] : M(0)) * l;
        n = n + (b ? l : M(0));
      }
      const M den = n > M(.01) ? n : M(.01);
#pragma unroll
      for (int c = 0; c < 3; ++c) u[c * len + g] = v[c] / den;
    } else if (a.kind == ASLHIP_VIS_NOSLIP_RHO) {
      M mv[Q];
      bool bv[Q];
#pragma unroll
      for (int i = 1; i < Q; ++i) {
        const float m = mx(g + offc<Q>(i, a.sx, a.sy));
        mv[i] = M(m);
        bv[i] = m > 0.f;
      }
      M gn[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        M acc = M(0);
        bool first = true;
#pragma unroll
        for (int i = 1; i < Q; ++i) {
          if (LT::c[i][k] == 0) continue;
          const M t = mv[i] * M(LT::gc(i) * double(LT::c[i][k]));
          acc = first ? t : acc + t;
          first = false;
        }
        gn[k] = acc;
      }
      M l2 = gn[0] * gn[0];
      l2 = l2 + gn[1] * gn[1];
      l2 = l2 + gn[2] * gn[2];
      const M ln = sizeof(M) == 4 ? M(__builtin_sqrtf(float(l2))) : M(__builtin_sqrt(double(l2)));
#pragma unroll
      for (int k = 0; k < 3; ++k) gn[k] = gn[k] / ln;
      M pr[Q], dc[Q];
      M n = M(0);
#pragma unroll
      for (int i = 1; i < Q; ++i) {
        const double cl = __builtin_sqrt(double(LT::c[i][0] * LT::c[i][0] + LT::c[i][1] * LT::c[i][1] +
                                                LT::c[i][2] * LT::c[i][2]));
        M p = M(0);
        bool first = true;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if (LT::c[i][d] == 0) continue;
          const M t = gn[d] * M(double(LT::c[i][d]) / cl);
          p = first ? t : p + t;
          first = false;
        }
        pr[i] = p;
        const M e = p + M(1.);
        dc[i] = e * e;
        n = n + (bv[i] ? dc[i] : M(0));
      }
      n = n > M(.001) ? n : M(.001);
      M r = M(0);
#pragma unroll
      for (int i = 1; i < Q; ++i) {
        const long long y = g + offc<Q>(i, a.sx, a.sy);
        r = r + (bv[i] ? (rho[wrap(y)] - pr[i] * M(0)) * dc[i] : M(0));
      }
      if (n > M(.1)) rho[g] = r / n;
    }
  }
}

// Many lattices in one launch: workgroup b works on chunk ent[b].y of
// lattice ent[b].x, whose descriptor lives in device memory.  One chunk per
// workgroup, with no loop around the body: a loop over several chunks kept
// values live across iterations and raised the kernel from 70 to 99 VGPRs,
// which cut the waves per SIMD from 7 to 4 at 256-lane workgroups.
// Odd steps take the entries in reverse order, as k_step does.
template <int Q, class S, bool ODD, bool PER, bool WIDE>
__global__ __launch_bounds__(STEP_BLOCK) void k_step_batch(const StepArgs *desc, const uint2 *ent) {
  const int b = static_cast<int>(blockIdx.x);
  const uint2 e = ent[ODD ? xcd_reverse(b, static_cast<int>(gridDim.x)) : b];
  step_body<Q, S, ODD, PER, WIDE>(desc[e.x], static_cast<int>(e.y));
}

// The same with LDS-staged tiles (entries are tiles of STEP_TILE cells).
template <int Q, class S, bool ODD>
__global__ __launch_bounds__(STEP_TILE) void k_step_batch_tile(const StepArgs *desc, const uint2 *ent) {
  const int b = static_cast<int>(blockIdx.x);
  const uint2 e = ent[ODD ? xcd_reverse(b, static_cast<int>(gridDim.x)) : b];
  step_body_tile<Q, S, ODD>(desc[e.x], static_cast<int>(e.y));
}

// Closure features of every interior cell of a periodic D3Q19 lattice: the
// populations the next step would collide (AA addressing of that step's
// parity, periodic wrap), their moments as collide() forms them, and the 7
// inputs of the closure network, written to out[7 * cell + p] with cells in
// ASL order over the interior.
template <class S, bool ODD>
__global__ __launch_bounds__(256) void k_closure_features(const StepArgs *args, float *out) {
  using CD = Codec<S>;
  constexpr int Q = 19;
  constexpr unsigned ES = sizeof(S);
  const StepArgs &a = *args;
  const long long c = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (c >= static_cast<long long>(a.nx) * a.ny * a.nz) return;
  const int z = static_cast<int>(c % a.nz) + 1;
  const int y = static_cast<int>((c / a.nz) % a.ny) + 1;
  const int x = static_cast<int>(c / (static_cast<long long>(a.ny) * a.nz)) + 1;
  const unsigned idx = static_cast<unsigned>(x * a.sx + y * a.sy + z);
  float v[Q];
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    // each direction from its own base (a.qs32 is 0 on lattices over 4 GiB)
    if (ODD) {
      const unsigned src = static_cast<unsigned>(wrap1(x - Lat<Q>::c[k][0], a.nx) * a.sx +
                                                 wrap1(y - Lat<Q>::c[k][1], a.ny) * a.sy +
                                                 wrap1(z - Lat<Q>::c[k][2], a.nz));
      v[k] = CD::ld(a.fd[inv_dir<Q>(k)], ES * src);
    } else {
      v[k] = CD::ld(a.fd[k], ES * idx);
    }
  }
  float r, m[3], f7[7];
  moments<Q, float>(v, r, m);
  closure_features<Q>(v, r, m, f7);
#pragma unroll
  for (int p = 0; p < 7; ++p) out[7 * c + p] = f7[p];
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
  constexpr unsigned ES = sizeof(S);
  auto fk = [&](int k) -> char * { return a.f + k * a.qs_bytes; };
  C coef = C(0), r = C(0);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    if (!((l.comp >> i) & 1u)) continue;
    const int kb = inv_dir<Q>(i);
    const int p = l.g + offc<Q>(i, a.sx, a.sy);
    const C fi = a.after_odd ? CD::ld(fk(kb), ES * l.g) : CD::ld(fk(i), ES * p);
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
      CD::st(fk(i), ES * p, val);  // next step is even: p reads slot (p, i)
    else
      CD::st(fk(kb), ES * l.g, val);  // next step is odd: p reads slot (g, inv i)
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
  uu = fmaT(C(1.5) * u[1], u[1], uu);
  uu = fmaT(C(1.5) * u[2], u[2], uu);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const C ai = i == 0 ? C(0) : dot_c<Q>(u, i);
    const C fi = fmaT(C(4.5), ai * ai, fmaT(C(3.), ai, C(1.))) - uu;
    CD::st(a.f + inv_dir<Q>(i) * a.qs_bytes, static_cast<unsigned>(idx) * sizeof(S), r * fi);
  }
  static_cast<M *>(a.rho)[idx] = M(r);
#pragma unroll
  for (int d = 0; d < 3; ++d) static_cast<M *>(a.u)[d * a.len + idx] = M(u[d]);
}

template <int Q, class S>
__global__ __launch_bounds__(BLOCK) void k_write_f(const WriteFArgs a) {
  using CD = Codec<S>;
  using C = typename CD::C;
  using M = typename CD::M;
  const int idx = blockIdx.x * BLOCK + threadIdx.x;
  if (idx >= a.len) return;
  const M *in = static_cast<const M *>(a.in);
  C v[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    v[i] = C(in[static_cast<long long>(i) * a.len + idx]);
    CD::st(a.f + inv_dir<Q>(i) * a.qs_bytes, static_cast<unsigned>(idx) * sizeof(S), v[i]);
  }
  C r, m[3];
  moments<Q, C>(v, r, m);
  static_cast<M *>(a.rho)[idx] = M(r);
#pragma unroll
  for (int d = 0; d < 3; ++d) static_cast<M *>(a.u)[d * a.len + idx] = M(a.compressible ? m[d] / r : m[d]);
}

// Cell whose slot i holds population i of cell idx after an odd step: idx + c_i,
// wrapped into the interior on periodic lattices.
template <int Q>
__device__ __forceinline__ int odd_slot(int idx, int i, int sx, int sy, int per, int nx, int ny, int nz) {
  if (!per) return idx + offc<Q>(i, sx, sy);
  const int x = idx / sx, rem = idx - x * sx, y = rem / sy, z = rem - y * sy;
  if (x < 1 || x > nx || y < 1 || y > ny || z < 1 || z > nz) return idx + offc<Q>(i, sx, sy);
  return wrap1(x + Lat<Q>::c[i][0], nx) * sx + wrap1(y + Lat<Q>::c[i][1], ny) * sy + wrap1(z + Lat<Q>::c[i][2], nz);
}

// The post-collision populations in ASL's view (f_i of cell x), from the AA
// state.  Slots outside the array (outer ghost layer) read as NaN.
template <int Q, class S>
__global__ __launch_bounds__(BLOCK) void k_unpermute(const UnpermuteArgs a) {
  using CD = Codec<S>;
  using M = typename CD::M;
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  if (c >= a.count) return;
  const int idx = a.first + c;
  M *out = static_cast<M *>(a.out);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    M v;
    if (a.after_odd) {
      const int src = odd_slot<Q>(idx, i, a.sx, a.sy, a.per, a.nx, a.ny, a.nz);
      v = (src >= 0 && src < a.len)
              ? M(CD::ld(a.f + i * a.qs_bytes, static_cast<unsigned>(src) * sizeof(S)))
              : M(__builtin_nan(""));
    } else {
      v = M(CD::ld(a.f + inv_dir<Q>(i) * a.qs_bytes, static_cast<unsigned>(idx) * sizeof(S)));
    }
    out[static_cast<long long>(i) * a.count + c] = v;
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
        fi = double(CD::ld(a.f + i * a.qs_bytes,
                           static_cast<unsigned>(odd_slot<Q>(idx, i, a.sx, a.sy, a.per, a.nx, a.ny, a.nz)) * sizeof(S)));
      else
        fi = double(CD::ld(a.f + inv_dir<Q>(i) * a.qs_bytes, static_cast<unsigned>(idx) * sizeof(S)));
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

template <int Q, class S, bool ODD, bool PER>
hipError_t step_w(bool wide, const StepLaunch &a, hipStream_t s) {
  const dim3 grid(a.nblocks), block(a.block);
  if (a.tile) {
    // the host sends only these: non-periodic, under 4 GiB, no per-cell relaxation, one cell per lane
    if (PER || wide || a.variant || a.cpl != 1 || a.block != STEP_TILE) return hipErrorInvalidValue;
    if constexpr (!PER) hipLaunchKernelGGL((k_step_tile<Q, S, ODD>), grid, block, 0, s, a.args);
    return hipGetLastError();
  }
  if (a.cpl > 1) {
    if (PER || wide || a.variant) return hipErrorInvalidValue;  // the host sends these one cell per lane
    if (a.cpl == 2)
      hipLaunchKernelGGL((k_step_run<Q, S, ODD, 2>), grid, block, 0, s, a.args);
    else if (a.cpl == 4)
      hipLaunchKernelGGL((k_step_run<Q, S, ODD, 4>), grid, block, 0, s, a.args);
    else
      return hipErrorInvalidValue;
    return hipGetLastError();
  }
  if (a.variant == 2) {  // the closure on the matrix cores: D3Q19 with FP32 collision (checked by the host)
    if constexpr (Q == 19 && sizeof(typename Codec<S>::C) == 4) {
      if (wide)
        hipLaunchKernelGGL((k_step_mfma<Q, S, ODD, PER, true>), grid, block, 0, s, a.args);
      else
        hipLaunchKernelGGL((k_step_mfma<Q, S, ODD, PER, false>), grid, block, 0, s, a.args);
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;
    }
  }
  if (a.variant == 1) {  // LBGKTurbulence or the FP32 closure
    if (wide)
      hipLaunchKernelGGL((k_step<Q, S, ODD, PER, true, 1>), grid, block, 0, s, a.args);
    else
      hipLaunchKernelGGL((k_step<Q, S, ODD, PER, false, 1>), grid, block, 0, s, a.args);
    return hipGetLastError();
  }
  if (wide)
    hipLaunchKernelGGL((k_step<Q, S, ODD, PER, true, 0>), grid, block, 0, s, a.args);
  else
    hipLaunchKernelGGL((k_step<Q, S, ODD, PER, false, 0>), grid, block, 0, s, a.args);
  return hipGetLastError();
}

template <int Q, class S, bool ODD, bool PER>
hipError_t batch_w(bool wide, const BatchLaunch &b, hipStream_t s) {
  const dim3 grid(b.nblocks), block(b.block);
  if (b.tile) {
    if (PER || wide || b.block != STEP_TILE) return hipErrorInvalidValue;  // the host sends only these
    if constexpr (!PER) hipLaunchKernelGGL((k_step_batch_tile<Q, S, ODD>), grid, block, 0, s, b.desc, b.ent);
    return hipGetLastError();
  }
  if (wide)
    hipLaunchKernelGGL((k_step_batch<Q, S, ODD, PER, true>), grid, block, 0, s, b.desc, b.ent);
  else
    hipLaunchKernelGGL((k_step_batch<Q, S, ODD, PER, false>), grid, block, 0, s, b.desc, b.ent);
  return hipGetLastError();
}

template <int Q, class S>
hipError_t step_q_s(bool odd, bool per, bool wide, const StepLaunch &a, hipStream_t s) {
  if (per) return odd ? step_w<Q, S, true, true>(wide, a, s) : step_w<Q, S, false, true>(wide, a, s);
  return odd ? step_w<Q, S, true, false>(wide, a, s) : step_w<Q, S, false, false>(wide, a, s);
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
  static hipError_t run(bool odd, bool per, bool wide, const StepLaunch &a, hipStream_t s) {
    return step_q_s<Q, S>(odd, per, wide, a, s);
  }
};
template <int Q, class S>
struct BatchF {
  static hipError_t run(bool odd, bool per, bool wide, const BatchLaunch &b, hipStream_t s) {
    if (per) return odd ? batch_w<Q, S, true, true>(wide, b, s) : batch_w<Q, S, false, true>(wide, b, s);
    return odd ? batch_w<Q, S, true, false>(wide, b, s) : batch_w<Q, S, false, false>(wide, b, s);
  }
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
struct PeekF {
  static hipError_t run(bool odd, bool per, const StepLaunch &a, void *out, hipStream_t s) {
    const dim3 grid(a.nblocks), block(a.block);
    if (per) {
      if (odd)
        hipLaunchKernelGGL((k_peek<Q, S, true, true>), grid, block, 0, s, a.args, out);
      else
        hipLaunchKernelGGL((k_peek<Q, S, false, true>), grid, block, 0, s, a.args, out);
    } else {
      if (odd)
        hipLaunchKernelGGL((k_peek<Q, S, true, false>), grid, block, 0, s, a.args, out);
      else
        hipLaunchKernelGGL((k_peek<Q, S, false, false>), grid, block, 0, s, a.args, out);
    }
    return hipGetLastError();
  }
};
template <int Q, class S>
struct SurfaceF {
  static hipError_t run(const SurfaceArgs &a, hipStream_t s) {
    hipLaunchKernelGGL((k_surface<Q, S>), dim3((a.len + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
  }
};
template <int Q, class S>
struct VisF {
  static hipError_t run(const VisArgs &a, hipStream_t s) {
    if (Q == 27 && a.kind != ASLHIP_VIS_VELOCITY) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_vis<Q, S>), dim3((a.len + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
  }
};
template <int Q, class S>
struct WriteFF {
  static hipError_t run(const WriteFArgs &a, hipStream_t s) {
    hipLaunchKernelGGL((k_write_f<Q, S>), dim3((a.len + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
  }
};
template <int Q, class S>
struct UnpermuteF {
  static hipError_t run(const UnpermuteArgs &a, hipStream_t s) {
    if (a.first < 0 || a.count < 1 || a.first + a.count > a.len) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_unpermute<Q, S>), dim3((a.count + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a);
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

// Staging memory of aslhip_read_f_noalloc: a file-scope device array, so the
// entry point needs no allocation.  In host code `scratch` only names the
// symbol; its device address on each GPU comes from hipGetSymbolAddress (the
// array's name is never passed to a kernel or a copy).
static __device__ __align__(256) unsigned char scratch[ASLHIP_NOALLOC_SCRATCH_BYTES];
namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex scratch_mu;
void *scratch_ptr[SCRATCH_MAX_DEVICES] = {};  // its address on each device (null: not looked up yet)
}  // namespace

hipError_t noalloc_scratch(void **ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    scratch_ptr[dev] = p;
  }
  *ptr = scratch_ptr[dev];
  return hipSuccess;
}

int reduce_blocks(int len) {
  const int nb = (len + BLOCK - 1) / BLOCK;
  return nb < RED_MAX_BLOCKS ? (nb < 1 ? 1 : nb) : RED_MAX_BLOCKS;
}

hipError_t launch_step(int q, int dtype, bool odd, bool periodic, bool wide, const StepLaunch &a, hipStream_t s) {
  return dispatch<StepF>(q, dtype, odd, periodic, wide, a, s);
}
hipError_t launch_step_batch(int q, int dtype, bool odd, bool periodic, bool wide, const BatchLaunch &b,
                             hipStream_t s) {
  return dispatch<BatchF>(q, dtype, odd, periodic, wide, b, s);
}
hipError_t launch_vlinks(int q, int dtype, const VLinkArgs &a, hipStream_t s) {
  return dispatch<VLinkF>(q, dtype, a, s);
}
hipError_t launch_init(int q, int dtype, const InitArgs &a, hipStream_t s) { return dispatch<InitF>(q, dtype, a, s); }
hipError_t launch_peek(int q, int dtype, bool odd, bool periodic, const StepLaunch &a, void *out, hipStream_t s) {
  return dispatch<PeekF>(q, dtype, odd, periodic, a, out, s);
}
hipError_t launch_surface(int q, int dtype, const SurfaceArgs &a, hipStream_t s) {
  return dispatch<SurfaceF>(q, dtype, a, s);
}
hipError_t launch_vis(int q, int dtype, const VisArgs &a, hipStream_t s) { return dispatch<VisF>(q, dtype, a, s); }
hipError_t launch_write_f(int q, int dtype, const WriteFArgs &a, hipStream_t s) {
  return dispatch<WriteFF>(q, dtype, a, s);
}
hipError_t launch_unpermute(int q, int dtype, const UnpermuteArgs &a, hipStream_t s) {
  return dispatch<UnpermuteF>(q, dtype, a, s);
}
template <class T>
void closure_features_w(bool odd, dim3 grid, const StepArgs *args, float *out, hipStream_t s) {
  if (odd)
    hipLaunchKernelGGL((k_closure_features<T, true>), grid, dim3(256), 0, s, args, out);
  else
    hipLaunchKernelGGL((k_closure_features<T, false>), grid, dim3(256), 0, s, args, out);
}

hipError_t launch_closure_features(int dtype, bool odd, const StepArgs *args, long long ncell, float *out,
                                  hipStream_t s) {
  const dim3 grid(static_cast<unsigned>((ncell + 255) / 256));
  switch (dtype) {
    case ASLHIP_F32: closure_features_w<float>(odd, grid, args, out, s); break;
    case ASLHIP_BF16: closure_features_w<unsigned short>(odd, grid, args, out, s); break;
    case ASLHIP_F16: closure_features_w<_Float16>(odd, grid, args, out, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_reduce(int q, int dtype, const ReduceArgs &a, double *out, hipStream_t s) {
  return dispatch<ReduceF>(q, dtype, a, out, s);
}

}  // namespace aslhip
