// lbgk_oracle.cpp — CPU restatement of ASL's LBGK path (test infrastructure).
//
// Only tests/, the smoke check and bench.py's CPU-baseline leg load this
// library.  It restates the reference algorithm; it is not the product.
// Every function cites the ASL source it follows (paths relative to the ASL
// tree).  Built with -ffp-contract=off so that every float/double operation
// rounds exactly once, in the order written here.
//
// Operation order.  ASL's kernel is generated from expression trees; the order
// of every sum and product below follows those trees
// (src/acl/aclMath/aclVectorOfElementsOperations.cxx:336-364: a dot product is
// a left-to-right sum a0*b0 + a1*b1 + ...).  Terms whose constant factor is an
// exact zero (c_i components) are left out: for finite populations x + (+-0)
// == x, so the value is unchanged.  Sums of products x*y + z are contracted
// to one fma (OpenCL's default FP_CONTRACT ON), at the same sites as in the
// GPU kernel; built with -mfma so std::fma is a single instruction.
#include "lbgk_oracle.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Round to the nearest BF16 value (ties to even), kept as a float.
float bf16_round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&x, &u, 4);
  return x;
}

// ---------------------------------------------------------------------------
// Lattice templates — src/math/aslTemplates.cxx.
// ---------------------------------------------------------------------------
struct Templ {
  int Q = 0;
  int c[27][3];
  double w[27];  // quasiparticlesCoefs: lc_i/6, +1 on the rest direction
  int inv[27];
  // laplasCoefs / gradientCoefs (D3Q15 and D3Q19 only: ASL's d3q27() has neither)
  bool has_lc = false;
  double lc[27], gc[27];
};

// VectorTemplate(int, AVec*, double* lc, double* gc): quasiparticlesCoefs[i] =
// lc[i]/6, then [0] += 1 (src/math/aslTemplates.cxx:38-70); invertVectors by
// pairwise search (:72-83).
void finish_templ(Templ &t, const double *lc) {
  for (int i = 0; i < t.Q; ++i) t.w[i] = lc[i] / 6.;
  t.w[0] += 1.;
  for (int i = 0; i < t.Q; ++i)
    for (int j = i + 1; j < t.Q; ++j)
      if (t.c[i][0] == -t.c[j][0] && t.c[i][1] == -t.c[j][1] && t.c[i][2] == -t.c[j][2]) {
        t.inv[i] = j;
        t.inv[j] = i;
      }
  t.inv[0] = 0;
}

Templ make_templ(int q) {
  Templ t;
  // d3q15(): src/math/aslTemplates.cxx:146-165
  static const int c15[15][3] = {{0, 0, 0},  {1, 0, 0},   {0, 1, 0},   {0, 0, 1},   {-1, 0, 0},
                                 {0, -1, 0}, {0, 0, -1},  {1, 1, 1},   {1, 1, -1},  {1, -1, 1},
                                 {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
  // d3q19() / first 19 of d3q27(): src/math/aslTemplates.cxx:168-200
  static const int c27[27][3] = {
      {0, 0, 0},   {1, 0, 0},   {0, 1, 0},  {0, 0, 1},   {-1, 0, 0},  {0, -1, 0},  {0, 0, -1},
      {1, 1, 0},   {1, -1, 0},  {-1, -1, 0}, {-1, 1, 0}, {0, 1, 1},   {0, 1, -1},  {0, -1, -1},
      {0, -1, 1},  {1, 0, 1},   {1, 0, -1}, {-1, 0, -1}, {-1, 0, 1},  {1, 1, 1},   {1, 1, -1},
      {1, -1, 1},  {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
  double lc[27];
  if (q == 15) {
    t.Q = 15;
    std::memcpy(t.c, c15, sizeof(c15));
    lc[0] = -14. / 3.;
    for (int i = 1; i < 7; ++i) lc[i] = 2. / 3.;
    for (int i = 7; i < 15; ++i) lc[i] = 1. / 12.;
    t.gc[0] = 0.;
    for (int i = 1; i < 7; ++i) t.gc[i] = 1. / 3.;
    for (int i = 7; i < 15; ++i) t.gc[i] = 1. / 24.;
    t.has_lc = true;
  } else if (q == 19) {
    t.Q = 19;
    std::memcpy(t.c, c27, sizeof(int) * 3 * 19);
    lc[0] = -4.;
    for (int i = 1; i < 7; ++i) lc[i] = 1. / 3.;
    for (int i = 7; i < 19; ++i) lc[i] = 1. / 6.;
    t.gc[0] = 0.;
    for (int i = 1; i < 7; ++i) t.gc[i] = 1. / 6.;
    for (int i = 7; i < 19; ++i) t.gc[i] = 1. / 12.;
    t.has_lc = true;
  } else if (q == 27) {
    // Extension: ASL builds d3q27() without weights.  Textbook D3Q27 weights
    // 8/27, 2/27, 1/54, 1/216, written as lc = 6 w (rest: 6 (w0 - 1)).
    t.Q = 27;
    std::memcpy(t.c, c27, sizeof(c27));
    lc[0] = 6. * (8. / 27. - 1.);
    for (int i = 1; i < 7; ++i) lc[i] = 6. * (2. / 27.);
    for (int i = 7; i < 19; ++i) lc[i] = 6. * (1. / 54.);
    for (int i = 19; i < 27; ++i) lc[i] = 6. * (1. / 216.);
  } else {
    throw std::invalid_argument("lattice template must be 15, 19 or 27");
  }
  finish_templ(t, lc);
  std::memcpy(t.lc, lc, sizeof(lc));
  if (q == 27) {  // exact textbook values (the lc/6 round trip is not exact)
    t.w[0] = 8. / 27.;
    for (int i = 1; i < 7; ++i) t.w[i] = 2. / 27.;
    for (int i = 7; i < 19; ++i) t.w[i] = 1. / 54.;
    for (int i = 19; i < 27; ++i) t.w[i] = 1. / 216.;
  }
  return t;
}

enum Kind {
  NOSLIP = 1,
  CONSTANT_PRESSURE = 2,
  CONSTANT_VELOCITY = 3,
  CONSTANT_PRESSURE_VELOCITY = 4,
  NOSLIP_MAP = 5,
  PRESSURE_VELOCITY_MAP = 6,
  VELOCITY_MAP = 7
};

template <typename T>
struct BC {
  int kind = 0;
  unsigned faces = 0;
  T p = 0;
  T u[3] = {0, 0, 0};
  bool has_rot = false;
  T omega[3] = {0, 0, 0}, centre[3] = {0, 0, 0};
  std::vector<float> map, cd;
  // slice point list (BCond::addPoint, src/num/aslBCond.cxx:51-56)
  std::vector<long> idx;
  std::vector<int> dir;
};

template <typename T>
struct Lattice {
  Templ vt;
  int N[3], S[3];
  long Tr[3];  // c2iTransformVector of the ghost-inclusive block
  long len;
  bool periodic = false;
  long pool_len = 0;
  std::vector<T> pool;  // Q * pool_len
  std::vector<long> shift, inc;
  std::vector<T> fper, fper2;  // periodic mode: plain arrays Q * len
  std::vector<T> rho, vel;
  T nu = T(0.1);
  bool compressible = true, comp_rho = true, comp_vel = true;
  // LBGK::setOmega (src/num/aslLBGK.cxx:252-255): angular velocity of the
  // rotating frame, constants rounded to T like ASL's literals
  bool has_omega = false;
  T om[3] = {0, 0, 0};
  // Learned closure (an extension; ASL has no counterpart): a 7-32-32-1
  // network of the pre-collision moments gives a per-cell relaxation-time
  // correction.  Parameters are BF16-rounded, evaluated in float.
  bool has_closure = false;
  std::vector<float> cl;          // CL_NPARAM values, layout of closure_eval
  float cl_mu[7], cl_sc[7], cl_out = 0.f;
  int cl_mode = 1;  // 0: FP32 chain; 1: BF16 inputs / hidden activations, exact sums
  // LBGKTurbulence (src/num/aslLBGK.cxx:298-355)
  bool turb = false;
  double dx = 1., pos0[3] = {0., 0., 0.};
  // Extension: velocity-map density averages skip neighbours in the outer
  // ghost layer (ASL reads them as whatever its pool memory holds there).
  bool vmap_skip_outer = false;
  std::vector<BC<T>> bcs;
  // per-direction constants as ASL prints them (double -> T)
  T cw[27];       // w_i
  T cwc[27][3];   // c_i[d] * w_i

  long c2i(int x, int y, int z) const { return x * Tr[0] + y * Tr[1] + z * Tr[2]; }
  long off(int i) const { return c2i(vt.c[i][0], vt.c[i][1], vt.c[i][2]); }
  T *win(int i) { return periodic ? &fper[(size_t)i * len] : &pool[(size_t)i * pool_len + shift[i]]; }

  Lattice(int nx, int ny, int nz, int q, unsigned per) : vt(make_templ(q)) {
    if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("block size must be positive");
    if (per != 0 && per != 7) throw std::invalid_argument("periodic wrap must be on all three axes");
    periodic = per == 7;
    N[0] = nx; N[1] = ny; N[2] = nz;
    for (int d = 0; d < 3; ++d) S[d] = N[d] + 2;  // offset(b, 1)
    // castTransformVector (src/data/aslBlocks.h:118-129): last index fastest
    Tr[2] = 1; Tr[1] = S[2]; Tr[0] = (long)S[1] * S[2];
    len = (long)S[0] * S[1] * S[2];
    for (int i = 0; i < vt.Q; ++i) {
      cw[i] = T(vt.w[i]);
      for (int d = 0; d < 3; ++d) cwc[i][d] = T(double(vt.c[i][d]) * vt.w[i]);
    }
    const T qnan = std::numeric_limits<T>::quiet_NaN();
    if (periodic) {
      fper.assign((size_t)vt.Q * len, qnan);
      fper2.assign((size_t)vt.Q * len, qnan);
    } else {
      // LBGK::createData (src/num/aslLBGK.cxx:129-158)
      shift.resize(vt.Q);
      inc.resize(vt.Q);
      long max_inc = 0;
      for (int i = 0; i < vt.Q; ++i) {
        inc[i] = -off(i);
        if (inc[i] > max_inc) max_inc = inc[i];
      }
      pool_len = 2 * len + 4 * max_inc;
      for (int i = 0; i < vt.Q; ++i) shift[i] = inc[i] > 0 ? 0 : pool_len - len;
      // cl::Buffer memory is never cleared: stale memory is NaN here.
      pool.assign((size_t)vt.Q * pool_len, qnan);
    }
    rho.assign(len, qnan);
    vel.assign((size_t)3 * len, qnan);
  }

  // LBGK::preProcessing + createCopyKernels (src/num/aslLBGK.cxx:160-185,
  // 237-250).  The window of direction i slides by inc_i; when it would leave
  // the pool it is first copied to the pool start (inc > 0) or end.
  void pre_processing() {
    for (int i = 0; i < vt.Q; ++i) {
      if (shift[i] + len + inc[i] > pool_len || shift[i] + inc[i] < 0) {
        long dst = inc[i] > 0 ? 0 : pool_len - len;
        T *base = &pool[(size_t)i * pool_len];
        std::memmove(base + dst, base + shift[i], sizeof(T) * len);  // regions never overlap
        shift[i] = dst;
      }
    }
    for (int i = 0; i < vt.Q; ++i) shift[i] += inc[i];
  }

  // The collision of one cell, LBGK::init0 (src/num/aslLBGK.cxx:187-235).
  // Every x*y + z of ASL's expression tree is contracted to fma(x, y, z), as an
  // OpenCL compiler does under the default FP_CONTRACT ON; the GPU kernel
  // contracts the same sites, so both round identically.
  // Features of one cell: the non-equilibrium second moments
  // Pi_ab = sum_i w_i c_ia c_ib f_i - (rho/3 delta_ab + m_a m_b / rho) in the
  // order xx, yy, zz, xy, xz, yz, then rho - 1.
  inline void closure_features(const T *f, T r, const T *m, T *x) const {
    static const int A[6] = {0, 1, 2, 0, 0, 1}, B[6] = {0, 1, 2, 1, 2, 2};
    T mr[3];
    for (int d = 0; d < 3; ++d) mr[d] = m[d] / r;
    const T third = T(1.) / T(3.);
    for (int p = 0; p < 6; ++p) {
      bool first = true;
      T acc = 0;
      for (int j = 0; j < vt.Q; ++j) {
        const int cc = vt.c[j][A[p]] * vt.c[j][B[p]];
        if (cc == 0) continue;
        const T wcc = T(vt.w[j] * cc);
        acc = first ? f[j] * wcc : std::fma(f[j], wcc, acc);
        first = false;
      }
      const T t = A[p] == B[p] ? std::fma(m[A[p]], mr[B[p]], r * third) : m[A[p]] * mr[B[p]];
      x[p] = acc - t;
    }
    x[6] = r - T(1.);
  }

  // The network: h1 = relu(W1 xn + b1), h2 = relu(W2 h1 + b2), y = w3.h2 + b3
  // with xn = (x - mu) * sc; the relaxation time becomes tau0 + y * out, at
  // least 1/2 + (tau0 - 1/2) / 2.
  // The BF16 evaluation the GPU runs on its matrix cores: inputs and first
  // hidden layer rounded to BF16, products exact, each unit's sum rounded once
  // to float (the hardware may round in its own order: agreement to a few
  // ulp, not bit for bit); the output layer in float in the GPU's order (the
  // 16 units of each wavefront half summed by fma, then the two halves added).
  inline T closure_w_bf16(const T *x, T tau0) const {
    const float *W1 = &cl[0], *b1 = &cl[224], *W2 = &cl[256], *b2 = &cl[1280], *w3 = &cl[1312];
    float xb[7], h1[32], h2[32];
    for (int j = 0; j < 7; ++j) xb[j] = bf16_round((float(x[j]) - cl_mu[j]) * cl_sc[j]);
    for (int o = 0; o < 32; ++o) {
      double a = double(b1[o]);
      for (int j = 0; j < 7; ++j) a += double(W1[o * 7 + j]) * double(xb[j]);
      const float af = float(a);
      h1[o] = bf16_round(af > 0.f ? af : 0.f);
    }
    for (int o = 0; o < 32; ++o) {
      double a = double(b2[o]);
      for (int j = 0; j < 32; ++j) a += double(W2[o * 32 + j]) * double(h1[j]);
      const float af = float(a);
      h2[o] = af > 0.f ? af : 0.f;
    }
    float p[2] = {0.f, 0.f};
    for (int half = 0; half < 2; ++half)
      for (int r = 0; r < 16; ++r) {
        const int u = 8 * (r / 4) + 4 * half + r % 4;
        p[half] = std::fma(w3[u], h2[u], p[half]);
      }
    const float y = (p[0] + p[1]) + cl[1344];
    const float tau = float(tau0) + y * cl_out;
    const float lo = 0.5f + 0.5f * (float(tau0) - 0.5f);
    return T(1.) / T(tau < lo ? lo : tau);
  }

  inline T closure_w(const T *x, T tau0) const {
    if (cl_mode == 1) return closure_w_bf16(x, tau0);
    const float *W1 = &cl[0], *b1 = &cl[224], *W2 = &cl[256], *b2 = &cl[1280], *w3 = &cl[1312];
    float xn[7], h1[32], h2[32];
    for (int j = 0; j < 7; ++j) xn[j] = (float(x[j]) - cl_mu[j]) * cl_sc[j];
    for (int o = 0; o < 32; ++o) {
      float a = b1[o];
      for (int j = 0; j < 7; ++j) a = std::fma(W1[o * 7 + j], xn[j], a);
      h1[o] = a > 0.f ? a : 0.f;
    }
    for (int o = 0; o < 32; ++o) {
      float a = b2[o];
      for (int j = 0; j < 32; ++j) a = std::fma(W2[o * 32 + j], h1[j], a);
      h2[o] = a > 0.f ? a : 0.f;
    }
    float y = cl[1344];
    for (int j = 0; j < 32; ++j) y = std::fma(w3[j], h2[j], y);
    const float tau = float(tau0) + y * cl_out;
    const float lo = 0.5f + 0.5f * (float(tau0) - 0.5f);
    return T(1.) / T(tau < lo ? lo : tau);
  }

  // aDotMomentum = momentum . c_i (a_i_dot_v, src/num/aslLBGK.cxx:65-72):
  // the nonzero components of c_i in order, signed, summed left to right.
  inline T c_dot(int i, const T *m) const {
    bool first = true;
    T a = 0;
    for (int d = 0; d < 3; ++d) {
      if (vt.c[i][d] == 0) continue;
      T t = vt.c[i][d] > 0 ? m[d] : -m[d];
      a = first ? t : a + t;
      first = false;
    }
    return a;
  }

  // fEq (src/num/aslLBGK.cxx:333-345) of direction i; mmr = 1.5 m.m (/ rho).
  inline T feq_dir(int i, T r, T a, T mmr) const {
    if (i == 0) return r - mmr;  // rho + 3*0 + 4.5*0/rho - mm/rho
    if (compressible) return (std::fma(T(3.), a, r) + (T(4.5) * (a * a)) / r) - mmr;
    return std::fma(T(4.5), a * a, std::fma(T(3.), a, r)) - mmr;
  }

  inline T mm15(const T *m) const {
    T mm = (T(1.5) * m[0]) * m[0];
    mm = std::fma(T(1.5) * m[1], m[1], mm);
    return std::fma(T(1.5) * m[2], m[2], mm);
  }

  // LBGKTurbulence::init0 (src/num/aslLBGK.cxx:312-318):
  //   w = 4 (fOld - fEq).(fOld - fEq) / fOld.fOld,
  //   w = w > 1 ? 1 / (3 min(nu w, .5) + .5) : 1 / (3 nu + .5).
  // ASL evaluates it before fEq is assigned (an uninitialised private
  // variable of the generated kernel); fEq here is the equilibrium of the
  // cell's own rho and momentum, before the rotating-frame term.
  inline T turbulence_w(const T *f, T r, const T *m, T w_plain) const {
    const int Q = vt.Q;
    const T mm = mm15(m);
    const T mmr = compressible ? mm / r : mm;
    T fl2 = f[0] * f[0];
    T d0 = f[0] - feq_dir(0, r, T(0), mmr);
    T dl2 = d0 * d0;
    for (int i = 1; i < Q; ++i) {
      fl2 = std::fma(f[i], f[i], fl2);
      const T di = f[i] - feq_dir(i, r, c_dot(i, m), mmr);
      dl2 = std::fma(di, di, dl2);
    }
    const T ratio = T(4.) * (dl2 / fl2);
    const T lim = std::min(nu * ratio, T(.5));
    return ratio > T(1.) ? T(1.) / (T(3.) * lim + T(.5)) : w_plain;
  }

  inline void collide(T *f, T w, T &rho_o, T *v_o) const {
    const int Q = vt.Q;
    // momentum = computeMomentum(fOld) (:51-63), pRho = computeRho(fOld) (:46-49)
    T m[3];
    for (int d = 0; d < 3; ++d) {
      bool first = true;
      T acc = 0;
      for (int j = 0; j < Q; ++j) {
        if (vt.c[j][d] == 0) continue;
        acc = first ? f[j] * cwc[j][d] : std::fma(f[j], cwc[j][d], acc);
        first = false;
      }
      m[d] = acc;
    }
    T r = f[0] * cw[0];
    for (int j = 1; j < Q; ++j) r = std::fma(f[j], cw[j], r);
    if (turb) {
      w = turbulence_w(f, r, m, w);
    } else if (has_closure) {
      T x[7];
      closure_features(f, r, m, x);
      w = closure_w(x, T(3.) * nu + T(.5));
    }
    // momentum += (2./w) * crossProduct(omega, momentum) (:207-208).  The
    // three components are three statements of the generated kernel, in
    // order (operator+=, src/acl/aclMath/aclVectorOfElementsOperations.cxx:249-259;
    // ExpressionContainer <<, src/acl/acl.cxx:578-584), so each one sees the
    // components updated before it.  crossProduct: :514-535.
    if (has_omega) {
      const T t = T(2.) / w;
      m[0] = std::fma(t, std::fma(om[1], m[2], -(om[2] * m[1])), m[0]);
      m[1] = std::fma(t, std::fma(om[2], m[0], -(om[0] * m[2])), m[1]);
      m[2] = std::fma(t, std::fma(om[0], m[1], -(om[1] * m[0])), m[2]);
    }
    rho_o = r;
    if (compressible)
      for (int d = 0; d < 3; ++d) v_o[d] = m[d] / r;
    else
      for (int d = 0; d < 3; ++d) v_o[d] = m[d];
    // 1.5 * momentum * momentum (dot product), then / pRho (compressible)
    const T mm = mm15(m);
    const T one_w = T(1.) - w;
    const T mmr = compressible ? mm / r : mm;
    for (int i = 0; i < Q; ++i) {
      const T feq = feq_dir(i, r, i ? c_dot(i, m) : T(0), mmr);
      f[i] = std::fma(f[i], one_w, w * feq);  // fOld*(1-w) + w*fEq
    }
  }

  void collide_all() {
    const int Q = vt.Q;
    const T w = T(1.) / (T(3.) * nu + T(.5));  // :203
    if (!periodic) {
      T *wp[27];
      for (int i = 0; i < Q; ++i) wp[i] = win(i);
#pragma omp parallel for schedule(static)
      for (long j = 0; j < len; ++j) {
        T f[27], r, v[3];
        for (int i = 0; i < Q; ++i) f[i] = wp[i][j];
        collide(f, w, r, v);
        if (comp_rho) rho[j] = r;
        if (comp_vel)
          for (int d = 0; d < 3; ++d) vel[(size_t)d * len + j] = v[d];
        for (int i = 0; i < Q; ++i) wp[i][j] = f[i];
      }
    } else {
      // Extension: periodic pull streaming over the interior cells.
#pragma omp parallel for collapse(2) schedule(static)
      for (int x = 1; x <= N[0]; ++x)
        for (int y = 1; y <= N[1]; ++y)
          for (int z = 1; z <= N[2]; ++z) {
            long j = c2i(x, y, z);
            T f[27], r, v[3];
            for (int i = 0; i < Q; ++i) {
              int xs = x - vt.c[i][0], ys = y - vt.c[i][1], zs = z - vt.c[i][2];
              xs = xs < 1 ? N[0] : (xs > N[0] ? 1 : xs);
              ys = ys < 1 ? N[1] : (ys > N[1] ? 1 : ys);
              zs = zs < 1 ? N[2] : (zs > N[2] ? 1 : zs);
              f[i] = fper[(size_t)i * len + c2i(xs, ys, zs)];
            }
            collide(f, w, r, v);
            if (comp_rho) rho[j] = r;
            if (comp_vel)
              for (int d = 0; d < 3; ++d) vel[(size_t)d * len + j] = v[d];
            for (int i = 0; i < Q; ++i) fper2[(size_t)i * len + j] = f[i];
          }
      fper.swap(fper2);
    }
  }

  // SingleKernelNM::execute = preProcessing + compute (src/num/aslSingleKernelNM.cxx:48-53)
  void lbgk_execute() {
    if (!periodic) pre_processing();
    collide_all();
  }

  // LBGKUtilities::initF(rho, vel) (src/num/aslLBGK.cxx:262-275):
  // f_i = rho * (1 + 3 a_i + 4.5 a_i^2 - 1.5 u.u), a_i = u . c_i
  void init_cell(long j, T r, const T *u) {
    T uu = (T(1.5) * u[0]) * u[0];
    uu = std::fma(T(1.5) * u[1], u[1], uu);
    uu = std::fma(T(1.5) * u[2], u[2], uu);
    for (int i = 0; i < vt.Q; ++i) {
      T a = i == 0 ? T(0) : dot_c(u, i);
      T fi = std::fma(T(4.5), a * a, std::fma(T(3.), a, T(1.))) - uu;
      win(i)[j] = r * fi;
    }
    rho[j] = r;
    for (int d = 0; d < 3; ++d) vel[(size_t)d * len + j] = u[d];
  }

  // v . c_i with the zero terms left out (u0*c0 + u1*c1 + u2*c2)
  T dot_c(const T *u, int i) const {
    bool first = true;
    T a = 0;
    for (int d = 0; d < 3; ++d) {
      if (vt.c[i][d] == 0) continue;
      T t = vt.c[i][d] > 0 ? u[d] : -u[d];
      a = first ? t : a + t;
      first = false;
    }
    return a;
  }

  // ---------------------------------------------------------------------
  // Boundary conditions
  // ---------------------------------------------------------------------
  // addSliceX0 ... addSliceZE (src/num/aslBCond.cxx:373-493): for the face
  // with inward direction e = vectors[1..6] and each k with c_k . e > 0, every
  // face-layer node p (tangential indices 1..S-2) gives the entry
  // (p - c_k, k).
  void add_slices(BC<T> &bc) {
    const int Q = vt.Q;
    static const int face_dir[6] = {1, 4, 2, 5, 3, 6};  // X0 XE Y0 YE Z0 ZE
    for (int f = 0; f < 6; ++f) {
      if (!(bc.faces & (1u << f))) continue;
      const int *e = vt.c[face_dir[f]];
      int axis = f / 2;
      int layer = (f % 2 == 0) ? 1 : S[axis] - 2;
      int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
      for (int k = 1; k < Q; ++k) {
        int dotp = vt.c[k][0] * e[0] + vt.c[k][1] * e[1] + vt.c[k][2] * e[2];
        if (dotp <= 0) continue;
        for (int i = 1; i < S[a1] - 1; ++i)
          for (int j = 1; j < S[a2] - 1; ++j) {
            int p[3];
            p[axis] = layer;
            p[a1] = i;
            p[a2] = j;
            bc.idx.push_back(c2i(p[0] - vt.c[k][0], p[1] - vt.c[k][1], p[2] - vt.c[k][2]));
            bc.dir.push_back(k);
          }
      }
    }
  }

  // Maps are float fields (the drivers' FlT); their compares are made in float.
  float map_safe(const std::vector<float> &m, long idx, float out) const {
    // generateVEOutOfBoundarySafe(a, outVal) (src/acl/aclMath/aclMathAlg.cxx:84-94)
    return (idx < 0 || idx > len - 1) ? out : m[(size_t)idx];
  }
  T f_safe(int i, long idx) {
    // generateVEOutOfBoundarySafe(a) (src/acl/aclMath/aclMathAlg.cxx:76-81)
    return win(i)[(len + idx) % len];
  }
  // ComputeSurfaceForceMap / ComputeSurfaceFluxMap::init + execute
  // (src/num/aslLBGKBC.cxx:623-736) on the pool state after the last step and
  // its BCs: per cell with |map| < 0.9999, over i = 1..Q-1,
  //   isBoundary0 = map(x) <= 0 && map(x + c_i) > 0, isBoundaryI = map(x) > 0 && map(x + c_i) <= 0,
  //   f0 = f_i(x), fI = f_inv(i)(x + c_i) (linear index wrapped, as generateDCFullSafe),
  //   force += select(0, (f0 + fI - 2) w_i c_i, isBoundary0) + select(0, (-f0 - fI + 2) w_i c_i, isBoundaryI),
  //   flux += select((f0 - fI) w_i, isBoundary0) + select((fI - f0) w_i, isBoundaryI).
  // Cells with |map| >= 0.9999 are left to the caller (0 here).
  void surface(const float *map, bool flux, double *out) {
    const int Q = vt.Q;
    std::vector<float> m(map, map + len);
#pragma omp parallel for schedule(static)
    for (long j = 0; j < len; ++j) {
      T acc[3] = {T(0), T(0), T(0)};
      const float m0 = m[(size_t)j];
      if (std::fabs(m0) < .9999f) {
        for (int i = 1; i < Q; ++i) {
          const long y = j + off(i);
          const float my = map_safe(m, y, -.9f);
          const bool b0 = m0 <= 0.f && my > 0.f;
          const bool bi = m0 > 0.f && my <= 0.f;
          const T f0 = win(i)[j];
          const T fI = f_safe(vt.inv[i], y);
          if (flux) {
            const T wi = T(vt.w[i]);
            acc[0] = acc[0] + (b0 ? (f0 - fI) * wi : T(0));
            acc[0] = acc[0] + (bi ? (fI - f0) * wi : T(0));
          } else {
            for (int d = 0; d < 3; ++d) {
              const T wc = T(vt.w[i] * double(vt.c[i][d]));
              acc[d] = acc[d] + (b0 ? ((f0 + fI) - T(2.)) * wc : T(0));
              acc[d] = acc[d] + (bi ? ((-f0 - fI) + T(2.)) * wc : T(0));
            }
          }
        }
      }
      if (flux)
        out[j] = double(acc[0]);
      else
        for (int d = 0; d < 3; ++d) out[(size_t)d * len + j] = double(acc[d]);
    }
  }

  // idx in the block's outer ghost layer (or outside the block)
  bool outer_layer(long idx) const {
    if (idx < 0 || idx >= len) return true;
    const long x = idx / Tr[0], y = (idx % Tr[0]) / Tr[1], z = (idx % Tr[0]) % Tr[1];
    return x == 0 || y == 0 || z == 0 || x == S[0] - 1 || y == S[1] - 1 || z == S[2] - 1;
  }

  bool comp_node(const BC<T> &bc, long idx) const {
    // isComputationNode (src/num/aslBCond.cxx:204-211;
    // src/math/aslDistanceFunctionAlg.cxx:45-48): map > 0 [&& cdomain > 0]
    bool r = map_safe(bc.map, idx, -.9f) > 0.f;
    if (!bc.cd.empty()) r = r && map_safe(bc.cd, idx, -1.f) > 0.f;
    return r;
  }

  // Visualisation fills of rho / u at the boundary nodes of a map, in place,
  // on cells with map <= 0 && map > -0.9999 (the kernels' outer condition).
  // Reads come only from cells with map > 0 and writes go only to cells with
  // map <= 0, so the in-place sweep equals ASL's one-kernel update.
  //  kind 1, generateBCNoSlipVel (src/num/aslLBGKBC.cxx:808-815) ->
  //    BCConstantValueMiddlePointMap(velocity, 0, map) (src/num/aslBasicBC.cxx:
  //    167-227): over i = 1..Q-1 with b_i = map(x + c_i) > 0,
  //      v += select(2*0 - v(x + c_i), b_i) * lc_i,  n += select(lc_i, b_i),
  //      v(x) = v / max(n, .01);
  //  kind 2, generateBCNoSlipRho (:818-825) -> BCConstantGradientMap(rho, 0,
  //    map) (src/num/aslBasicBC.cxx:283-351): the map normal
  //      g_k = sum_i map(x + c_i) gc_i c_i[k]   (gradient, src/math/aslTemplateVE.cxx:210-222)
  //    normalised by its length (gcNormalize, aclVectorOfElementsOperations.cxx:
  //    1213-1244), d_i = (g . c_i/|c_i| + 1)^2 and
  //      n = max(sum_i select(d_i, b_i), .001),
  //      r = sum_i select((rho(x + c_i) - (g . c_i/|c_i|) * 0) d_i, b_i),
  //      rho(x) = r / n where n > 0.1 (else unchanged);
  //  kind 3, generateBCVelocityVel (:845-850) -> BCValuePFMap(velocity, v, map)
  //    (src/num/aslBasicBC.cxx:124-145): v(x) = the position function at the
  //    node's physical position (Index2PositionACL of the ghost-inclusive block).
  // Map reads outside the block are -0.9 (generateDCFullSafe(map, -.9)); field
  // reads wrap the linear index (generateDCFullSafe(data)).  Sums run left to
  // right in the order of the directions, dot products over the non-zero
  // components; no operation is contracted.
  void vis_fill(int kind, const float *mapp, const T *u, const T *rot, bool has_rot) {
    const int Q = vt.Q;
    if ((kind == 1 || kind == 2) && !vt.has_lc) throw std::invalid_argument("d3q27 has no laplace/gradient coefficients");
    std::vector<float> m(mapp, mapp + len);
    auto mx = [&](long idx) { return map_safe(m, idx, -.9f); };
    auto wrap = [&](long idx) { return (size_t)((len + idx) % len); };
    T chat[27][3];
    for (int i = 0; i < Q; ++i) {
      const double l = std::sqrt(double(vt.c[i][0] * vt.c[i][0] + vt.c[i][1] * vt.c[i][1] + vt.c[i][2] * vt.c[i][2]));
      for (int d = 0; d < 3; ++d) chat[i][d] = i ? T(double(vt.c[i][d]) / l) : T(0);
    }
#pragma omp parallel for schedule(static)
    for (long g = 0; g < len; ++g) {
      const float mg = m[(size_t)g];
      if (!(mg <= 0.f && mg > -.9999f)) continue;
      if (kind == 1) {
        T v[3] = {T(0), T(0), T(0)}, n = T(0);
        for (int i = 1; i < Q; ++i) {
          const bool b = mx(g + off(i)) > 0.f;
          const size_t y = wrap(g + off(i));
          for (int c = 0; c < 3; ++c) v[c] = v[c] + (b ? T(0) - vel[(size_t)c * len + y] : T(0)) * T(vt.lc[i]);
          n = n + (b ? T(vt.lc[i]) : T(0));
        }
        const T den = n > T(.01) ? n : T(.01);
        for (int c = 0; c < 3; ++c) vel[(size_t)c * len + g] = v[c] / den;
      } else if (kind == 2) {
        T gn[3];
        for (int k = 0; k < 3; ++k) {
          T acc = T(0);
          bool first = true;
          for (int i = 1; i < Q; ++i) {
            if (vt.c[i][k] == 0) continue;
            const T t = T(mx(g + off(i))) * T(vt.gc[i] * double(vt.c[i][k]));
            acc = first ? t : acc + t;
            first = false;
          }
          gn[k] = acc;
        }
        T l2 = gn[0] * gn[0];
        l2 = l2 + gn[1] * gn[1];
        l2 = l2 + gn[2] * gn[2];
        const T ln = std::sqrt(l2);
        for (int k = 0; k < 3; ++k) gn[k] = gn[k] / ln;
        T pr[27], dc[27];
        T n = T(0);
        for (int i = 1; i < Q; ++i) {
          T a = T(0);
          bool first = true;
          for (int d = 0; d < 3; ++d) {
            if (vt.c[i][d] == 0) continue;
            const T t = gn[d] * chat[i][d];
            a = first ? t : a + t;
            first = false;
          }
          pr[i] = a;
          const T e = a + T(1.);
          dc[i] = e * e;
          n = n + (mx(g + off(i)) > 0.f ? dc[i] : T(0));
        }
        n = n > T(.001) ? n : T(.001);  // a NaN count stays below the 0.1 test either way
        T r = T(0);
        for (int i = 1; i < Q; ++i) {
          const bool b = mx(g + off(i)) > 0.f;
          r = r + (b ? (rho[wrap(g + off(i))] - pr[i] * T(0)) * dc[i] : T(0));
        }
        if (n > T(.1)) rho[(size_t)g] = r / n;
      } else {
        const long I[3] = {g / Tr[0], (g % Tr[0]) / Tr[1], (g % Tr[0]) % Tr[1]};
        T uw[3] = {u[0], u[1], u[2]};
        if (has_rot) {
          T pos[3];
          for (int d = 0; d < 3; ++d) pos[d] = T(pos0[d] - dx) + T(dx) * T(I[d]);
          const T d0 = pos[0] - rot[3], d1 = pos[1] - rot[4], d2 = pos[2] - rot[5];
          uw[0] = rot[1] * d2 - rot[2] * d1;
          uw[1] = rot[2] * d0 - rot[0] * d2;
          uw[2] = rot[0] * d1 - rot[1] * d0;
        }
        for (int c = 0; c < 3; ++c) vel[(size_t)c * len + g] = uw[c];
      }
    }
  }

  void execute_bc(BC<T> &bc) {
    const int Q = vt.Q;
    switch (bc.kind) {
      case NOSLIP:  // BCNoSlip::init (src/num/aslLBGKBC.cxx:91-113)
        for (size_t e = 0; e < bc.idx.size(); ++e) {
          int k = bc.dir[e];
          long g = bc.idx[e];
          win(k)[g] = win(vt.inv[k])[g + off(k)];
        }
        break;
      case CONSTANT_PRESSURE:  // BCConstantPressure::init (:121-142)
        for (size_t e = 0; e < bc.idx.size(); ++e) win(bc.dir[e])[bc.idx[e]] = bc.p;
        break;
      case CONSTANT_VELOCITY:  // BCConstantVelocity::init (:149-172)
        for (size_t e = 0; e < bc.idx.size(); ++e) {
          int k = bc.dir[e];
          long g = bc.idx[e];
          T add = T(6.) * dot_c(bc.u, k);
          win(k)[g] = win(vt.inv[k])[g + off(k)] + add;
        }
        break;
      case CONSTANT_PRESSURE_VELOCITY:  // BCConstantPressureVelocity::init (:182-202)
        for (size_t e = 0; e < bc.idx.size(); ++e) {
          int k = bc.dir[e];
          win(k)[bc.idx[e]] = bc.p + T(3.) * dot_c(bc.u, k);
        }
        break;
      case NOSLIP_MAP:  // BCNoSlipMap::init (:219-254)
#pragma omp parallel for schedule(static)
        for (long j = 0; j < len; ++j) {
          const float mj = bc.map[(size_t)j];
          if (!(mj <= 0.f && mj > -.9999f)) continue;
          for (int i = 1; i < Q; ++i)
            if (comp_node(bc, j + off(i))) win(i)[j] = f_safe(vt.inv[i], j + off(i));
        }
        break;
      case PRESSURE_VELOCITY_MAP: {  // BCConstantPressureVelocityMap::init (:386-432)
        T feq[27];
        const T *u = bc.u;
        const T p = bc.p;
        T uu = u[0] * u[0];
        uu = uu + u[1] * u[1];
        uu = uu + u[2] * u[2];
        for (int i = 1; i < Q; ++i) {
          T a = dot_c(u, i);
          if (compressible)
            feq[i] = ((p + (T(3.) * p) * a) + ((T(4.5) * p) * a) * a) - (T(1.5) * p) * uu;
          else {
            // 3.*velocity*c_i is the dot product of (3u) with c_i
            T u3[3] = {T(3.) * u[0], T(3.) * u[1], T(3.) * u[2]};
            feq[i] = ((p + dot_c(u3, i)) + (T(4.5) * a) * a) - T(1.5) * uu;
          }
        }
#pragma omp parallel for schedule(static)
        for (long j = 0; j < len; ++j) {
          const float mj = bc.map[(size_t)j];
          if (!(mj <= 0.f && mj > -.9999f)) continue;
          for (int i = 1; i < Q; ++i)
            if (comp_node(bc, j + off(i))) win(i)[j] = feq[i];
        }
        break;
      }
      case VELOCITY_MAP: {  // BCVelocityMap::init (:290-348)
#pragma omp parallel for schedule(static)
        for (long j = 0; j < len; ++j) {
          const float mj = bc.map[(size_t)j];
          if (!(mj <= 0.f && mj > -.9999f)) continue;
          T coef = T(0), r = T(0);
          for (int i = 1; i < Q; ++i) {
            if (!comp_node(bc, j + off(i))) continue;
            if (vmap_skip_outer && outer_layer(j + off(i))) continue;
            coef = coef + cw[i];
            r = r + f_safe(vt.inv[i], j + off(i)) * cw[i];
          }
          r = r / coef;
          // Index2PositionACL over the ghost-inclusive block
          // (src/math/aslIndex2Position.cxx:64-108)
          long I[3] = {j / Tr[0], (j % Tr[0]) / Tr[1], (j % Tr[0]) % Tr[1]};
          T pos[3];
          for (int d = 0; d < 3; ++d) pos[d] = T(pos0[d] - dx) + T(dx) * T(I[d]);
          T uw[3] = {bc.u[0], bc.u[1], bc.u[2]};
          if (bc.has_rot) {
            // PFRotationField: crossProduct(axis, pos - c)
            // (src/math/aslPositionFunction.cxx:180-190; cross product
            // src/acl/aclMath/aclVectorOfElementsOperations.cxx:514-535)
            T d0 = pos[0] - bc.centre[0], d1 = pos[1] - bc.centre[1], d2 = pos[2] - bc.centre[2];
            uw[0] = bc.omega[1] * d2 - bc.omega[2] * d1;
            uw[1] = bc.omega[2] * d0 - bc.omega[0] * d2;
            uw[2] = bc.omega[0] * d1 - bc.omega[1] * d0;
          }
          T mom[3] = {uw[0] * r, uw[1] * r, uw[2] * r};
          for (int i = 1; i < Q; ++i)
            if (comp_node(bc, j + off(i))) win(i)[j] = r + dot_c(mom, i) * T(3.);
        }
        break;
      }
      default:
        throw std::invalid_argument("unknown boundary-condition kind");
    }
  }
};

struct Handle {
  bool fp64;
  std::unique_ptr<Lattice<float>> f;
  std::unique_ptr<Lattice<double>> d;
};

template <typename F>
auto dispatch(void *h, F fn) {
  Handle *H = static_cast<Handle *>(h);
  if (H->fp64) return fn(*H->d);
  return fn(*H->f);
}

template <typename T>
int add_slice(Lattice<T> &L, int kind, unsigned faces, double p, const double *u) {
  if (L.periodic) throw std::invalid_argument("slice BCs need a non-periodic lattice");
  if (kind < NOSLIP || kind > CONSTANT_PRESSURE_VELOCITY) throw std::invalid_argument("not a slice BC kind");
  BC<T> bc;
  bc.kind = kind;
  bc.faces = faces;
  bc.p = T(p);
  for (int d = 0; d < 3; ++d) bc.u[d] = u ? T(u[d]) : T(0);
  L.add_slices(bc);
  L.bcs.push_back(std::move(bc));
  return (int)L.bcs.size() - 1;
}

template <typename T>
int add_map(Lattice<T> &L, int kind, const float *map, const float *cd, double p, const double *u,
            const double *rot) {
  if (L.periodic) throw std::invalid_argument("map BCs need a non-periodic lattice");
  if (kind < NOSLIP_MAP || kind > VELOCITY_MAP) throw std::invalid_argument("not a map BC kind");
  BC<T> bc;
  bc.kind = kind;
  bc.p = T(p);
  for (int d = 0; d < 3; ++d) bc.u[d] = u ? T(u[d]) : T(0);
  bc.map.assign(map, map + L.len);
  if (cd) bc.cd.assign(cd, cd + L.len);
  if (rot) {
    bc.has_rot = true;
    for (int d = 0; d < 3; ++d) {
      bc.omega[d] = T(rot[d]);
      bc.centre[d] = T(rot[3 + d]);
    }
  }
  L.bcs.push_back(std::move(bc));
  return (int)L.bcs.size() - 1;
}

}  // namespace

extern "C" {

void *orc_create(int nx, int ny, int nz, int q, int fp64, unsigned periodic) {
  try {
    Handle *h = new Handle();
    h->fp64 = fp64 != 0;
    if (h->fp64)
      h->d.reset(new Lattice<double>(nx, ny, nz, q, periodic));
    else
      h->f.reset(new Lattice<float>(nx, ny, nz, q, periodic));
    return h;
  } catch (...) {
    return nullptr;
  }
}

void orc_destroy(void *h) { delete static_cast<Handle *>(h); }

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

void orc_set_viscosity(void *h, double nu) {
  dispatch(h, [&](auto &L) { L.nu = decltype(L.nu)(nu); });
}
void orc_set_compressible(void *h, int c) {
  dispatch(h, [&](auto &L) { L.compressible = c != 0; });
}
void orc_set_omega(void *h, const double *om) {
  dispatch(h, [&](auto &L) {
    using T = std::remove_reference_t<decltype(L.om[0])>;
    L.has_omega = om != nullptr;
    for (int d = 0; d < 3; ++d) L.om[d] = om ? T(om[d]) : T(0);
  });
}
int orc_set_closure(void *h, const float *params, const float *mu, const float *sc, float out) {
  int rc = 0;
  dispatch(h, [&](auto &L) {
    if (sizeof(L.om[0]) != sizeof(float) || L.vt.Q != 19) {
      rc = -1;
      return;
    }
    L.has_closure = params != nullptr;
    if (!params) return;
    L.cl.resize(1345);
    for (int i = 0; i < 1345; ++i) L.cl[i] = bf16_round(params[i]);
    for (int j = 0; j < 7; ++j) {
      L.cl_mu[j] = mu[j];
      L.cl_sc[j] = sc[j];
    }
    L.cl_out = out;
  });
  return rc;
}

void orc_surface(void *h, const float *map, int flux, double *out) {
  dispatch(h, [&](auto &L) { L.surface(map, flux != 0, out); });
}

// Visualisation fills (Lattice::vis_fill): kind 1 no-slip velocity, 2 no-slip
// density, 3 velocity from a position function (u, or rot = {axis, centre}).
int orc_vis_fill(void *h, int kind, const float *map, const double u[3], const double rot[6]) {
  try {
    dispatch(h, [&](auto &L) {
      using T = std::remove_reference_t<decltype(L.rho[0])>;
      T uu[3], rr[6];
      for (int d = 0; d < 3; ++d) uu[d] = u ? T(u[d]) : T(0);
      for (int d = 0; d < 6; ++d) rr[d] = rot ? T(rot[d]) : T(0);
      L.vis_fill(kind, map, uu, rr, rot != nullptr);
      return 0;
    });
  } catch (...) {
    return -1;
  }
  return 0;
}

void orc_set_turbulence(void *h, int on) {
  dispatch(h, [&](auto &L) { L.turb = on != 0; });
}

void orc_set_closure_mode(void *h, int mode) {
  dispatch(h, [&](auto &L) { L.cl_mode = mode; });
}

// Closure features of every interior cell (periodic lattices): the inputs the
// next collision would see, 7 per cell, cells in ASL order.
int orc_closure_features(void *h, float *out) {
  int rc = 0;
  dispatch(h, [&](auto &L) {
    using T = std::remove_reference_t<decltype(L.om[0])>;
    if (!L.periodic || L.vt.Q != 19) {
      rc = -1;
      return;
    }
    const int Q = L.vt.Q;
#pragma omp parallel for collapse(2) schedule(static)
    for (int x = 1; x <= L.N[0]; ++x)
      for (int y = 1; y <= L.N[1]; ++y)
        for (int z = 1; z <= L.N[2]; ++z) {
          T f[27];
          for (int i = 0; i < Q; ++i) {
            int xs = x - L.vt.c[i][0], ys = y - L.vt.c[i][1], zs = z - L.vt.c[i][2];
            xs = xs < 1 ? L.N[0] : (xs > L.N[0] ? 1 : xs);
            ys = ys < 1 ? L.N[1] : (ys > L.N[1] ? 1 : ys);
            zs = zs < 1 ? L.N[2] : (zs > L.N[2] ? 1 : zs);
            f[i] = L.fper[(size_t)i * L.len + L.c2i(xs, ys, zs)];
          }
          T m[3];
          for (int d = 0; d < 3; ++d) {
            bool first = true;
            T acc = 0;
            for (int j = 0; j < Q; ++j) {
              if (L.vt.c[j][d] == 0) continue;
              acc = first ? f[j] * L.cwc[j][d] : std::fma(f[j], L.cwc[j][d], acc);
              first = false;
            }
            m[d] = acc;
          }
          T r = f[0] * L.cw[0];
          for (int j = 1; j < Q; ++j) r = std::fma(f[j], L.cw[j], r);
          T xf[7];
          L.closure_features(f, r, m, xf);
          const size_t c = ((size_t)(x - 1) * L.N[1] + (y - 1)) * L.N[2] + (z - 1);
          for (int p = 0; p < 7; ++p) out[7 * c + p] = float(xf[p]);
        }
  });
  return rc;
}

void orc_set_compute_macro(void *h, int r, int v) {
  dispatch(h, [&](auto &L) {
    L.comp_rho = r != 0;
    L.comp_vel = v != 0;
  });
}
void orc_set_vmap_skip_outer(void *h, int on) {
  dispatch(h, [&](auto &L) { L.vmap_skip_outer = on != 0; });
}
void orc_set_geometry(void *h, double dx, const double pos[3]) {
  dispatch(h, [&](auto &L) {
    L.dx = dx;
    for (int d = 0; d < 3; ++d) L.pos0[d] = pos[d];
  });
}

void orc_init_equilibrium(void *h, double rho, const double u[3]) {
  dispatch(h, [&](auto &L) {
    using T = decltype(L.nu);
    T uu[3] = {T(u[0]), T(u[1]), T(u[2])};
    for (long j = 0; j < L.len; ++j) L.init_cell(j, T(rho), uu);
  });
}

void orc_init_equilibrium_field(void *h, const double *rho, const double *u) {
  dispatch(h, [&](auto &L) {
    using T = decltype(L.nu);
    for (long j = 0; j < L.len; ++j) {
      T uu[3] = {T(u[j]), T(u[L.len + j]), T(u[2 * L.len + j])};
      L.init_cell(j, rho ? T(rho[j]) : T(1), uu);
    }
  });
}

int orc_bc_add_slice(void *h, int kind, unsigned faces, double p, const double u[3]) {
  try {
    return dispatch(h, [&](auto &L) { return add_slice(L, kind, faces, p, u); });
  } catch (...) {
    return -1;
  }
}

int orc_bc_add_map(void *h, int kind, const float *map, const float *cd, double p, const double u[3],
                   const double rot[6]) {
  try {
    return dispatch(h, [&](auto &L) { return add_map(L, kind, map, cd, p, u, rot); });
  } catch (...) {
    return -1;
  }
}

int orc_bc_execute(void *h, int id) {
  return dispatch(h, [&](auto &L) {
    if (id < 0 || id >= (int)L.bcs.size()) return -1;
    L.execute_bc(L.bcs[id]);
    return 0;
  });
}

void orc_lbgk_execute(void *h) {
  dispatch(h, [&](auto &L) { L.lbgk_execute(); });
}

void orc_run(void *h, int nsteps, const int *order, int norder) {
  dispatch(h, [&](auto &L) {
    for (int s = 0; s < nsteps; ++s) {
      L.lbgk_execute();
      for (int b = 0; b < norder; ++b) L.execute_bc(L.bcs[order[b]]);
    }
  });
}

void orc_read_rho(void *h, double *out) {
  dispatch(h, [&](auto &L) {
    for (long j = 0; j < L.len; ++j) out[j] = double(L.rho[j]);
  });
}
void orc_read_velocity(void *h, double *out) {
  dispatch(h, [&](auto &L) {
    for (long j = 0; j < 3 * L.len; ++j) out[j] = double(L.vel[j]);
  });
}
void orc_read_f(void *h, double *out) {
  dispatch(h, [&](auto &L) {
    for (int i = 0; i < L.vt.Q; ++i) {
      auto *w = L.win(i);
      for (long j = 0; j < L.len; ++j) out[(size_t)i * L.len + j] = double(w[j]);
    }
  });
}
void orc_write_f(void *h, const double *in) {
  dispatch(h, [&](auto &L) {
    using T = decltype(L.nu);
    for (int i = 0; i < L.vt.Q; ++i) {
      auto *w = L.win(i);
      for (long j = 0; j < L.len; ++j) w[j] = T(in[(size_t)i * L.len + j]);
    }
  });
}

long orc_slice_entries(void *h, int id) {
  return dispatch(h, [&](auto &L) -> long {
    if (id < 0 || id >= (int)L.bcs.size()) return -1;
    return (long)L.bcs[id].idx.size();
  });
}

// bit0: ghost node (map <= 0); bit1: the map guard (map <= 0 && map > -0.9999);
// bit2: computation node (map > 0 [&& cdomain > 0]).
void orc_map_masks(void *h, int id, uint8_t *out) {
  dispatch(h, [&](auto &L) {
    auto &bc = L.bcs[id];
    for (long j = 0; j < L.len; ++j) {
      const float m = bc.map[(size_t)j];
      uint8_t v = 0;
      if (m <= 0.f) v |= 1;
      if (m <= 0.f && m > -.9999f) v |= 2;
      if (L.comp_node(bc, j)) v |= 4;
      out[j] = v;
    }
  });
}

// FP64 sums over cells with mask != 0 of the current populations:
// {sum rho, sum m_x, sum m_y, sum m_z, sum |m|^2/(2 rho)}.
void orc_conservation(void *h, const uint8_t *mask, double out[5]) {
  dispatch(h, [&](auto &L) {
    double s[5] = {0, 0, 0, 0, 0};
    for (long j = 0; j < L.len; ++j) {
      if (!mask[j]) continue;
      double r = 0, m[3] = {0, 0, 0};
      for (int i = 0; i < L.vt.Q; ++i) {
        double fi = double(L.win(i)[j]) * L.vt.w[i];
        r += fi;
        for (int d = 0; d < 3; ++d) m[d] += fi * L.vt.c[i][d];
      }
      s[0] += r;
      for (int d = 0; d < 3; ++d) s[1 + d] += m[d];
      s[4] += (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) / (2. * r);
    }
    for (int k = 0; k < 5; ++k) out[k] = s[k];
  });
}

int orc_template(int q, int *c, double *w, int *inv) {
  try {
    Templ t = make_templ(q);
    for (int i = 0; i < t.Q; ++i) {
      for (int d = 0; d < 3; ++d) c[3 * i + d] = t.c[i][d];
      w[i] = t.w[i];
      inv[i] = t.inv[i];
    }
    return t.Q;
  } catch (...) {
    return -1;
  }
}

}  // extern "C"
