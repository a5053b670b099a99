// lattice.cpp — C ABI of the gfx950 LBGK library (include/aslhip.h).
//
// Host-side state of one lattice: device buffers, the boundary-condition
// list, the driver's execute order, and the finalize step that turns ASL's
// ghost-fill BCs into per-link resolution tables for the fused step kernel.
//
// How ASL's BC semantics map onto links.  ASL streams by pulling
// f_k(x) <- f_k(x - c_k) after the BCs have overwritten some ghost entries
// (g, k) (src/num/aslLBGK.cxx:237-250; src/num/aslLBGKBC.cxx).  So the value a
// computation node x receives along direction k is the BC value of entry
// (x - c_k, k) if some executed BC writes that entry (the last one executed
// wins), and otherwise the collided population of x - c_k.  finalize()
// resolves every incoming link of every computed cell that way, in the
// driver's execute order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/aslhip.h"
#include "kernels.hpp"
#include "lattice_consts.hpp"
#include "status.hpp"

using namespace aslhip;

namespace {
thread_local std::string g_err;
}  // namespace

int aslhip::fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

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

struct Templ {
  int Q;
  int c[27][3];
  double w[27];
  int inv[27];
};

template <int Q>
Templ make_templ_q() {
  Templ t;
  t.Q = Q;
  for (int i = 0; i < Q; ++i) {
    for (int d = 0; d < 3; ++d) t.c[i][d] = Lat<Q>::c[i][d];
    t.w[i] = Lat<Q>::w(i);
    t.inv[i] = inv_dir<Q>(i);
  }
  return t;
}

Templ make_templ(int q) {
  if (q == 15) return make_templ_q<15>();
  if (q == 27) return make_templ_q<27>();
  return make_templ_q<19>();
}

struct BCDesc {
  int kind = 0;
  unsigned faces = 0;
  double p = 0, u[3] = {0, 0, 0};
  bool has_rot = false;
  double rot[6] = {0, 0, 0, 0, 0, 0};
  std::vector<float> map, cd;
  bool is_map() const { return kind >= ASLHIP_BC_NOSLIP_MAP; }
};

// v . c_i with zero terms left out, in type C.
template <class C>
C dot_c(const Templ &t, const C *u, int i) {
  C a = C(0);
  bool first = true;
  for (int d = 0; d < 3; ++d) {
    if (t.c[i][d] == 0) continue;
    C v = t.c[i][d] > 0 ? u[d] : -u[d];
    a = first ? v : a + v;
    first = false;
  }
  return a;
}

}  // namespace

struct aslhip_lattice {
  int n[3] = {0, 0, 0};
  int S[3] = {0, 0, 0};
  int sx = 0, sy = 0, len = 0;
  int Q = 19, dtype = ASLHIP_F32;
  bool periodic = false;
  Templ t;
  size_t selem = 4, melem = 4, celem = 4;  // storage / macroscopic / compute element bytes
  long long qs_bytes = 0;
  hipStream_t stream = nullptr;
  int device = -1;  // the device current at creation (its memory and stream live there)
  // device buffers
  char *f = nullptr;
  void *rho = nullptr, *u = nullptr;
  unsigned char *code = nullptr;
  unsigned short *pat = nullptr;
  LinkRec *rec = nullptr;
  unsigned char *mode = nullptr;
  void *val = nullptr;
  VLink *vl = nullptr;
  VLinkBC *vlbc = nullptr;
  double *red_partial = nullptr, *red_out = nullptr;
  // [2][STEP_F_COUNT]: step arguments of even and odd steps for each value of
  // StepArgs::flags (the two parities differ when only even steps run tiles)
  StepArgs *d_args = nullptr;
  std::vector<StepArgs> h_args;  // host copy of d_args (empty: not uploaded yet)
  // Captured runs of plain steps (the velocity-map link fills included):
  // full chunks of ASLHIP_GRAPH_CHUNK steps and the shorter remainders of
  // step counts that are not a multiple of it, each for a starting parity
  // and the step flags it was captured with; dropped after finalize.
  struct Graph {
    int n;
    bool odd;
    unsigned fl;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  int nvl = 0;
  // parameters
  double nu = 0.1;
  bool compressible = true, comp_rho = true, comp_vel = true;
  bool has_omega = false;
  double omega[3] = {0., 0., 0.};
  // learned closure (network parameters in device memory)
  bool has_closure = false;
  bool turbulence = false;   // LBGKTurbulence
  bool closure_mfma = true;  // evaluate the network with BF16 matrix instructions
  ClosureParams *d_cl = nullptr;
  float *d_feat = nullptr;  // closure features scratch (7 per interior cell)
  double dx = 1., pos0[3] = {0., 0., 0.};
  // a stream reading rho / u through aslhip_macro_device: the next work that
  // writes them waits for what it has queued by then
  hipStream_t consumer = nullptr;
  hipEvent_t ev_macro = nullptr, ev_consumer = nullptr;
  // visualisation fills (aslhip_vis_add): kind, device copy of the map, velocity
  struct Vis {
    int kind;
    float *map;
    double u[3];
    bool has_rot;
    double rot[6];
  };
  std::vector<Vis> vis;
  // BCs and order
  std::vector<BCDesc> bcs;
  std::vector<int> order, pending;
  bool dirty = true;
  bool initialized = false;
  bool fresh = false;        // no step since init
  bool bc_since_init = false;
  bool next_odd = true;
  // a step requested with ASLHIP_STEP_DEFER that has not run yet
  bool deferred = false;
  bool deferred_raw = false;
  bool deferred_first = false;
  // launch tuning
  int tune_block = 512;   // workgroup size (measured best with runs of 4 chunks per XCD)
  bool force_wide = false;  // 64-bit addressing on any size (ASLHIP_TUNE_WIDE)
  int tune_remap = 4;  // chunk runs per XCD (measured best for graph chunks; 64 for 16-bit storage; the autotuner may choose others)
  int tune_cpl = 0;    // cells per lane along z: 0 automatic, else 1, 2 or 4 (ASLHIP_TUNE_CPL)
  int tune_lds = -1;   // LDS-staged tiles: 0 off, 1 every step, 2 even steps, -1 automatic (ASLHIP_TUNE_LDS)
  bool vmap_skip_outer = false;  // velocity-map averages skip outer-layer neighbours (opt-in)
  // face rule: per face slot * 4 + mode of the winning slice BC, and values
  int fface[6] = {0, 0, 0, 0, 0, 0};
  float fval32[6][27] = {};
  double fval64[6][27] = {};
  // stats
  long long n_counted = 0, n_computed = 0, n_boundary = 0, n_patterns = 0;
  std::vector<unsigned char> h_code;

  int idx(int x, int y, int z) const { return x * sx + y * sy + z; }
};

namespace {

int check(aslhip_lattice *h) {
  if (!h) return fail(ASLHIP_EINVAL, "ASL ERROR: null lattice handle");
  return 0;
}

// Destroys the captured runs once the stream has finished any replay of them.
void drop_graphs(aslhip_lattice *h) {
  if (!h->graphs.empty() && h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto &g : h->graphs) (void)hipGraphExecDestroy(g.exec);
  h->graphs.clear();
}

void free_all(aslhip_lattice *h) {
  drop_graphs(h);
  void *ptrs[] = {h->f,  h->rho,  h->u,           h->code,         h->pat, h->rec, h->mode,
                  h->val, h->vl, h->vlbc, h->red_partial, h->red_out, h->d_args, h->d_cl, h->d_feat};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (auto &v : h->vis)
    if (v.map) (void)hipFree(v.map);
  if (h->ev_macro) (void)hipEventDestroy(h->ev_macro);
  if (h->ev_consumer) (void)hipEventDestroy(h->ev_consumer);
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

// Work about to be queued on the lattice's stream may overwrite rho / u (and
// the populations): it first waits for whatever the consumer stream named in
// aslhip_macro_device has queued up to now.
int order_after_consumer(aslhip_lattice *h) {
  if (!h->consumer) return 0;
  HIPCHK(hipEventRecord(h->ev_consumer, h->consumer));
  HIPCHK(hipStreamWaitEvent(h->stream, h->ev_consumer, 0));
  h->consumer = nullptr;
  return 0;
}

// Relaxation rate w = 1/(3 nu + 1/2) in the compute type
// (src/num/aslLBGK.cxx:203).
// StepArgs::flags bits fixed by the lattice's parameters.
unsigned base_flags(const aslhip_lattice *h) {
  return (h->compressible ? STEP_F_COMPRESSIBLE : 0u) | (h->has_omega ? STEP_F_OMEGA : 0u) |
         (h->has_closure ? STEP_F_CLOSURE : 0u) | (h->has_closure && h->closure_mfma ? STEP_F_CLOSURE_MFMA : 0u) |
         (h->turbulence ? STEP_F_TURB : 0u);
}

double relaxation(const aslhip_lattice *h) {
  if (h->dtype == ASLHIP_F64) return 1. / (3. * h->nu + .5);
  const float nu = static_cast<float>(h->nu);
  return static_cast<double>(1.f / (3.f * nu + .5f));
}

// Slot values of one BC in the compute type C.
template <class C>
void bc_values(const aslhip_lattice *h, const BCDesc &b, unsigned char &md, C *val) {
  const Templ &t = h->t;
  const C u[3] = {C(b.u[0]), C(b.u[1]), C(b.u[2])};
  const C p = C(b.p);
  for (int k = 0; k < t.Q; ++k) val[k] = C(0);
  switch (b.kind) {
    case ASLHIP_BC_NOSLIP:
    case ASLHIP_BC_NOSLIP_MAP:
      md = MODE_BOUNCE;
      break;
    case ASLHIP_BC_CONSTANT_VELOCITY:  // f_kbar(p) + 6.*(velocity*c_k)
      md = MODE_BOUNCE;
      for (int k = 1; k < t.Q; ++k) val[k] = C(6.) * dot_c(t, u, k);
      break;
    case ASLHIP_BC_CONSTANT_PRESSURE:
      md = MODE_CONST;
      for (int k = 1; k < t.Q; ++k) val[k] = p;
      break;
    case ASLHIP_BC_CONSTANT_PRESSURE_VELOCITY:  // pressure + 3.*(velocity*c_k)
      md = MODE_CONST;
      for (int k = 1; k < t.Q; ++k) val[k] = p + C(3.) * dot_c(t, u, k);
      break;
    case ASLHIP_BC_PRESSURE_VELOCITY_MAP: {  // src/num/aslLBGKBC.cxx:400-415
      md = MODE_CONST;
      C uu = u[0] * u[0];
      uu = uu + u[1] * u[1];
      uu = uu + u[2] * u[2];
      for (int k = 1; k < t.Q; ++k) {
        const C a = dot_c(t, u, k);
        if (h->compressible) {
          val[k] = ((p + (C(3.) * p) * a) + ((C(4.5) * p) * a) * a) - (C(1.5) * p) * uu;
        } else {
          const C u3[3] = {C(3.) * u[0], C(3.) * u[1], C(3.) * u[2]};
          val[k] = ((p + dot_c(t, u3, k)) + (C(4.5) * a) * a) - C(1.5) * uu;
        }
      }
      break;
    }
    case ASLHIP_BC_VELOCITY_MAP:
      md = MODE_LINK;
      break;
    default:
      md = MODE_NONE;
  }
}

// Out-of-block-safe map read (src/acl/aclMath/aclMathAlg.cxx:84-94).
inline float map_safe(const std::vector<float> &m, long long i, long long len, float out) {
  return (i < 0 || i > len - 1) ? out : m[(size_t)i];
}

// isComputationNode (src/num/aslBCond.cxx:204-211): map > 0 [&& cdomain > 0].
inline bool comp_node(const BCDesc &b, long long i, long long len) {
  bool r = map_safe(b.map, i, len, -.9f) > 0.f;
  if (!b.cd.empty()) r = r && map_safe(b.cd, i, len, -1.f) > 0.f;
  return r;
}

struct PatHash {
  size_t operator()(const std::string &s) const { return std::hash<std::string>()(s); }
};

// allow_open: the next step reads ghost links raw (the first step after
// initF with no BC executed before it), so links no BC fills are allowed; the
// lattice is finalized again before any later step.
int finalize_impl(aslhip_lattice *h, bool allow_open = false) {
  drop_graphs(h);  // they hold the link tables replaced below
  const Templ &t = h->t;
  const int Q = t.Q, len = h->len;
  std::vector<unsigned char> code(len, CODE_SKIP);
  std::vector<unsigned short> pat(len, 0);
  std::vector<LinkRec> recs(1);
  std::memset(&recs[0], 0, sizeof(LinkRec));
  std::vector<VLink> vlinks;
  std::vector<VLinkBC> vlbc(256);
  const int nord = static_cast<int>(h->order.size());
  if (nord > 254) return fail(ASLHIP_EINVAL, "ASL ERROR: more than 254 boundary-condition executions per step");
  int off[27];
  for (int k = 0; k < Q; ++k) off[k] = t.c[k][0] * h->sx + t.c[k][1] * h->sy + t.c[k][2];
  auto interior = [&](int x, int y, int z) {
    return x >= 1 && x <= h->n[0] && y >= 1 && y <= h->n[1] && z >= 1 && z <= h->n[2];
  };

  h->n_counted = h->n_computed = h->n_boundary = 0;
  if (h->periodic) {
    if (nord > 0 || !h->bcs.empty())
      return fail(ASLHIP_EINVAL, "ASL ERROR: boundary conditions on a periodic lattice");
    for (int x = 1; x <= h->n[0]; ++x)
      for (int y = 1; y <= h->n[1]; ++y)
        for (int z = 1; z <= h->n[2]; ++z) code[h->idx(x, y, z)] = CODE_FLUID;
    h->n_counted = h->n_computed = (long long)h->n[0] * h->n[1] * h->n[2];
  } else {
    // BC of each order position; map BCs applied.
    std::vector<const BCDesc *> ob(nord);
    for (int o = 0; o < nord; ++o) ob[o] = &h->bcs[h->order[o]];
    std::vector<int> map_pos;
    for (int o = 0; o < nord; ++o)
      if (ob[o]->is_map()) map_pos.push_back(o);
    static const int face_dir[6] = {1, 4, 2, 5, 3, 6};  // X0 XE Y0 YE Z0 ZE: vectors[1..6]

    // Slot (order position + 1) of the last executed BC that writes entry
    // (x - c_k, k), for interior x; 0 if none.
    auto resolve = [&](int x, int y, int z, int k) -> int {
      const int i = h->idx(x, y, z);
      const int p[3] = {x, y, z};
      for (int o = nord - 1; o >= 0; --o) {
        const BCDesc &b = *ob[o];
        if (!b.is_map()) {
          // addSliceX0 ... addSliceZE (src/num/aslBCond.cxx:373-493)
          for (int fc = 0; fc < 6; ++fc) {
            if (!(b.faces & (1u << fc))) continue;
            const int axis = fc / 2;
            const int layer = (fc % 2 == 0) ? 1 : h->S[axis] - 2;
            const int *e = t.c[face_dir[fc]];
            if (p[axis] == layer && t.c[k][0] * e[0] + t.c[k][1] * e[1] + t.c[k][2] * e[2] > 0) return o + 1;
          }
        } else {
          // map guard and predicates (src/num/aslLBGKBC.cxx:235-251)
          const float mg = b.map[(size_t)(i - off[k])];
          if (mg <= 0.f && mg > -.9999f && comp_node(b, i, len)) return o + 1;
        }
      }
      return 0;
    };

    std::vector<unsigned char> computed(len, 0), counted(len, 0);
    for (int x = 1; x <= h->n[0]; ++x)
      for (int y = 1; y <= h->n[1]; ++y)
        for (int z = 1; z <= h->n[2]; ++z) {
          const int i = h->idx(x, y, z);
          bool solid = false;
          for (int o : map_pos)
            if (ob[o]->map[(size_t)i] <= 0.f) solid = true;
          computed[i] = counted[i] = solid ? 0 : 1;
        }
    // Closure: a cell that is not computed but sends an unresolved link into
    // a computed cell must be computed too (ASL collides every cell).
    for (bool changed = true; changed;) {
      changed = false;
      for (int x = 0; x < h->S[0]; ++x)
        for (int y = 0; y < h->S[1]; ++y)
          for (int z = 0; z < h->S[2]; ++z) {
            const int s = h->idx(x, y, z);
            if (computed[s]) continue;
            for (int k = 1; k < Q; ++k) {
              const int X = x + t.c[k][0], Y = y + t.c[k][1], Z = z + t.c[k][2];
              if (!interior(X, Y, Z)) continue;
              if (!computed[h->idx(X, Y, Z)] || resolve(X, Y, Z, k) != 0) continue;
              if (!interior(x, y, z)) {
                if (allow_open) continue;
                return fail(ASLHIP_EUNDEF,
                            "ASL ERROR: no boundary condition fills the ghost link (%d,%d,%d) -> (%d,%d,%d) "
                            "(ASL would stream uninitialised memory into the block)",
                            x - 1, y - 1, z - 1, X - 1, Y - 1, Z - 1);
              }
              computed[s] = 1;
              changed = true;
              break;
            }
          }
    }

    // Face rule: the last executed slice BC on each face claims the links
    // that enter through that face (see StepArgs).
    int fslot[6] = {0, 0, 0, 0, 0, 0};
    for (int o = 0; o < nord; ++o)
      if (!ob[o]->is_map())
        for (int fc = 0; fc < 6; ++fc)
          if (ob[o]->faces & (1u << fc)) fslot[fc] = o + 1;
    auto face_rule = [&](int x, int y, int z, int k) -> int {
      const int p[3] = {x, y, z};
      int best = 0;
      for (int d = 0; d < 3; ++d) {
        const int cd = t.c[k][d];
        if (cd == 0) continue;
        const int fc = 2 * d + (cd > 0 ? 0 : 1);
        const bool on = (fc % 2 == 0) ? p[d] == 1 : p[d] == h->n[d];
        if (on && fslot[fc] > best) best = fslot[fc];
      }
      return best;
    };
    for (int fc = 0; fc < 6; ++fc) {
      h->fface[fc] = 0;
      for (int k = 0; k < 27; ++k) {
        h->fval32[fc][k] = 0.f;
        h->fval64[fc][k] = 0.;
      }
      if (!fslot[fc]) continue;
      const BCDesc &b = *ob[fslot[fc] - 1];
      unsigned char md = MODE_NONE;
      double v64[27];
      float v32[27];
      bc_values<double>(h, b, md, v64);
      bc_values<float>(h, b, md, v32);
      h->fface[fc] = fslot[fc] * 4 + md;
      for (int k = 0; k < Q; ++k) {
        h->fval32[fc][k] = v32[k];
        h->fval64[fc][k] = v64[k];
      }
    }

    // Cells that some map BC treats as ghost nodes (map <= 0), any layer.
    std::vector<unsigned char> solid_any(len, 0);
    for (int o : map_pos)
      for (int i = 0; i < len; ++i)
        if (ob[o]->map[(size_t)i] <= 0.f) solid_any[i] = 1;
    std::unordered_map<std::string, int, PatHash> pid;
    for (int x = 1; x <= h->n[0]; ++x)
      for (int y = 1; y <= h->n[1]; ++y)
        for (int z = 1; z <= h->n[2]; ++z) {
          const int i = h->idx(x, y, z);
          if (!computed[i]) continue;
          LinkRec r;
          std::memset(&r, 0, sizeof r);
          bool any = false;
          // A link into x can only be BC-resolved if x is in a face layer or
          // its source x - c_k is a map ghost node.
          bool near = x == 1 || x == h->n[0] || y == 1 || y == h->n[1] || z == 1 || z == h->n[2];
          for (int k = 1; k < Q && !near && !map_pos.empty(); ++k) near = solid_any[i - off[k]] != 0;
          for (int k = 1; k < Q && near; ++k) {
            const int s = resolve(x, y, z, k);
            r.slot[k] = static_cast<unsigned char>(s);
            // the step kernel resolves this link itself unless the face
            // rule disagrees with the full resolution
            any = any || s != face_rule(x, y, z, k);
          }
          // Outgoing link k of a computed ghost cell: skip it if its receiver
          // resolves it (computation nodes are never the source of a BC entry).
          for (int k = 1; k < Q && !counted[i]; ++k) {
            const int X = x + t.c[k][0], Y = y + t.c[k][1], Z = z + t.c[k][2];
            if (interior(X, Y, Z) && computed[h->idx(X, Y, Z)] && resolve(X, Y, Z, k) != 0) {
              r.skip |= 1u << k;
              any = true;
            }
          }
          ++h->n_computed;
          if (counted[i]) ++h->n_counted;
          if (!any && counted[i]) {
            code[i] = CODE_FLUID;
            continue;
          }
          code[i] = counted[i] ? CODE_FLUID_BC : CODE_GHOST;
          if (counted[i]) ++h->n_boundary;
          std::string key(reinterpret_cast<const char *>(&r), sizeof r);
          auto it = pid.find(key);
          int id;
          if (it == pid.end()) {
            id = static_cast<int>(recs.size());
            if (id > 65535) return fail(ASLHIP_EINVAL, "ASL ERROR: more than 65535 distinct boundary link patterns");
            recs.push_back(r);
            pid.emplace(key, id);
          } else {
            id = it->second;
          }
          pat[i] = static_cast<unsigned short>(id);
        }

    // Boundary-link entries of velocity map BCs.
    for (int o = 0; o < nord; ++o) {
      const BCDesc &b = *ob[o];
      if (b.kind != ASLHIP_BC_VELOCITY_MAP) continue;
      VLinkBC &vb = vlbc[o + 1];
      for (int d = 0; d < 3; ++d) {
        vb.u[d] = b.u[d];
        vb.omega[d] = b.rot[d];
        vb.centre[d] = b.rot[3 + d];
      }
      vb.has_rot = b.has_rot ? 1 : 0;
      for (int g = 0; g < len; ++g) {
        const float mg = b.map[(size_t)g];
        if (!(mg <= 0.f && mg > -.9999f)) continue;
        VLink l{g, 0u, 0u, o + 1};
        const int gx = g / h->sx, gy = (g % h->sx) / h->sy, gz = g % h->sy;
        for (int i = 1; i < Q; ++i) {
          const long long p = (long long)g + off[i];
          if (!comp_node(b, p, len)) continue;
          l.comp |= 1u << i;
          const int X = gx + t.c[i][0], Y = gy + t.c[i][1], Z = gz + t.c[i][2];
          if (interior(X, Y, Z) && computed[h->idx(X, Y, Z)] && h->idx(X, Y, Z) == p &&
              resolve(X, Y, Z, i) == o + 1)
            l.own |= 1u << i;
        }
        if (!l.own) continue;
        for (int i = 1; i < Q; ++i) {
          if (!((l.comp >> i) & 1u)) continue;
          const int X = gx + t.c[i][0], Y = gy + t.c[i][1], Z = gz + t.c[i][2];
          if (!interior(X, Y, Z) && h->vmap_skip_outer) {
            l.comp &= ~(1u << i);  // left out of the density average (opt-in)
            continue;
          }
          if (!interior(X, Y, Z) || !computed[h->idx(X, Y, Z)])
            return fail(ASLHIP_EUNDEF,
                        "ASL ERROR: velocity map BC at ghost (%d,%d,%d) averages a neighbour that is not a "
                        "computed cell",
                        gx - 1, gy - 1, gz - 1);
        }
        vlinks.push_back(l);
      }
    }
  }
  h->n_patterns = static_cast<long long>(recs.size()) - 1;

  // Slot tables.
  std::vector<unsigned char> mode(256, MODE_NONE);
  std::vector<char> val(256 * (size_t)Q * h->celem, 0);
  for (int o = 0; o < nord; ++o) {
    const BCDesc &b = h->bcs[h->order[o]];
    if (h->celem == 8)
      bc_values<double>(h, b, mode[o + 1], reinterpret_cast<double *>(val.data()) + (size_t)(o + 1) * Q);
    else
      bc_values<float>(h, b, mode[o + 1], reinterpret_cast<float *>(val.data()) + (size_t)(o + 1) * Q);
  }

  // Upload.
  HIPCHK(hipMemcpyAsync(h->code, code.data(), len, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->pat, pat.data(), (size_t)len * 2, hipMemcpyHostToDevice, h->stream));
  if (h->rec) HIPCHK(hipFree(h->rec));
  h->rec = nullptr;
  HIPCHK(hipMalloc(&h->rec, recs.size() * sizeof(LinkRec)));
  HIPCHK(hipMemcpyAsync(h->rec, recs.data(), recs.size() * sizeof(LinkRec), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->mode, mode.data(), 256, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->val, val.data(), val.size(), hipMemcpyHostToDevice, h->stream));
  if (h->vl) HIPCHK(hipFree(h->vl));
  h->vl = nullptr;
  h->nvl = static_cast<int>(vlinks.size());
  if (h->nvl) {
    HIPCHK(hipMalloc(&h->vl, vlinks.size() * sizeof(VLink)));
    HIPCHK(hipMemcpyAsync(h->vl, vlinks.data(), vlinks.size() * sizeof(VLink), hipMemcpyHostToDevice, h->stream));
  }
  HIPCHK(hipMemcpyAsync(h->vlbc, vlbc.data(), 256 * sizeof(VLinkBC), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->h_code.swap(code);
  h->dirty = allow_open;
  return 0;
}

int run_vlinks(aslhip_lattice *h, bool after_odd) {
  if (!h->nvl) return 0;
  VLinkArgs a;
  a.f = h->f;
  a.qs_bytes = h->qs_bytes;
  a.sx = h->sx;
  a.sy = h->sy;
  a.links = h->vl;
  a.nlinks = h->nvl;
  a.bc = h->vlbc;
  a.dx = h->dx;
  for (int d = 0; d < 3; ++d) a.pos0[d] = h->pos0[d];
  a.after_odd = after_odd ? 1 : 0;
  HIPCHK(launch_vlinks(h->Q, h->dtype, a, h->stream));
  return 0;
}

bool is_wide(const aslhip_lattice *h) { return h->force_wide || h->qs_bytes * h->Q > 0xffffffffLL; }

// Cells per lane the step kernel runs with: runs of 2 or 4 cells along z on
// non-periodic lattices under 4 GiB without a per-cell relaxation, when nz is
// a multiple of the run; else one cell per lane (the longest run that divides
// nz).  Automatic: one cell per lane.
// LDS-staged tiles apply to non-periodic lattices under 4 GiB without a
// per-cell relaxation unless runs of 2 or 4 cells per lane are asked for:
// 0 none, 1 every step, 2 even steps only (odd steps one cell per lane).
// Automatic: every step for FP32 storage whose populations outgrow the
// 256 MiB Infinity Cache (the sphere channel, 331 MB: 0.1225 against 0.128
// ms per step); none for lattices the cache holds (64^3 cavity, 22 MB: 15 us
// against 11 us) and for FP16 / BF16 / FP64 storage (the direct kernel is
// faster there).
constexpr long long TILE_MIN_BYTES = 256LL << 20;
int effective_lds(const aslhip_lattice *h) {
  const bool big = h->qs_bytes * h->Q > TILE_MIN_BYTES;
  const int mode = h->tune_lds < 0 ? (h->dtype == ASLHIP_F32 && big ? 1 : 0) : h->tune_lds;
  return h->tune_cpl <= 1 && !(h->periodic || is_wide(h) || h->has_closure || h->turbulence) ? mode : 0;
}

int effective_cpl(const aslhip_lattice *h) {
  if (h->periodic || is_wide(h) || h->has_closure || h->turbulence || effective_lds(h)) return 1;
  int c = h->tune_cpl;
  if (c == 0) c = 1;
  while (c > 1 && h->n[2] % c) c /= 2;
  return c;
}

// Workgroup size the step kernel runs with (runs of cells: at most
// STEP_RUN_BLOCK; LDS-staged tiles on every step: STEP_TILE; tiles on even
// steps only: the odd steps' workgroup).
int step_workgroup(const aslhip_lattice *h) {
  if (effective_lds(h) == 1) return STEP_TILE;
  return effective_cpl(h) > 1 && h->tune_block > STEP_RUN_BLOCK ? STEP_RUN_BLOCK : h->tune_block;
}

// Kernel arguments of one step of h (flags filled in by the caller).  With
// block and cpl 0 the lattice's own tuning for a step of that parity (LDS-
// staged tiles included), else one cell per lane in workgroups of that size.
StepArgs step_args(const aslhip_lattice *h, int block = 0, int cpl = 0, bool odd = false) {
  StepArgs a;
  std::memset(&a, 0, sizeof a);  // compared bytewise by upload_args
  for (int k = 0; k < 27; ++k) a.fd[k] = k < h->Q ? h->f + k * h->qs_bytes : nullptr;
  a.qs32 = is_wide(h) ? 0u : static_cast<unsigned>(h->qs_bytes);
  for (int fc = 0; fc < 6; ++fc) {
    a.fface[fc] = h->fface[fc];
    for (int k = 0; k < 27; ++k) {
      a.fval32[fc][k] = h->fval32[fc][k];
      a.fval64[fc][k] = h->fval64[fc][k];
    }
  }
  a.sx = h->sx;
  a.sy = h->sy;
  a.nx = h->n[0];
  a.ny = h->n[1];
  a.nz = h->n[2];
  a.len = h->len;
  a.cpl = cpl > 0 ? cpl : effective_cpl(h);
  a.block = block > 0 ? block : h->tune_block;
  if (a.cpl > 1 && a.block > STEP_RUN_BLOCK) a.block = STEP_RUN_BLOCK;
  const int lds = effective_lds(h);
  a.tile = block <= 0 && cpl <= 0 && (lds == 1 || (lds == 2 && !odd)) ? 1 : 0;
  if (a.tile) a.block = STEP_TILE;

  a.nzv = h->n[2] / a.cpl;
  // lanes per plane: every cell of the rows (ghost columns included), or the runs of the interior
  const int lanes = a.cpl > 1 ? h->n[1] * a.nzv : h->n[1] * h->sy;
  a.bpp = (lanes + a.block - 1) / a.block;
  a.nblocks = a.bpp * h->n[0];
  a.zmagic = ((1ULL << 40) + (unsigned long long)h->sy - 1) / (unsigned long long)h->sy;
  a.vmagic = ((1ULL << 40) + (unsigned long long)a.nzv - 1) / (unsigned long long)a.nzv;
  a.remap = h->tune_remap;
  a.code = h->code;
  a.pat = h->pat;
  a.rec = h->rec;
  a.mode = h->mode;
  a.val = h->val;
  a.rho = h->rho;
  a.u = h->u;
  a.w = relaxation(h);
  for (int d = 0; d < 3; ++d) {
    a.om32[d] = static_cast<float>(h->omega[d]);
    a.om64[d] = h->omega[d];
  }
  a.cl = h->d_cl;
  a.cl_tau0 = 3.f * static_cast<float>(h->nu) + .5f;
  a.nu32 = static_cast<float>(h->nu);
  a.nu64 = h->nu;
  a.flags = base_flags(h);
  return a;
}

// Runs nsteps steps now.  raw_first: the first one reads ghost links raw
// (no BC executed before it since init); macro_last: the last one writes rho/u.
// The execute order of the BCs since the last step becomes the active one;
// the tables are rebuilt if anything changed.
int prepare(aslhip_lattice *h, bool allow_open) {
  if (!h->pending.empty()) {
    if (h->pending != h->order) {
      h->order = h->pending;
      h->dirty = true;
    }
    h->pending.clear();
  }
  return h->dirty ? finalize_impl(h, allow_open) : 0;
}

// Brings the device copies of the step arguments (one per flags value) up to
// date; they change with the BC tables, viscosity, compressibility and tuning.
int upload_args(aslhip_lattice *h) {
  const StepArgs base[2] = {step_args(h, 0, 0, false), step_args(h, 0, 0, true)};
  if (!h->h_args.empty()) {
    if (std::memcmp(&base[0], &h->h_args[base[0].flags], sizeof(StepArgs)) == 0 &&
        std::memcmp(&base[1], &h->h_args[STEP_F_COUNT + base[1].flags], sizeof(StepArgs)) == 0)
      return 0;
    // the previous upload reads h_args: let it finish before overwriting
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  h->h_args.resize(2 * STEP_F_COUNT);
  for (int p = 0; p < 2; ++p)
    for (unsigned fl = 0; fl < STEP_F_COUNT; ++fl) {
      h->h_args[p * STEP_F_COUNT + fl] = base[p];
      h->h_args[p * STEP_F_COUNT + fl].flags = fl;
    }
  // stream-ordered: steps queued before still see the previous arguments
  HIPCHK(hipMemcpyAsync(h->d_args, h->h_args.data(), 2 * STEP_F_COUNT * sizeof(StepArgs), hipMemcpyHostToDevice,
                        h->stream));
  return 0;
}

int launch_one(aslhip_lattice *h, unsigned fl, std::vector<hipEvent_t> *ev, int s) {
  StepLaunch a;
  const int p0 = h->next_odd ? STEP_F_COUNT : 0;  // this parity's argument table
  const StepArgs &A = h->h_args[p0];
  a.nblocks = A.nblocks;
  a.block = A.block;
  a.args = h->d_args + p0 + fl;
  // LBGKTurbulence takes precedence over a closure in the collision
  a.variant = h->turbulence ? 1 : (h->has_closure ? (h->closure_mfma ? 2 : 1) : 0);
  a.cpl = A.cpl;
  a.tile = A.tile;
  if (ev) HIPCHK(hipEventRecord((*ev)[2 * s], h->stream));
  HIPCHK(launch_step(h->Q, h->dtype, h->next_odd, h->periodic, is_wide(h), a, h->stream));
  if (ev) HIPCHK(hipEventRecord((*ev)[2 * s + 1], h->stream));
  const bool was_odd = h->next_odd;
  h->next_odd = !h->next_odd;
  return run_vlinks(h, was_odd);
}

// At most this many captured remainders (runs shorter than a chunk) are kept.
constexpr size_t MAX_TAIL_GRAPHS = 6;

// The captured run of n plain steps starting at parity odd with the current
// step flags, captured now if there is none.
int get_graph(aslhip_lattice *h, int n, bool odd, hipGraphExec_t *out) {
  const unsigned fl = base_flags(h);
  for (const auto &g : h->graphs)
    if (g.n == n && g.odd == odd && g.fl == fl) {
      *out = g.exec;
      return 0;
    }
  // A run captured with other flags (frame, compressibility, relaxation
  // model) names the old kernel variant and argument slot, and the oldest
  // remainder makes room for a new one: destroyed once replays of them have
  // finished.
  auto drop = [h](size_t i) -> int {
    HIPCHK(hipStreamSynchronize(h->stream));
    (void)hipGraphExecDestroy(h->graphs[i].exec);
    h->graphs.erase(h->graphs.begin() + static_cast<long>(i));
    return 0;
  };
  for (size_t i = 0; i < h->graphs.size(); ++i)
    if (h->graphs[i].n == n && h->graphs[i].odd == odd) {
      if (int r = drop(i)) return r;
      break;
    }
  if (n != ASLHIP_GRAPH_CHUNK) {
    size_t tails = 0, oldest = 0;
    for (size_t i = h->graphs.size(); i-- > 0;)
      if (h->graphs[i].n != ASLHIP_GRAPH_CHUNK) {
        ++tails;
        oldest = i;
      }
    if (tails >= MAX_TAIL_GRAPHS)
      if (int r = drop(oldest)) return r;
  }
  const bool keep = h->next_odd;
  h->next_odd = odd;
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int r = 0;
  for (int s = 0; s < n && !r; ++s) r = launch_one(h, fl, nullptr, s);
  hipError_t e = hipStreamEndCapture(h->stream, &graph);
  h->next_odd = keep;
  if (r) {
    if (graph) (void)hipGraphDestroy(graph);
    return r;
  }
  if (e != hipSuccess) return fail(ASLHIP_EHIP, "ASL ERROR: graph capture: %s", hipGetErrorString(e));
  hipGraphExec_t g = nullptr;
  e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(ASLHIP_EHIP, "ASL ERROR: graph instantiate: %s", hipGetErrorString(e));
  h->graphs.push_back({n, odd, fl, g});
  *out = g;
  return 0;
}

// graph: run whole chunks of plain steps (not the raw first one, not the
// one writing rho/u) as hipGraph replays.  consume: the BC executions
// recorded since the last step become the active order first (a deferred
// step runs with the order folded in when it was requested instead: the
// executions recorded since then belong to the step after it).
int run_steps(aslhip_lattice *h, int nsteps, bool first_after_init, bool raw_first, bool macro_last,
              std::vector<hipEvent_t> *ev, bool graph = false, bool consume = true) {
  if (nsteps <= 0) return 0;
  if (int r = order_after_consumer(h)) return r;
  const bool allow_open = raw_first && nsteps == 1;
  if (int r = consume ? prepare(h, allow_open) : (h->dirty ? finalize_impl(h, allow_open) : 0)) return r;
  if (first_after_init && !raw_first) {
    // the BCs were applied to the initial state: velocity-map links too
    int r = run_vlinks(h, false);
    if (r) return r;
  }
  if (int r = upload_args(h)) return r;
  for (int s = 0; s < nsteps;) {
    const bool raw = s == 0 && raw_first;
    const int plain = nsteps - s - (macro_last ? 1 : 0);  // plain steps left
    if (graph && !ev && !raw && plain >= 1) {
      const int n = plain < ASLHIP_GRAPH_CHUNK ? plain : ASLHIP_GRAPH_CHUNK;
      hipGraphExec_t g = nullptr;
      if (int r = get_graph(h, n, h->next_odd, &g)) return r;
      HIPCHK(hipGraphLaunch(g, h->stream));
      s += n;
      if (n & 1) h->next_odd = !h->next_odd;
      continue;
    }
    unsigned fl = base_flags(h);
    if (raw) fl |= STEP_F_RAW;
    if (s == nsteps - 1 && macro_last) {
      if (h->comp_rho) fl |= STEP_F_RHO;
      if (h->comp_vel) fl |= STEP_F_VEL;
    }
    if (int r = launch_one(h, fl, ev, s)) return r;
    ++s;
  }
  return 0;
}

// Runs a deferred step, writing rho and u (the host is about to look).  It
// uses the BC order folded in when it was requested.
int flush_deferred(aslhip_lattice *h) {
  if (!h->deferred) return 0;
  h->deferred = false;
  return run_steps(h, 1, h->deferred_first, h->deferred_raw, true, nullptr, false, false);
}

int step_impl(aslhip_lattice *h, int nsteps, unsigned flags, std::vector<hipEvent_t> *ev) {
  if (nsteps < 0) return fail(ASLHIP_EINVAL, "ASL ERROR: negative step count");
  if (!h->initialized) return fail(ASLHIP_ESTATE, "ASL ERROR: LBGK::execute before initF");
  if (h->deferred) {  // a newer step follows: no need for its rho/u
    h->deferred = false;
    int r = run_steps(h, 1, h->deferred_first, h->deferred_raw, false, nullptr, false, false);
    if (r) return r;
  }
  // the first step after initF reads raw ghost values unless a BC ran before it
  const bool first = h->fresh && nsteps > 0;
  const bool raw = first && !h->bc_since_init;
  if (nsteps > 0) h->fresh = false;
  const bool graph = (flags & ASLHIP_STEP_GRAPH) != 0;
  if ((flags & ASLHIP_STEP_DEFER) && nsteps > 0) {
    int r = run_steps(h, nsteps - 1, first, raw, false, ev, graph);
    if (r) return r;
    h->deferred_first = first && nsteps == 1;
    h->deferred_raw = raw && nsteps == 1;
    // the deferred step takes the BC executions recorded up to now; later
    // ones stay pending for the step after it
    if (int r2 = prepare(h, h->deferred_raw)) return r2;
    h->deferred = true;
    return 0;
  }
  return run_steps(h, nsteps, first, raw, (flags & ASLHIP_STEP_WRITE_MACRO) != 0, ev, graph);
}

int copy_out(aslhip_lattice *h, const void *dev, void *host, size_t bytes) {
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost));
  return 0;
}

}  // namespace

int aslhip::read_f_through(aslhip_lattice *h, void *out, void *buf, size_t buf_bytes) {
  if (int r = flush_deferred(h)) return r;
  const size_t cell_bytes = (size_t)h->Q * h->melem;  // one cell's Q values
  const size_t per = std::min(buf_bytes / cell_bytes, (size_t)h->len);
  if (per < 1) return fail(ASLHIP_EINVAL, "ASL ERROR: read_f: staging buffer smaller than one cell");
  UnpermuteArgs a;
  a.f = h->f;
  a.qs_bytes = h->qs_bytes;
  a.sx = h->sx;
  a.sy = h->sy;
  a.len = h->len;
  a.per = h->periodic ? 1 : 0;
  a.nx = h->n[0];
  a.ny = h->n[1];
  a.nz = h->n[2];
  a.after_odd = h->next_odd ? 0 : 1;
  a.out = buf;
  char *o = static_cast<char *>(out);
  for (size_t first = 0; first < (size_t)h->len; first += per) {
    const size_t count = std::min(per, (size_t)h->len - first);
    a.first = static_cast<int>(first);
    a.count = static_cast<int>(count);
    HIPCHK(launch_unpermute(h->Q, h->dtype, a, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    // piece i of the view: cells first .. first + count - 1 of direction i
    HIPCHK(hipMemcpy2D(o + first * h->melem, (size_t)h->len * h->melem, buf, count * h->melem, count * h->melem,
                       (size_t)h->Q, hipMemcpyDeviceToHost));
  }
  return 0;
}

extern "C" {

int aslhip_last_error(char *buf, size_t n) {
  if (buf && n) {
    std::strncpy(buf, g_err.c_str(), n - 1);
    buf[n - 1] = 0;
  }
  return static_cast<int>(g_err.size());
}

int aslhip_device_info(char *buf, size_t n, int *cu_count, double *hbm_bytes) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, dev));
  if (buf && n) {
    // some runtimes leave the marketing name empty: then the CU count names the part
    char name[256] = {0};
    if (p.name[0]) std::snprintf(name, sizeof name, "%s", p.name);
    else if (hipDeviceGetName(name, sizeof name, dev) != hipSuccess || !name[0])
      std::snprintf(name, sizeof name, "AMD GPU with %d CUs", p.multiProcessorCount);
    std::snprintf(buf, n, "%s (%s)", name, p.gcnArchName);
  }
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = static_cast<double>(p.totalGlobalMem);
  return 0;
}

int aslhip_lattice_create(aslhip_lattice **out, int nx, int ny, int nz, int q, int dtype, unsigned flags) {
  if (!out) return fail(ASLHIP_EINVAL, "ASL ERROR: null output handle");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1) return fail(ASLHIP_EINVAL, "ASL ERROR: block size must be positive");
  if (q != 15 && q != 19 && q != 27) return fail(ASLHIP_EINVAL, "ASL ERROR: lattice template must be D3Q15, D3Q19 or D3Q27");
  if (dtype < ASLHIP_F32 || dtype > ASLHIP_F16) return fail(ASLHIP_EINVAL, "ASL ERROR: unknown population type %d", dtype);
  const unsigned per = flags & 7u;
  if (per != 0 && per != 7) return fail(ASLHIP_EINVAL, "ASL ERROR: periodic wrap must be on all three axes");
  const long long S0 = nx + 2LL, S1 = ny + 2LL, S2 = nz + 2LL;
  const long long len = S0 * S1 * S2;
  if (len > 0x7fffffffLL - 4096) return fail(ASLHIP_EINVAL, "ASL ERROR: block too large for one lattice");
  aslhip_lattice *h = new aslhip_lattice();
  h->n[0] = nx;
  h->n[1] = ny;
  h->n[2] = nz;
  h->S[0] = (int)S0;
  h->S[1] = (int)S1;
  h->S[2] = (int)S2;
  h->sy = (int)S2;
  h->sx = (int)(S1 * S2);
  h->len = (int)len;
  h->Q = q;
  h->dtype = dtype;
  h->periodic = per == 7;
  h->t = make_templ(q);
  h->selem = dtype == ASLHIP_F64 ? 8 : (dtype == ASLHIP_F32 ? 4 : 2);
  // 16-bit storage: runs of 64 chunks per XCD measured fastest on the sphere
  // channel (0.068 against 0.076 ms per step with runs of 4)
  if (h->selem == 2) h->tune_remap = 64;
  h->melem = h->celem = dtype == ASLHIP_F64 ? 8 : 4;
  // direction arrays 256-byte aligned
  const long long qs_elems = (len + 127) / 128 * 128;
  // The step kernel finds a cell's row as (off * ceil(2^40/Sz)) >> 40, exact
  // while (cells per plane) * Sz < 2^40.
  if ((double)S1 * S2 * S2 >= 1099511627776.0)
    return fail(ASLHIP_EINVAL, "ASL ERROR: block planes too large");
  if (qs_elems * (dtype == ASLHIP_F64 ? 8 : 4) > 0xffffffffLL)
    return fail(ASLHIP_EINVAL, "ASL ERROR: block too large: one direction array must stay below 4 GiB");
  h->qs_bytes = qs_elems * (long long)h->selem;
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipStreamCreate(&h->stream);
  if (e == hipSuccess) e = hipMalloc(&h->f, (size_t)h->qs_bytes * q);
  if (e == hipSuccess) e = hipMalloc(&h->rho, (size_t)len * h->melem);
  if (e == hipSuccess) e = hipMalloc(&h->u, (size_t)len * 3 * h->melem);
  if (e == hipSuccess) e = hipMalloc(&h->code, (size_t)len);
  if (e == hipSuccess) e = hipMalloc(&h->pat, (size_t)len * 2);
  if (e == hipSuccess) e = hipMalloc(&h->mode, 256);
  if (e == hipSuccess) e = hipMalloc(&h->val, 256 * (size_t)q * h->celem);
  if (e == hipSuccess) e = hipMalloc(&h->vlbc, 256 * sizeof(VLinkBC));
  if (e == hipSuccess) e = hipMalloc(&h->d_args, 2 * STEP_F_COUNT * sizeof(StepArgs));
  if (e == hipSuccess) e = hipMalloc(&h->red_partial, (size_t)reduce_blocks((int)len) * 5 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->red_out, 5 * sizeof(double));
  if (e != hipSuccess) {
    free_all(h);
    delete h;
    return fail(e == hipErrorOutOfMemory ? ASLHIP_ENOMEM : ASLHIP_EHIP, "ASL ERROR: lattice allocation: %s",
                hipGetErrorString(e));
  }
  *out = h;
  return 0;
}

int aslhip_lattice_destroy(aslhip_lattice *h) {
  if (!h) return 0;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_all(h);
  delete h;
  return 0;
}

int aslhip_set_viscosity(aslhip_lattice *h, double nu) {
  if (int r = check(h)) return r;
  if (!(nu >= 0.) && !std::isinf(nu)) return fail(ASLHIP_EINVAL, "ASL ERROR: viscosity must be >= 0");
  if (int r = flush_deferred(h)) return r;  // a pending step runs with the old value
  h->nu = nu;
  return 0;
}

int aslhip_set_compressible(aslhip_lattice *h, int c) {
  if (int r = check(h)) return r;
  if (int r = flush_deferred(h)) return r;
  h->compressible = c != 0;
  h->dirty = true;  // pressure-velocity map values depend on it
  return 0;
}

int aslhip_set_omega(aslhip_lattice *h, const double omega[3]) {
  if (int r = check(h)) return r;
  if (omega && !(std::isfinite(omega[0]) && std::isfinite(omega[1]) && std::isfinite(omega[2])))
    return fail(ASLHIP_EINVAL, "ASL ERROR: setOmega: non-finite angular velocity");
  if (int r = flush_deferred(h)) return r;  // the pending step runs with the old frame
  h->has_omega = omega != nullptr;
  for (int d = 0; d < 3; ++d) h->omega[d] = omega ? omega[d] : 0.;
  return 0;
}

int aslhip_set_vmap_skip_outer(aslhip_lattice *h, int on) {
  if (int r = check(h)) return r;
  if (int r = flush_deferred(h)) return r;  // a pending step uses the current link tables
  if (h->vmap_skip_outer != (on != 0)) h->dirty = true;
  h->vmap_skip_outer = on != 0;
  return 0;
}

int aslhip_set_geometry(aslhip_lattice *h, double dx, const double position[3]) {
  if (int r = check(h)) return r;
  if (!position) return fail(ASLHIP_EINVAL, "ASL ERROR: set_geometry: null position");
  if (!(dx > 0.) || !std::isfinite(dx) ||
      !(std::isfinite(position[0]) && std::isfinite(position[1]) && std::isfinite(position[2])))
    return fail(ASLHIP_EINVAL, "ASL ERROR: set_geometry: bad dx or position");
  if (int r = flush_deferred(h)) return r;  // a pending step's BCs use the old positions
  h->dx = dx;
  for (int d = 0; d < 3; ++d) h->pos0[d] = position[d];
  h->dirty = true;  // captured graphs hold the velocity-link arguments by value
  return 0;
}

int aslhip_closure_set(aslhip_lattice *h, const float *params, const float mu[7], const float sc[7], float out_scale) {
  if (int r = check(h)) return r;
  if (params && (h->Q != 19 || h->dtype == ASLHIP_F64))
    return fail(ASLHIP_EINVAL, "ASL ERROR: the closure needs a D3Q19 lattice with FP32 collision");
  if (params && h->turbulence) return fail(ASLHIP_EINVAL, "ASL ERROR: LBGKTurbulence and the closure exclude each other");
  if (params && (!mu || !sc || !std::isfinite(out_scale) || out_scale < 0.f))
    return fail(ASLHIP_EINVAL, "ASL ERROR: closure_set: bad normalisation or output scale");
  if (int r = flush_deferred(h)) return r;  // a pending step runs with the old relaxation
  if (!params) {
    h->has_closure = false;
    return 0;
  }
  ClosureParams c;
  for (int i = 0; i < CL_NPARAM; ++i) {
    if (!std::isfinite(params[i])) return fail(ASLHIP_EINVAL, "ASL ERROR: closure_set: non-finite parameter %d", i);
    c.p[i] = bf16_round(params[i]);
  }
  for (int j = 0; j < CL_NIN; ++j) {
    c.mu[j] = mu[j];
    c.sc[j] = sc[j];
  }
  c.out = out_scale;
  pack_closure(c);
  if (!h->d_cl) HIPCHK(hipMalloc(&h->d_cl, sizeof(ClosureParams)));
  HIPCHK(hipStreamSynchronize(h->stream));  // steps queued earlier read the old parameters
  HIPCHK(hipMemcpy(h->d_cl, &c, sizeof c, hipMemcpyHostToDevice));
  h->has_closure = true;
  return 0;
}

int aslhip_set_turbulence(aslhip_lattice *h, int on) {
  if (int r = check(h)) return r;
  if (on && h->has_closure) return fail(ASLHIP_EINVAL, "ASL ERROR: LBGKTurbulence and the closure exclude each other");
  if (int r = flush_deferred(h)) return r;
  h->turbulence = on != 0;
  return 0;
}

int aslhip_closure_set_mode(aslhip_lattice *h, int mode) {
  if (int r = check(h)) return r;
  if (mode != ASLHIP_CLOSURE_FP32 && mode != ASLHIP_CLOSURE_BF16_MFMA)
    return fail(ASLHIP_EINVAL, "ASL ERROR: unknown closure mode %d", mode);
  if (int r = flush_deferred(h)) return r;
  h->closure_mfma = mode == ASLHIP_CLOSURE_BF16_MFMA;
  return 0;
}

int aslhip_closure_features(aslhip_lattice *h, float *out) {
  if (int r = check(h)) return r;
  if (!out) return fail(ASLHIP_EINVAL, "ASL ERROR: closure_features needs an output buffer");
  if (h->Q != 19 || h->dtype == ASLHIP_F64 || !h->periodic)
    return fail(ASLHIP_EINVAL, "ASL ERROR: closure features need a periodic D3Q19 lattice with FP32 collision");
  if (!h->initialized) return fail(ASLHIP_ESTATE, "ASL ERROR: closure_features before initF");
  if (int r = flush_deferred(h)) return r;
  if (int r = prepare(h, false)) return r;
  if (int r = upload_args(h)) return r;
  const long long ncell = (long long)h->n[0] * h->n[1] * h->n[2];
  if (!h->d_feat) HIPCHK(hipMalloc(&h->d_feat, (size_t)ncell * CL_NIN * sizeof(float)));
  HIPCHK(launch_closure_features(h->dtype, h->next_odd, h->d_args, ncell, h->d_feat, h->stream));
  return copy_out(h, h->d_feat, out, (size_t)ncell * CL_NIN * sizeof(float));
}

int aslhip_set_compute_macro(aslhip_lattice *h, int comp_rho, int comp_vel) {
  if (int r = check(h)) return r;
  h->comp_rho = comp_rho != 0;
  h->comp_vel = comp_vel != 0;
  return 0;
}

int aslhip_init_equilibrium(aslhip_lattice *h, double rho, const double u[3]) {
  if (int r = check(h)) return r;
  if (int r = flush_deferred(h)) return r;
  if (int r = order_after_consumer(h)) return r;
  InitArgs a;
  a.f = h->f;
  a.qs_bytes = h->qs_bytes;
  a.len = h->len;
  a.rho0 = rho;
  for (int d = 0; d < 3; ++d) a.u0[d] = u ? u[d] : 0.;
  a.rho_in = a.u_in = nullptr;
  a.rho = h->rho;
  a.u = h->u;
  HIPCHK(launch_init(h->Q, h->dtype, a, h->stream));
  h->initialized = true;
  h->fresh = true;
  h->bc_since_init = false;
  h->next_odd = true;
  h->pending.clear();
  return 0;
}

int aslhip_init_equilibrium_field(aslhip_lattice *h, const void *rho, const void *u) {
  if (int r = check(h)) return r;
  if (int r = flush_deferred(h)) return r;
  if (int r = order_after_consumer(h)) return r;
  if (!u) return fail(ASLHIP_EINVAL, "ASL ERROR: initF needs a velocity field");
  void *drho = nullptr, *du = nullptr;
  const size_t n = (size_t)h->len * h->melem;
  HIPCHK(hipMalloc(&du, 3 * n));
  HIPCHK(hipMemcpyAsync(du, u, 3 * n, hipMemcpyHostToDevice, h->stream));
  if (rho) {
    HIPCHK(hipMalloc(&drho, n));
    HIPCHK(hipMemcpyAsync(drho, rho, n, hipMemcpyHostToDevice, h->stream));
  }
  InitArgs a;
  a.f = h->f;
  a.qs_bytes = h->qs_bytes;
  a.len = h->len;
  a.rho0 = 1.;
  a.u0[0] = a.u0[1] = a.u0[2] = 0.;
  a.rho_in = drho;
  a.u_in = du;
  a.rho = h->rho;
  a.u = h->u;
  HIPCHK(launch_init(h->Q, h->dtype, a, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipFree(du));
  if (drho) HIPCHK(hipFree(drho));
  h->initialized = true;
  h->fresh = true;
  h->bc_since_init = false;
  h->next_odd = true;
  h->pending.clear();
  return 0;
}

int aslhip_template(int q, int *c, double *w, int *inv) {
  if (q != 15 && q != 19 && q != 27) return fail(ASLHIP_EINVAL, "ASL ERROR: no template with %d directions", q);
  const Templ t = make_templ(q);
  for (int i = 0; i < q; ++i) {
    for (int d = 0; d < 3; ++d)
      if (c) c[3 * i + d] = t.c[i][d];
    if (w) w[i] = t.w[i];
    if (inv) inv[i] = t.inv[i];
  }
  return 0;
}

int aslhip_write_f(aslhip_lattice *h, const void *f) {
  if (int r = check(h)) return r;
  if (!f) return fail(ASLHIP_EINVAL, "ASL ERROR: write_f needs populations");
  if (int r = flush_deferred(h)) return r;
  if (int r = order_after_consumer(h)) return r;
  void *d = nullptr;
  const size_t bytes = (size_t)h->len * h->Q * h->melem;
  HIPCHK(hipMalloc(&d, bytes));
  HIPCHK(hipMemcpyAsync(d, f, bytes, hipMemcpyHostToDevice, h->stream));
  WriteFArgs a;
  a.f = h->f;
  a.qs_bytes = h->qs_bytes;
  a.len = h->len;
  a.compressible = h->compressible ? 1 : 0;
  a.in = d;
  a.rho = h->rho;
  a.u = h->u;
  HIPCHK(launch_write_f(h->Q, h->dtype, a, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipFree(d));
  h->initialized = true;
  h->fresh = true;
  h->bc_since_init = false;
  h->next_odd = true;
  h->pending.clear();
  return 0;
}

int aslhip_bc_add_slice(aslhip_lattice *h, int kind, unsigned faces, double p, const double u[3], int *bc_id) {
  if (int r = check(h)) return r;
  if (h->periodic) return fail(ASLHIP_EINVAL, "ASL ERROR: slice BCs need a non-periodic lattice");
  if (kind < ASLHIP_BC_NOSLIP || kind > ASLHIP_BC_CONSTANT_PRESSURE_VELOCITY)
    return fail(ASLHIP_EINVAL, "ASL ERROR: %d is not a slice BC kind", kind);
  if (faces == 0 || faces > 63u) return fail(ASLHIP_EINVAL, "ASL ERROR: bad face mask %u", faces);
  BCDesc b;
  b.kind = kind;
  b.faces = faces;
  b.p = p;
  for (int d = 0; d < 3; ++d) b.u[d] = u ? u[d] : 0.;
  h->bcs.push_back(std::move(b));
  if (bc_id) *bc_id = static_cast<int>(h->bcs.size()) - 1;
  h->dirty = true;
  return 0;
}

int aslhip_bc_add_map(aslhip_lattice *h, int kind, const float *map, const float *cdomain, double p,
                      const double u[3], const double rot[6], int *bc_id) {
  if (int r = check(h)) return r;
  if (h->periodic) return fail(ASLHIP_EINVAL, "ASL ERROR: map BCs need a non-periodic lattice");
  if (kind < ASLHIP_BC_NOSLIP_MAP || kind > ASLHIP_BC_VELOCITY_MAP)
    return fail(ASLHIP_EINVAL, "ASL ERROR: %d is not a map BC kind", kind);
  if (!map) return fail(ASLHIP_EINVAL, "ASL ERROR: map BC without a map");
  BCDesc b;
  b.kind = kind;
  b.p = p;
  for (int d = 0; d < 3; ++d) b.u[d] = u ? u[d] : 0.;
  if (rot) {
    b.has_rot = true;
    for (int d = 0; d < 6; ++d) b.rot[d] = rot[d];
  }
  b.map.assign(map, map + h->len);
  if (cdomain) b.cd.assign(cdomain, cdomain + h->len);
  h->bcs.push_back(std::move(b));
  if (bc_id) *bc_id = static_cast<int>(h->bcs.size()) - 1;
  h->dirty = true;
  return 0;
}

int aslhip_vis_add(aslhip_lattice *h, int kind, const float *map, const double u[3], const double rot[6],
                   int *vis_id) {
  if (int r = check(h)) return r;
  if (kind < ASLHIP_VIS_NOSLIP_VEL || kind > ASLHIP_VIS_VELOCITY)
    return fail(ASLHIP_EINVAL, "ASL ERROR: %d is not a visualisation fill kind", kind);
  if (!map) return fail(ASLHIP_EINVAL, "ASL ERROR: visualisation fill without a map");
  if (h->Q == 27 && kind != ASLHIP_VIS_VELOCITY)
    return fail(ASLHIP_EINVAL, "ASL ERROR: d3q27 has no laplace / gradient coefficients for this fill");
  aslhip_lattice::Vis v{kind, nullptr, {0., 0., 0.}, rot != nullptr, {0., 0., 0., 0., 0., 0.}};
  for (int d = 0; d < 3; ++d) v.u[d] = u ? u[d] : 0.;
  if (rot)
    for (int d = 0; d < 6; ++d) v.rot[d] = rot[d];
  const size_t bytes = (size_t)h->len * sizeof(float);
  if (hipMalloc(&v.map, bytes) != hipSuccess) return fail(ASLHIP_ENOMEM, "ASL ERROR: visualisation fill: out of memory");
  if (hipMemcpy(v.map, map, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(v.map);
    return fail(ASLHIP_EHIP, "ASL ERROR: visualisation fill: map upload failed");
  }
  h->vis.push_back(v);
  if (vis_id) *vis_id = static_cast<int>(h->vis.size()) - 1;
  return 0;
}

int aslhip_vis_execute(aslhip_lattice *h, int vis_id) {
  if (int r = check(h)) return r;
  if (vis_id < 0 || vis_id >= (int)h->vis.size())
    return fail(ASLHIP_EINVAL, "ASL ERROR: unknown visualisation fill id %d", vis_id);
  if (!h->initialized) return fail(ASLHIP_ESTATE, "ASL ERROR: visualisation fill before initF");
  // rho and u must hold the last step's values (a deferred step writes them)
  if (int r = flush_deferred(h)) return r;
  if (int r = order_after_consumer(h)) return r;
  const aslhip_lattice::Vis &v = h->vis[vis_id];
  VisArgs a;
  a.map = v.map;
  a.rho = h->rho;
  a.u = h->u;
  a.kind = v.kind;
  a.len = h->len;
  a.sx = h->sx;
  a.sy = h->sy;
  for (int d = 0; d < 3; ++d) a.u0[d] = v.u[d];
  for (int d = 0; d < 6; ++d) a.rot[d] = v.rot[d];
  a.has_rot = v.has_rot ? 1 : 0;
  a.dx = h->dx;
  for (int d = 0; d < 3; ++d) a.pos0[d] = h->pos0[d];
  HIPCHK(launch_vis(h->Q, h->dtype, a, h->stream));
  return 0;
}

int aslhip_bc_execute(aslhip_lattice *h, int bc_id) {
  if (int r = check(h)) return r;
  if (bc_id < 0 || bc_id >= (int)h->bcs.size()) return fail(ASLHIP_EINVAL, "ASL ERROR: unknown BC id %d", bc_id);
  // a BC executed twice in one cycle counts at its last position
  auto it = std::find(h->pending.begin(), h->pending.end(), bc_id);
  if (it != h->pending.end()) h->pending.erase(it);
  h->pending.push_back(bc_id);
  if (h->fresh) h->bc_since_init = true;
  return 0;
}

int aslhip_finalize(aslhip_lattice *h) {
  if (int r = check(h)) return r;
  // a deferred step keeps the order it was requested with
  if (int r = flush_deferred(h)) return r;
  if (!h->pending.empty()) {
    if (h->pending != h->order) {
      h->order = h->pending;
      h->dirty = true;
    }
    h->pending.clear();
  }
  return h->dirty ? finalize_impl(h) : 0;
}

int aslhip_step(aslhip_lattice *h, int nsteps, unsigned flags) {
  if (int r = check(h)) return r;
  return step_impl(h, nsteps, flags, nullptr);
}

int aslhip_synchronize(aslhip_lattice *h) {
  if (int r = check(h)) return r;
  if (int r = flush_deferred(h)) return r;
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int aslhip_read_rho(aslhip_lattice *h, void *out) {
  if (int r = check(h)) return r;
  if (int r = flush_deferred(h)) return r;
  return copy_out(h, h->rho, out, (size_t)h->len * h->melem);
}

int aslhip_read_velocity(aslhip_lattice *h, void *out) {
  if (int r = check(h)) return r;
  if (int r = flush_deferred(h)) return r;
  return copy_out(h, h->u, out, (size_t)h->len * 3 * h->melem);
}

int aslhip_macro_device(aslhip_lattice *h, void *stream, const void **rho, const void **u) {
  if (int r = check(h)) return r;
  if (!rho && !u) return fail(ASLHIP_EINVAL, "ASL ERROR: macro_device: no output address");
  if (int r = flush_deferred(h)) return r;
  if (!stream) {
    HIPCHK(hipStreamSynchronize(h->stream));  // readable from any stream until the next step
  } else {
    if (!h->ev_macro) HIPCHK(hipEventCreateWithFlags(&h->ev_macro, hipEventDisableTiming));
    if (!h->ev_consumer) HIPCHK(hipEventCreateWithFlags(&h->ev_consumer, hipEventDisableTiming));
    hipStream_t c = static_cast<hipStream_t>(stream);
    if (h->consumer && h->consumer != c) {
      // the earlier consumer's reads still come before the next step
      if (int r = order_after_consumer(h)) return r;
    }
    HIPCHK(hipEventRecord(h->ev_macro, h->stream));
    HIPCHK(hipStreamWaitEvent(c, h->ev_macro, 0));
    h->consumer = c;
  }
  if (rho) *rho = h->rho;
  if (u) *u = h->u;
  return 0;
}

int aslhip_read_f_noalloc(aslhip_lattice *h, void *out) {
  if (int r = check(h)) return r;
  if (!out) return fail(ASLHIP_EINVAL, "ASL ERROR: read_f_noalloc needs an output buffer");
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != h->device)
    return fail(ASLHIP_EINVAL, "ASL ERROR: read_f_noalloc: the lattice lives on device %d, device %d is current",
                h->device, dev);
  if ((size_t)h->Q * h->melem > ASLHIP_NOALLOC_SCRATCH_BYTES)
    return fail(ASLHIP_EINVAL, "ASL ERROR: read_f_noalloc: one cell's populations exceed the %u-byte staging array",
                ASLHIP_NOALLOC_SCRATCH_BYTES);
  if (h->deferred && (h->dirty || !h->pending.empty()))
    return fail(ASLHIP_ESTATE, "ASL ERROR: read_f_noalloc: the pending step needs its boundary tables rebuilt "
                               "(an allocation); call aslhip_finalize first");
  void *buf = nullptr;
  HIPCHK(noalloc_scratch(&buf));
  int r = aslhip::read_f_through(h, out, buf, ASLHIP_NOALLOC_SCRATCH_BYTES);
  (void)hipStreamSynchronize(h->stream);
  return r;
}

int aslhip_read_f(aslhip_lattice *h, void *out) {
  if (int r = check(h)) return r;
  if (!out) return fail(ASLHIP_EINVAL, "ASL ERROR: read_f needs an output buffer");
  if (int r = flush_deferred(h)) return r;
  void *d = nullptr;
  const size_t bytes = (size_t)h->len * h->Q * h->melem;
  HIPCHK(hipMalloc(&d, bytes));
  int r = aslhip::read_f_through(h, out, d, bytes);
  (void)hipStreamSynchronize(h->stream);
  HIPCHK(hipFree(d));
  return r;
}

}  // extern "C"

namespace {

// ComputeSurfaceForceMap / ComputeSurfaceFluxMap on the device: post-collision
// view, incoming view (peek), then the surface kernel.
int surface_impl(aslhip_lattice *h, const float *map, void *out, bool flux) {
  if (int r = check(h)) return r;
  if (!map || !out) return fail(ASLHIP_EINVAL, "ASL ERROR: surface force/flux needs a map and an output field");
  if (!h->initialized) return fail(ASLHIP_ESTATE, "ASL ERROR: surface force/flux before initF");
  if (is_wide(h)) return fail(ASLHIP_EINVAL, "ASL ERROR: surface force/flux on lattices over 4 GiB is not supported");
  if (int r = flush_deferred(h)) return r;
  if (int r = prepare(h, false)) return r;
  if (int r = upload_args(h)) return r;
  const size_t view = (size_t)h->len * h->Q * h->melem;
  const size_t outb = (size_t)h->len * (flux ? 1 : 3) * h->melem;
  void *post = nullptr, *in = nullptr, *dout = nullptr;
  float *dmap = nullptr;
  StepArgs *dpk = nullptr;  // the step arguments with one cell per lane, for the peek
  StepArgs pk = step_args(h, 0, 1);
  pk.flags = base_flags(h) & ~(STEP_F_CLOSURE | STEP_F_CLOSURE_MFMA);
  hipError_t e = hipMalloc(&post, view);
  if (e == hipSuccess) e = hipMalloc(&in, view);
  if (e == hipSuccess) e = hipMalloc(&dout, outb);
  if (e == hipSuccess) e = hipMalloc(&dmap, (size_t)h->len * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dpk, sizeof(StepArgs));
  int r = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(dpk, &pk, sizeof(StepArgs), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dmap, map, (size_t)h->len * sizeof(float), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemsetAsync(in, 0, view, h->stream);
  if (e == hipSuccess) {
    UnpermuteArgs u;
    u.f = h->f;
    u.qs_bytes = h->qs_bytes;
    u.sx = h->sx;
    u.sy = h->sy;
    u.len = h->len;
    u.per = h->periodic ? 1 : 0;
    u.nx = h->n[0];
    u.ny = h->n[1];
    u.nz = h->n[2];
    u.after_odd = h->next_odd ? 0 : 1;
    u.first = 0;
    u.count = h->len;
    u.out = post;
    e = launch_unpermute(h->Q, h->dtype, u, h->stream);
  }
  if (e == hipSuccess) {
    StepLaunch a;
    a.nblocks = pk.nblocks;
    a.block = pk.block;
    a.args = dpk;
    a.variant = 0;
    a.cpl = 1;
    a.tile = 0;
    e = launch_peek(h->Q, h->dtype, h->next_odd, h->periodic, a, in, h->stream);
  }
  if (e == hipSuccess) {
    SurfaceArgs s;
    s.map = dmap;
    s.post = post;
    s.in = in;
    s.out = dout;
    s.sx = h->sx;
    s.sy = h->sy;
    s.len = h->len;
    s.flux = flux ? 1 : 0;
    e = launch_surface(h->Q, h->dtype, s, h->stream);
  }
  if (e == hipSuccess) r = copy_out(h, dout, out, outb);
  else r = fail(ASLHIP_EHIP, "ASL ERROR: surface force/flux: %s", hipGetErrorString(e));
  (void)hipStreamSynchronize(h->stream);
  void *ptrs[] = {post, in, dout, dmap, dpk};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  return r;
}

}  // namespace

extern "C" {

int aslhip_surface_force(aslhip_lattice *h, const float *map, void *out) { return surface_impl(h, map, out, false); }

int aslhip_surface_flux(aslhip_lattice *h, const float *map, void *out) { return surface_impl(h, map, out, true); }

int aslhip_read_cell_types(aslhip_lattice *h, uint8_t *out) {
  if (int r = check(h)) return r;
  if (h->dirty) {
    int r = aslhip_finalize(h);
    if (r) return r;
  }
  return copy_out(h, h->code, out, (size_t)h->len);
}

int aslhip_reduce_conservation(aslhip_lattice *h, double out[5]) {
  if (int r = check(h)) return r;
  if (int r = flush_deferred(h)) return r;
  if (h->dirty) {
    int r = aslhip_finalize(h);
    if (r) return r;
  }
  ReduceArgs a;
  a.f = h->f;
  a.qs_bytes = h->qs_bytes;
  a.sx = h->sx;
  a.sy = h->sy;
  a.len = h->len;
  a.per = h->periodic ? 1 : 0;
  a.nx = h->n[0];
  a.ny = h->n[1];
  a.nz = h->n[2];
  a.after_odd = h->next_odd ? 0 : 1;
  a.code = h->code;
  a.partial = h->red_partial;
  HIPCHK(launch_reduce(h->Q, h->dtype, a, h->red_out, h->stream));
  return copy_out(h, h->red_out, out, 5 * sizeof(double));
}

int aslhip_stats(aslhip_lattice *h, int64_t out[4]) {
  if (int r = check(h)) return r;
  if (h->dirty) {
    int r = aslhip_finalize(h);
    if (r) return r;
  }
  out[0] = h->n_counted;
  out[1] = h->n_computed;
  out[2] = h->n_boundary;
  out[3] = h->n_patterns;
  return 0;
}

int aslhip_get_tuning(aslhip_lattice *h, int *block, int *remap) {
  if (int r = check(h)) return r;
  if (block) *block = step_workgroup(h);
  if (remap) *remap = h->tune_remap;
  return 0;
}

int aslhip_set_tuning(aslhip_lattice *h, int knob, int value) {
  if (int r = check(h)) return r;
  if (knob == ASLHIP_TUNE_WIDE) {
    if (value != 0 && value != 1) return fail(ASLHIP_EINVAL, "ASL ERROR: ASLHIP_TUNE_WIDE takes 0 or 1");
    if (int r = flush_deferred(h)) return r;
    if ((value != 0) != h->force_wide) drop_graphs(h);  // captured launches name the variant
    h->force_wide = value != 0;
    return 0;
  }
  if (knob == ASLHIP_TUNE_XCD_REMAP) {
    if (value < 0 || value > 4096) return fail(ASLHIP_EINVAL, "ASL ERROR: XCD run length %d outside 0..4096", value);
    h->tune_remap = value;
    return 0;
  }
  if (knob == ASLHIP_TUNE_CPL) {
    if (value != 0 && value != 1 && value != 2 && value != 4)
      return fail(ASLHIP_EINVAL, "ASL ERROR: cells per lane %d is not 0 (automatic), 1, 2 or 4", value);
    if (int r = flush_deferred(h)) return r;
    if (value != h->tune_cpl) drop_graphs(h);  // captured launches name the variant and grid
    h->tune_cpl = value;
    return 0;
  }
  if (knob == ASLHIP_TUNE_BLOCK) {
    if (value < 64 || value > STEP_BLOCK || value % 64)
      return fail(ASLHIP_EINVAL, "ASL ERROR: workgroup size %d is not a multiple of 64 in 64..%d", value, STEP_BLOCK);
    if (int r = flush_deferred(h)) return r;
    if (value != h->tune_block) drop_graphs(h);  // captured launches carry the old grid
    h->tune_block = value;
    return 0;
  }
  if (knob == ASLHIP_TUNE_LDS) {
    if (value < 0 || value > 2) return fail(ASLHIP_EINVAL, "ASL ERROR: ASLHIP_TUNE_LDS takes 0, 1 or 2");
    if (int r = flush_deferred(h)) return r;
    if (value != h->tune_lds) drop_graphs(h);  // captured launches name the variant and grid
    h->tune_lds = value;
    return 0;
  }
  return fail(ASLHIP_EINVAL, "ASL ERROR: unknown tuning knob %d", knob);
}

int aslhip_tuning_value(aslhip_lattice *h, int knob, int *value) {
  if (int r = check(h)) return r;
  if (!value) return fail(ASLHIP_EINVAL, "ASL ERROR: tuning_value needs an output");
  switch (knob) {
    case ASLHIP_TUNE_XCD_REMAP: *value = h->tune_remap; return 0;
    case ASLHIP_TUNE_BLOCK: *value = step_workgroup(h); return 0;
    case ASLHIP_TUNE_WIDE: *value = is_wide(h) ? 1 : 0; return 0;
    case ASLHIP_TUNE_CPL: *value = effective_cpl(h); return 0;
    case ASLHIP_TUNE_LDS: *value = effective_lds(h); return 0;
    default: return fail(ASLHIP_EINVAL, "ASL ERROR: unknown tuning knob %d", knob);
  }
}

int aslhip_autotune(aslhip_lattice *h, int nsteps, int *best_block, int *best_remap) {
  if (int r = check(h)) return r;
  if (nsteps < 1) return fail(ASLHIP_EINVAL, "ASL ERROR: autotune needs at least one step per candidate");
  nsteps += nsteps & 1;
  const int remaps[6] = {0, 2, 4, 8, 16, 64};
  double best = 1e300;
  int bb = h->tune_block, br = h->tune_remap, bc = h->tune_cpl, bl = effective_lds(h);
  auto trial = [&](int lds, int cpl, int remap, int blk) -> int {
    if (int r = aslhip_set_tuning(h, ASLHIP_TUNE_LDS, lds)) return r;
    if (int r = aslhip_set_tuning(h, ASLHIP_TUNE_CPL, cpl)) return r;
    if (int r = aslhip_set_tuning(h, ASLHIP_TUNE_XCD_REMAP, remap)) return r;
    if (int r = aslhip_set_tuning(h, ASLHIP_TUNE_BLOCK, blk)) return r;
    double tot = 0., ker = 0.;
    if (int r = aslhip_time_steps(h, nsteps, 0u, &tot, &ker, nullptr)) return r;
    if (ker < best) {
      best = ker;
      bb = blk;
      br = remap;
      bc = cpl;
      bl = lds;
    }
    return 0;
  };
  // LDS-staged tiles where the lattice can run them: on every step (fixed
  // workgroup), on even steps (workgroup of the odd steps' direct kernel)
  if (int r = aslhip_set_tuning(h, ASLHIP_TUNE_CPL, 0)) return r;
  if (int r = aslhip_set_tuning(h, ASLHIP_TUNE_LDS, 1)) return r;
  if (effective_lds(h)) {
    for (int remap : remaps)
      if (int r = trial(1, 0, remap, h->tune_block)) return r;
    for (int remap : remaps)
      for (int blk : {256, 512, 1024})
        if (int r = trial(2, 0, remap, blk)) return r;
  }
  // direct: the distinct cells-per-lane values this lattice can run
  if (int r = aslhip_set_tuning(h, ASLHIP_TUNE_LDS, 0)) return r;
  std::vector<int> cpls;
  for (int c : {1, 2, 4}) {
    if (int r = aslhip_set_tuning(h, ASLHIP_TUNE_CPL, c)) return r;
    if (effective_cpl(h) == c) cpls.push_back(c);
  }
  for (int cpl : cpls)
    for (int remap : remaps)
      for (int blk : {256, 512, 1024})
        if (int r = trial(0, cpl, remap, blk)) return r;
  if (int r = aslhip_set_tuning(h, ASLHIP_TUNE_LDS, bl)) return r;
  if (int r = aslhip_set_tuning(h, ASLHIP_TUNE_CPL, bc)) return r;
  if (int r = aslhip_set_tuning(h, ASLHIP_TUNE_XCD_REMAP, br)) return r;
  if (int r = aslhip_set_tuning(h, ASLHIP_TUNE_BLOCK, bb)) return r;
  if (best_block) *best_block = step_workgroup(h);
  if (best_remap) *best_remap = br;
  return 0;
}

int aslhip_time_steps(aslhip_lattice *h, int nsteps, unsigned flags, double *ms_total, double *ms_kernel,
                      double *ms_each) {
  if (int r = check(h)) return r;
  if (int r = flush_deferred(h)) return r;
  flags &= ~ASLHIP_STEP_DEFER;
  if (nsteps < 1) return fail(ASLHIP_EINVAL, "ASL ERROR: time_steps needs at least one step");
  if (h->dirty || !h->pending.empty()) {
    int r = aslhip_finalize(h);
    if (r) return r;
  }
  // replayed graphs have no events between their launches
  const bool graph = (flags & ASLHIP_STEP_GRAPH) != 0;
  std::vector<hipEvent_t> ev(graph ? 0 : 2 * (size_t)nsteps);
  for (auto &e : ev) HIPCHK(hipEventCreate(&e));
  hipEvent_t t0, t1;
  HIPCHK(hipEventCreate(&t0));
  HIPCHK(hipEventCreate(&t1));
  HIPCHK(hipEventRecord(t0, h->stream));
  int r = step_impl(h, nsteps, flags, graph ? nullptr : &ev);
  HIPCHK(hipEventRecord(t1, h->stream));
  HIPCHK(hipEventSynchronize(t1));
  if (r == 0) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, t0, t1));
    if (ms_total) *ms_total = ms;
    double ksum = 0.;
    for (int s = 0; s < (graph ? 0 : nsteps); ++s) {
      float k = 0.f;
      HIPCHK(hipEventElapsedTime(&k, ev[2 * s], ev[2 * s + 1]));
      ksum += k;
      if (ms_each) ms_each[s] = k;
    }
    if (graph && ms_each)
      for (int s = 0; s < nsteps; ++s) ms_each[s] = 0.;
    if (ms_kernel) *ms_kernel = ksum;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return r;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches of lattices
// ---------------------------------------------------------------------------
struct aslhip_batch {
  std::vector<aslhip_lattice *> lat;
  int Q = 19, dtype = ASLHIP_F32;
  bool periodic = false, wide = false;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  StepArgs *desc = nullptr;  // [2][n]: plain steps, then steps writing rho/u
  uint2 *ent = nullptr;      // per workgroup: lattice, chunk
  int nblocks = 0;
  int block = STEP_BLOCK;    // workgroup size of all lattices of the batch (direct kernel)
  int tune_lds = -1;         // LDS-staged tiles: 1 on, 0 off, -1 automatic (FP32 storage)
  // Graph chunks go group by group: each group of consecutive lattices (at
  // most group_mb MiB of populations; 0: one group) takes all the chunk's
  // steps before the next starts, so a group that fits the 256 MiB Infinity
  // Cache is stepped from cache.  Entry ranges [first, second) of each group.
  int group_mb = 0;
  std::vector<std::pair<int, int>> groups;
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // chunks starting with an even / odd step
  std::vector<StepArgs> hdesc;                    // host copy of desc (source of the async upload)
};

namespace {

// Tiles for the batch's steps: non-periodic batches under 4 GiB per lattice;
// automatic as for single lattices (FP32 storage, the batch's populations
// beyond the 256 MiB Infinity Cache).
bool batch_tile(const aslhip_batch *b) {
  long long bytes = 0;
  for (const aslhip_lattice *h : b->lat) bytes += h->qs_bytes * h->Q;
  const bool on = b->tune_lds < 0 ? b->dtype == ASLHIP_F32 && bytes > TILE_MIN_BYTES : b->tune_lds != 0;
  return on && !b->periodic && !b->wide;
}

// One lattice's step arguments in the batch's launch shape.
StepArgs batch_args(const aslhip_batch *b, const aslhip_lattice *h) {
  const bool tile = batch_tile(b);
  StepArgs a = step_args(h, tile ? STEP_TILE : b->block, 1);
  a.tile = tile ? 1 : 0;
  return a;
}

void batch_free(aslhip_batch *b) {
  for (auto &g : b->graph)
    if (g) (void)hipGraphExecDestroy(g);
  void *ptrs[] = {b->desc, b->ent};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

// One step of the lattices whose workgroup entries are [e0, e1).
int batch_launch(aslhip_batch *b, bool odd, bool macro, int e0, int e1) {
  BatchLaunch L;
  L.desc = b->desc + (macro ? b->lat.size() : 0);
  L.ent = b->ent + e0;
  L.nblocks = e1 - e0;
  L.tile = batch_tile(b) ? 1 : 0;
  L.block = L.tile ? STEP_TILE : b->block;
  HIPCHK(launch_step_batch(b->Q, b->dtype, odd, b->periodic, b->wide, L, b->stream));
  return 0;
}

// The workgroup table for the batch's workgroup size (the lattices in order).
int batch_tables(aslhip_batch *b) {
  std::vector<uint2> ent;
  b->groups.clear();
  const double budget = 1048576.0 * b->group_mb;
  size_t gfirst = 0;
  double gbytes = 0.;
  for (size_t l = 0; l < b->lat.size(); ++l) {
    const aslhip_lattice *h = b->lat[l];
    const double bytes = static_cast<double>(h->Q) * h->selem * h->len;
    if (b->group_mb > 0 && ent.size() > gfirst && gbytes + bytes > budget) {
      b->groups.emplace_back(static_cast<int>(gfirst), static_cast<int>(ent.size()));
      gfirst = ent.size();
      gbytes = 0.;
    }
    gbytes += bytes;
    const StepArgs a = batch_args(b, b->lat[l]);
    for (int c = 0; c < a.nblocks; ++c) ent.push_back(make_uint2(static_cast<unsigned>(l), static_cast<unsigned>(c)));
  }
  b->groups.emplace_back(static_cast<int>(gfirst), static_cast<int>(ent.size()));
  // replays in flight read the old tables and graphs: finish them first
  HIPCHK(hipStreamSynchronize(b->stream));
  for (auto &g : b->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  if (b->ent) HIPCHK(hipFree(b->ent));
  b->ent = nullptr;
  HIPCHK(hipMalloc(&b->ent, ent.size() * sizeof(uint2)));
  HIPCHK(hipMemcpy(b->ent, ent.data(), ent.size() * sizeof(uint2), hipMemcpyHostToDevice));
  b->nblocks = static_cast<int>(ent.size());
  return 0;
}

int batch_graph(aslhip_batch *b, bool odd) {
  hipGraphExec_t &g = b->graph[odd ? 1 : 0];
  if (g) return 0;
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int r = 0;
  for (const auto &g : b->groups) {
    bool o = odd;
    for (int s = 0; s < ASLHIP_GRAPH_CHUNK && !r; ++s, o = !o) r = batch_launch(b, o, false, g.first, g.second);
  }
  hipError_t e = hipStreamEndCapture(b->stream, &graph);
  if (r) {
    if (graph) (void)hipGraphDestroy(graph);
    return r;
  }
  if (e != hipSuccess) return fail(ASLHIP_EHIP, "ASL ERROR: graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(ASLHIP_EHIP, "ASL ERROR: graph instantiate: %s", hipGetErrorString(e));
  return 0;
}

int batch_step_impl(aslhip_batch *b, int nsteps, unsigned flags) {
  if (nsteps < 0) return fail(ASLHIP_EINVAL, "ASL ERROR: negative step count");
  if (nsteps == 0) return 0;
  // Lattices fresh from initF take their first step on their own (it may
  // read ghost links raw); all of them then share one parity.
  bool any_fresh = false;
  for (auto *h : b->lat) {
    if (!h->initialized) return fail(ASLHIP_ESTATE, "ASL ERROR: batch lattice without initF");
    if (int r = flush_deferred(h)) return r;
    any_fresh = any_fresh || h->fresh;
  }
  if (any_fresh) {
    for (auto *h : b->lat)
      if (int r = step_impl(h, 1, (nsteps == 1 && (flags & ASLHIP_STEP_WRITE_MACRO)) ? ASLHIP_STEP_WRITE_MACRO : 0u,
                            nullptr))
        return r;
    if (--nsteps == 0) return 0;
  }
  const size_t n = b->lat.size();
  // the previous upload may still be in flight from this buffer
  HIPCHK(hipStreamSynchronize(b->stream));
  std::vector<StepArgs> &desc = b->hdesc;
  desc.resize(2 * n);
  const bool odd = b->lat[0]->next_odd;
  for (size_t l = 0; l < n; ++l) {
    aslhip_lattice *h = b->lat[l];
    if (int r = prepare(h, false)) return r;
    if (h->nvl) return fail(ASLHIP_EINVAL, "ASL ERROR: velocity map BCs are not supported in batches");
    if (h->has_closure || h->turbulence)
      return fail(ASLHIP_EINVAL, "ASL ERROR: per-cell relaxation (closure, turbulence) is not supported in batches");
    if (h->next_odd != odd) return fail(ASLHIP_ESTATE, "ASL ERROR: batch lattices are at different step parities");
    if (int r = order_after_consumer(h)) return r;
    desc[l] = batch_args(b, h);
    desc[n + l] = desc[l];
    if (h->comp_rho) desc[n + l].flags |= STEP_F_RHO;
    if (h->comp_vel) desc[n + l].flags |= STEP_F_VEL;
    // work queued on the lattice's own stream comes first
    HIPCHK(hipEventRecord(b->done, h->stream));
    HIPCHK(hipStreamWaitEvent(b->stream, b->done, 0));
  }
  HIPCHK(hipMemcpyAsync(b->desc, desc.data(), desc.size() * sizeof(StepArgs), hipMemcpyHostToDevice, b->stream));
  const bool macro = (flags & ASLHIP_STEP_WRITE_MACRO) != 0;
  int left = nsteps;
  bool o = odd;
  if (flags & ASLHIP_STEP_GRAPH) {
    while (left - (macro ? 1 : 0) >= ASLHIP_GRAPH_CHUNK) {
      if (int r = batch_graph(b, o)) return r;
      HIPCHK(hipGraphLaunch(b->graph[o ? 1 : 0], b->stream));
      left -= ASLHIP_GRAPH_CHUNK;  // an even count: the parity is unchanged
    }
  }
  for (; left > 0; --left, o = !o)
    if (int r = batch_launch(b, o, macro && left == 1, 0, b->nblocks)) return r;
  for (auto *h : b->lat) h->next_odd = (nsteps % 2) ? !odd : odd;
  // the lattices' own streams see the result in order
  HIPCHK(hipEventRecord(b->done, b->stream));
  for (auto *h : b->lat) HIPCHK(hipStreamWaitEvent(h->stream, b->done, 0));
  return 0;
}

}  // namespace

extern "C" {

int aslhip_batch_create(aslhip_batch **out, aslhip_lattice *const *lattices, int n) {
  if (!out || !lattices || n < 1) return fail(ASLHIP_EINVAL, "ASL ERROR: batch needs at least one lattice");
  *out = nullptr;
  if (n > 65535) return fail(ASLHIP_EINVAL, "ASL ERROR: at most 65535 lattices per batch");
  aslhip_batch *b = new aslhip_batch();
  b->lat.assign(lattices, lattices + n);
  b->Q = lattices[0]->Q;
  b->dtype = lattices[0]->dtype;
  b->periodic = lattices[0]->periodic;
  b->wide = is_wide(lattices[0]);
  for (int l = 0; l < n; ++l) {
    aslhip_lattice *h = lattices[l];
    if (!h || h->Q != b->Q || h->dtype != b->dtype || h->periodic != b->periodic || is_wide(h) != b->wide) {
      delete b;
      return fail(ASLHIP_EINVAL, "ASL ERROR: batch lattices must share template, population type and periodicity");
    }
  }
  hipError_t e = hipStreamCreate(&b->stream);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&b->desc, 2 * (size_t)n * sizeof(StepArgs));
  if (e != hipSuccess) {
    batch_free(b);
    delete b;
    return fail(ASLHIP_EHIP, "ASL ERROR: batch set-up: %s", hipGetErrorString(e));
  }
  if (int r = batch_tables(b)) {
    batch_free(b);
    delete b;
    return r;
  }
  *out = b;
  return 0;
}

int aslhip_batch_set_tuning(aslhip_batch *b, int knob, int value) {
  if (!b) return fail(ASLHIP_EINVAL, "ASL ERROR: null batch handle");
  if (knob == ASLHIP_TUNE_GROUP_MB) {
    if (value < 0 || value > (1 << 20)) return fail(ASLHIP_EINVAL, "ASL ERROR: group size %d MiB out of 0..2^20", value);
    if (value == b->group_mb) return 0;
    b->group_mb = value;
    return batch_tables(b);
  }
  if (knob == ASLHIP_TUNE_LDS) {
    if (value != 0 && value != 1) return fail(ASLHIP_EINVAL, "ASL ERROR: batches take ASLHIP_TUNE_LDS 0 or 1");
    if (value == b->tune_lds) return 0;
    b->tune_lds = value;
    return batch_tables(b);
  }
  if (knob != ASLHIP_TUNE_BLOCK)
    return fail(ASLHIP_EINVAL, "ASL ERROR: batches tune the workgroup size, the tiles and the group size only");
  if (value < 64 || value > STEP_BLOCK || value % 64)
    return fail(ASLHIP_EINVAL, "ASL ERROR: workgroup size %d is not a multiple of 64 in 64..%d", value, STEP_BLOCK);
  if (value == b->block) return 0;
  b->block = value;
  return batch_tables(b);
}

int aslhip_batch_autotune(aslhip_batch *b, int nsteps, int *best_block) {
  if (!b) return fail(ASLHIP_EINVAL, "ASL ERROR: null batch handle");
  if (nsteps < 1) return fail(ASLHIP_EINVAL, "ASL ERROR: autotune needs at least one step per candidate");
  nsteps += nsteps & 1;
  double best = 1e300;
  int bb = b->block, bl = batch_tile(b) ? 1 : 0;
  auto trial = [&](int lds, int blk) -> int {
    if (int r = aslhip_batch_set_tuning(b, ASLHIP_TUNE_LDS, lds)) return r;
    if (int r = aslhip_batch_set_tuning(b, ASLHIP_TUNE_BLOCK, blk)) return r;
    double ms = 0.;
    if (int r = aslhip_batch_time_steps(b, nsteps, 0u, &ms)) return r;
    if (ms < best) {
      best = ms;
      bb = blk;
      bl = lds;
    }
    return 0;
  };
  if (!b->periodic && !b->wide)
    if (int r = trial(1, b->block)) return r;  // tiles: fixed workgroup
  for (int blk : {256, 512, 1024})
    if (int r = trial(0, blk)) return r;
  if (int r = aslhip_batch_set_tuning(b, ASLHIP_TUNE_LDS, bl)) return r;
  if (int r = aslhip_batch_set_tuning(b, ASLHIP_TUNE_BLOCK, bb)) return r;
  if (best_block) *best_block = bl ? STEP_TILE : bb;
  return 0;
}

int aslhip_batch_autotune_groups(aslhip_batch *b, int chunks, int *best_mb) {
  if (!b) return fail(ASLHIP_EINVAL, "ASL ERROR: null batch handle");
  if (chunks < 1) return fail(ASLHIP_EINVAL, "ASL ERROR: autotune needs at least one chunk per candidate");
  double best = 1e300;
  int bm = b->group_mb;
  for (int mb : {0, 64, 128, 160, 192, 224, 256, 384}) {
    if (int r = aslhip_batch_set_tuning(b, ASLHIP_TUNE_GROUP_MB, mb)) return r;
    // the first chunk captures the graphs; the timed ones replay them
    if (int r = batch_step_impl(b, ASLHIP_GRAPH_CHUNK, ASLHIP_STEP_GRAPH)) return r;
    double ms = 0.;
    if (int r = aslhip_batch_time_steps(b, chunks * ASLHIP_GRAPH_CHUNK, ASLHIP_STEP_GRAPH, &ms)) return r;
    if (ms < best) {
      best = ms;
      bm = mb;
    }
  }
  if (int r = aslhip_batch_set_tuning(b, ASLHIP_TUNE_GROUP_MB, bm)) return r;
  if (best_mb) *best_mb = bm;
  return 0;
}

int aslhip_batch_destroy(aslhip_batch *b) {
  if (!b) return 0;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_free(b);
  delete b;
  return 0;
}

int aslhip_batch_step(aslhip_batch *b, int nsteps, unsigned flags) {
  if (!b) return fail(ASLHIP_EINVAL, "ASL ERROR: null batch handle");
  return batch_step_impl(b, nsteps, flags);
}

int aslhip_batch_time_steps(aslhip_batch *b, int nsteps, unsigned flags, double *ms_total) {
  if (!b) return fail(ASLHIP_EINVAL, "ASL ERROR: null batch handle");
  hipEvent_t t0, t1;
  HIPCHK(hipEventCreate(&t0));
  HIPCHK(hipEventCreate(&t1));
  // start from an idle batch stream (previous work of the lattices included)
  for (auto *h : b->lat) {
    HIPCHK(hipEventRecord(b->done, h->stream));
    HIPCHK(hipStreamWaitEvent(b->stream, b->done, 0));
  }
  HIPCHK(hipEventRecord(t0, b->stream));
  int r = batch_step_impl(b, nsteps, flags);
  HIPCHK(hipEventRecord(t1, b->stream));
  HIPCHK(hipEventSynchronize(t1));
  float ms = 0.f;
  if (r == 0) HIPCHK(hipEventElapsedTime(&ms, t0, t1));
  if (ms_total) *ms_total = ms;
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return r;
}

}  // extern "C"
