// Level-set maps of triangulated surfaces: the geometry input of
// asl::readSurface (src/readers/aslVTKFormatReaders.h:66-78), whose
// surfaceToData (src/readers/aslVTKFormatReaders.cxx:132-217) samples the
// unsigned distance to the surface on the nodes of a grid (a distance field
// started at the cap value, so that no node is farther than the cap) and
// rescales it into ASL's map convention, map <= 0 within the shell around
// the surface.
//
// On the GPU the work goes by triangle, not by node: the host cuts each
// triangle's node box (its bounding box grown by the cap) into pieces of
// DIST_PIECE nodes; a workgroup takes one piece, evaluates the exact squared
// point-triangle distance in double for each node, rounds it to float and
// keeps the minimum with a 32-bit atomic min on the float's bits (valid for
// non-negative floats).  Nodes outside every box stay at the cap.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aslhip.h"
#include "status.hpp"

namespace aslhip {
namespace {

constexpr int DIST_BLOCK = 256;
constexpr long long DIST_PIECE = 16384;  // nodes per workgroup (64 per lane)

struct TriBox {
  double v[9];  // vertices a, b, c
  int lo[3];    // node box, inclusive
  int n[3];     // nodes per axis in the box
};

struct Piece {
  int tri;
  int first;  // first node of the box (box-linear, last axis fastest)
};

__host__ __device__ inline double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Squared distance from p to the segment a-b (a point when a == b).
__host__ __device__ inline double seg_d2(const double *p, const double *a, const double *b) {
  double ab[3], ap[3];
  for (int d = 0; d < 3; ++d) {
    ab[d] = b[d] - a[d];
    ap[d] = p[d] - a[d];
  }
  const double l2 = dot3(ab, ab);
  double t = l2 > 0. ? dot3(ap, ab) / l2 : 0.;
  t = t < 0. ? 0. : (t > 1. ? 1. : t);
  double r[3];
  for (int d = 0; d < 3; ++d) r[d] = p[d] - (a[d] + t * ab[d]);
  return dot3(r, r);
}

// Squared distance from p to the triangle (a, b, c): the closest point by
// Voronoi regions of the vertices, edges and face (C. Ericson, Real-Time
// Collision Detection, 2005, sec. 5.1.5).  Triangles whose angle at a has
// sin^2 <= 1e-20 are taken as their three edges: for such slivers the region
// tests lose their meaning, and the edges are within 1e-10 edge lengths.
__host__ __device__ inline double tri_d2(const double *p, const double *t) {
  const double *a = t, *b = t + 3, *c = t + 6;
  double ab[3], ac[3], ap[3], bp[3], cp[3];
  for (int d = 0; d < 3; ++d) {
    ab[d] = b[d] - a[d];
    ac[d] = c[d] - a[d];
    ap[d] = p[d] - a[d];
    bp[d] = p[d] - b[d];
    cp[d] = p[d] - c[d];
  }
  const double nx = ab[1] * ac[2] - ab[2] * ac[1], ny = ab[2] * ac[0] - ab[0] * ac[2],
               nz = ab[0] * ac[1] - ab[1] * ac[0];
  if (nx * nx + ny * ny + nz * nz <= 1e-20 * (dot3(ab, ab) * dot3(ac, ac))) {
    const double e0 = seg_d2(p, a, b), e1 = seg_d2(p, b, c), e2 = seg_d2(p, c, a);
    const double m = e0 < e1 ? e0 : e1;
    return m < e2 ? m : e2;
  }
  double q[3];
  const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  const double vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
  if (d1 <= 0. && d2 <= 0.) {  // vertex a
    for (int d = 0; d < 3; ++d) q[d] = a[d];
  } else if (d3 >= 0. && d4 <= d3) {  // vertex b
    for (int d = 0; d < 3; ++d) q[d] = b[d];
  } else if (vc <= 0. && d1 >= 0. && d3 <= 0.) {  // edge ab
    const double v = d1 / (d1 - d3);
    for (int d = 0; d < 3; ++d) q[d] = a[d] + v * ab[d];
  } else if (d6 >= 0. && d5 <= d6) {  // vertex c
    for (int d = 0; d < 3; ++d) q[d] = c[d];
  } else if (vb <= 0. && d2 >= 0. && d6 <= 0.) {  // edge ac
    const double w = d2 / (d2 - d6);
    for (int d = 0; d < 3; ++d) q[d] = a[d] + w * ac[d];
  } else if (va <= 0. && d4 - d3 >= 0. && d5 - d6 >= 0.) {  // edge bc
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int d = 0; d < 3; ++d) q[d] = b[d] + w * (c[d] - b[d]);
  } else {  // face
    const double den = 1. / (va + vb + vc);
    const double v = vb * den, w = vc * den;
    for (int d = 0; d < 3; ++d) q[d] = a[d] + ab[d] * v + ac[d] * w;
  }
  double r[3];
  for (int d = 0; d < 3; ++d) r[d] = p[d] - q[d];
  return dot3(r, r);
}

__global__ __launch_bounds__(DIST_BLOCK) void k_dist_fill(unsigned *d2, long long n, unsigned v) {
  const long long i = static_cast<long long>(blockIdx.x) * DIST_BLOCK + threadIdx.x;
  if (i < n) d2[i] = v;
}

__global__ __launch_bounds__(DIST_BLOCK) void k_dist_tri(const TriBox *tb, const Piece *pc, unsigned *d2, int sy,
                                                       int sz, double p0, double p1, double p2, double h0, double h1,
                                                       double h2) {
  const Piece P = pc[blockIdx.x];
  const TriBox &T = tb[P.tri];
  double t[9];
  for (int k = 0; k < 9; ++k) t[k] = T.v[k];
  const long long count = static_cast<long long>(T.n[0]) * T.n[1] * T.n[2];
  long long end = P.first + DIST_PIECE;
  if (end > count) end = count;
  for (long long m = P.first + threadIdx.x; m < end; m += DIST_BLOCK) {
    const int k = static_cast<int>(m % T.n[2]);
    const long long r = m / T.n[2];
    const int j = static_cast<int>(r % T.n[1]);
    const int i = static_cast<int>(r / T.n[1]);
    const int gi = T.lo[0] + i, gj = T.lo[1] + j, gk = T.lo[2] + k;
    const double p[3] = {p0 + h0 * static_cast<double>(gi), p1 + h1 * static_cast<double>(gj),
                         p2 + h2 * static_cast<double>(gk)};
    const float f = static_cast<float>(tri_d2(p, t));
    const long long idx = (static_cast<long long>(gi) * sy + gj) * sz + gk;
    atomicMin(d2 + idx, __float_as_uint(f));
  }
}

__global__ __launch_bounds__(DIST_BLOCK) void k_dist_map(const unsigned *d2, float *map, long long n, float shift,
                                                       float div1, float div2) {
  const long long i = static_cast<long long>(blockIdx.x) * DIST_BLOCK + threadIdx.x;
  if (i < n) map[i] = ((sqrtf(__uint_as_float(d2[i])) - shift) / div1) / div2;
}

// The node box of a triangle: nodes within cap of its bounding box per
// axis, clamped to the grid; false when empty.
bool node_box(const double *v, const int size[3], const double pos[3], const double h[3], double cap, int lo[3],
              int n[3]) {
  for (int d = 0; d < 3; ++d) {
    double mn = v[d], mx = v[d];
    for (int k = 1; k < 3; ++k) {
      mn = std::fmin(mn, v[3 * k + d]);
      mx = std::fmax(mx, v[3 * k + d]);
    }
    const double a = std::ceil((mn - cap - pos[d]) / h[d]), b = std::floor((mx + cap - pos[d]) / h[d]);
    const double l = std::fmax(a, 0.), u = std::fmin(b, static_cast<double>(size[d] - 1));
    if (!(l <= u)) return false;
    lo[d] = static_cast<int>(l);
    n[d] = static_cast<int>(u) - lo[d] + 1;
  }
  return true;
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

struct StreamHold {
  hipStream_t s = nullptr;
  ~StreamHold() {
    if (s) (void)hipStreamDestroy(s);
  }
};

// STL (binary or ASCII, src/readers/aslVTKFormatReaders.cxx:266-305 reads it
// with VTK's reader, points in float): triangles as 9 doubles each, the
// coordinates rounded to float.  tri == nullptr: count only.
int stl_parse(const char *path, double *tri, long long cap_tri, long long *ntri) {
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return fail(ASLHIP_EINVAL, "ASL ERROR: readSurface: cannot open %s", path);
  std::fseek(fp, 0, SEEK_END);
  const long size = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  std::vector<char> buf(static_cast<size_t>(size > 0 ? size : 0));
  const size_t got = size > 0 ? std::fread(buf.data(), 1, buf.size(), fp) : 0;
  std::fclose(fp);
  if (got != buf.size()) return fail(ASLHIP_EINVAL, "ASL ERROR: readSurface: cannot read %s", path);
  // binary: 80-byte header, uint32 count, 50 bytes per triangle
  if (buf.size() >= 84) {
    unsigned cnt;
    std::memcpy(&cnt, buf.data() + 80, 4);
    if (84 + 50ull * cnt == buf.size()) {
      *ntri = cnt;
      if (!tri) return 0;
      if (cap_tri < static_cast<long long>(cnt)) return fail(ASLHIP_EINVAL, "ASL ERROR: readSurface: buffer too small");
      for (unsigned t = 0; t < cnt; ++t)
        for (int k = 0; k < 9; ++k) {
          float f;
          std::memcpy(&f, buf.data() + 84 + 50ull * t + 12 + 4 * k, 4);  // after the normal
          tri[9ll * t + k] = f;
        }
      return 0;
    }
  }
  // ASCII: "vertex x y z" lines, three per facet
  std::string s(buf.begin(), buf.end());
  long long nv = 0;
  size_t at = 0;
  while ((at = s.find("vertex", at)) != std::string::npos) {
    at += 6;
    const char *c = s.c_str() + at;
    char *e = nullptr;
    double x[3];
    for (int d = 0; d < 3; ++d) {
      x[d] = std::strtod(c, &e);
      if (e == c) return fail(ASLHIP_EINVAL, "ASL ERROR: readSurface: %s is not an STL file", path);
      c = e;
    }
    if (tri) {
      if (nv / 3 >= cap_tri) return fail(ASLHIP_EINVAL, "ASL ERROR: readSurface: buffer too small");
      for (int d = 0; d < 3; ++d) tri[3 * nv + d] = static_cast<float>(x[d]);
    }
    ++nv;
  }
  if (nv == 0 || nv % 3) return fail(ASLHIP_EINVAL, "ASL ERROR: readSurface: %s is not an STL file", path);
  *ntri = nv / 3;
  return 0;
}

}  // namespace
}  // namespace aslhip

using namespace aslhip;

extern "C" {

int aslhip_stl_count(const char *path, long long *ntri) {
  if (!path || !ntri) return fail(ASLHIP_EINVAL, "ASL ERROR: readSurface: null argument");
  return stl_parse(path, nullptr, 0, ntri);
}

int aslhip_stl_read(const char *path, double *tri, long long ntri) {
  if (!path || !tri) return fail(ASLHIP_EINVAL, "ASL ERROR: readSurface: null argument");
  long long n = 0;
  if (int r = stl_parse(path, tri, ntri, &n)) return r;
  if (n != ntri) return fail(ASLHIP_EINVAL, "ASL ERROR: readSurface: %s holds %lld triangles, not %lld", path, n, ntri);
  return 0;
}

int aslhip_surface_distance_map(const double *tri, long long ntri, const int size[3], const double pos[3],
                                const double h[3], double cap, double shift, double div1, double div2, float *map) {
  if (!size || !pos || !h || !map || (ntri > 0 && !tri) || ntri < 0)
    return fail(ASLHIP_EINVAL, "ASL ERROR: surface map: null argument");
  for (int d = 0; d < 3; ++d)
    if (size[d] < 1 || !(h[d] > 0.) || !std::isfinite(pos[d]))
      return fail(ASLHIP_EINVAL, "ASL ERROR: surface map: bad grid");
  if (!(cap > 0.) || !std::isfinite(cap) || !std::isfinite(shift) || div1 == 0. || div2 == 0.)
    return fail(ASLHIP_EINVAL, "ASL ERROR: surface map: bad cap or scaling");
  for (long long k = 0; k < 9 * ntri; ++k)
    if (!std::isfinite(tri[k])) return fail(ASLHIP_EINVAL, "ASL ERROR: surface map: non-finite vertex");
  const long long n = static_cast<long long>(size[0]) * size[1] * size[2];
  std::vector<TriBox> boxes;
  std::vector<Piece> pieces;
  for (long long t = 0; t < ntri; ++t) {
    TriBox b;
    for (int k = 0; k < 9; ++k) b.v[k] = tri[9 * t + k];
    if (!node_box(b.v, size, pos, h, cap, b.lo, b.n)) continue;
    const long long cnt = static_cast<long long>(b.n[0]) * b.n[1] * b.n[2];
    if (cnt > 0x7fffffffll) return fail(ASLHIP_EINVAL, "ASL ERROR: surface map: triangle box too large");
    for (long long f = 0; f < cnt; f += DIST_PIECE)
      pieces.push_back(Piece{static_cast<int>(boxes.size()), static_cast<int>(f)});
    boxes.push_back(b);
  }
  StreamHold st;
  HIPCHK(hipStreamCreate(&st.s));
  DevBuf d2, dmap, dbox, dpc;
  HIPCHK(hipMalloc(&d2.p, n * sizeof(unsigned)));
  HIPCHK(hipMalloc(&dmap.p, n * sizeof(float)));
  const float cap2 = static_cast<float>(cap * cap);
  unsigned cap2bits;
  std::memcpy(&cap2bits, &cap2, 4);
  const unsigned nb = static_cast<unsigned>((n + DIST_BLOCK - 1) / DIST_BLOCK);
  hipLaunchKernelGGL(k_dist_fill, dim3(nb), dim3(DIST_BLOCK), 0, st.s, static_cast<unsigned *>(d2.p), n, cap2bits);
  HIPCHK(hipGetLastError());
  if (!pieces.empty()) {
    HIPCHK(hipMalloc(&dbox.p, boxes.size() * sizeof(TriBox)));
    HIPCHK(hipMalloc(&dpc.p, pieces.size() * sizeof(Piece)));
    HIPCHK(hipMemcpyAsync(dbox.p, boxes.data(), boxes.size() * sizeof(TriBox), hipMemcpyHostToDevice, st.s));
    HIPCHK(hipMemcpyAsync(dpc.p, pieces.data(), pieces.size() * sizeof(Piece), hipMemcpyHostToDevice, st.s));
    for (size_t first = 0; first < pieces.size(); first += 0x7fffffffu) {
      const size_t cnt = std::min<size_t>(pieces.size() - first, 0x7fffffffu);
      hipLaunchKernelGGL(k_dist_tri, dim3(static_cast<unsigned>(cnt)), dim3(DIST_BLOCK), 0, st.s,
                         static_cast<const TriBox *>(dbox.p), static_cast<const Piece *>(dpc.p) + first,
                         static_cast<unsigned *>(d2.p), size[1], size[2], pos[0], pos[1], pos[2], h[0], h[1], h[2]);
      HIPCHK(hipGetLastError());
    }
  }
  hipLaunchKernelGGL(k_dist_map, dim3(nb), dim3(DIST_BLOCK), 0, st.s, static_cast<const unsigned *>(d2.p),
                     static_cast<float *>(dmap.p), n, static_cast<float>(shift), static_cast<float>(div1),
                     static_cast<float>(div2));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(map, dmap.p, n * sizeof(float), hipMemcpyDeviceToHost, st.s));
  HIPCHK(hipStreamSynchronize(st.s));
  return 0;
}

}  // extern "C"
