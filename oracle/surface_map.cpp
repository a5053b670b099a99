// CPU restatement of the surface distance map (test infrastructure only).
//
// asl::readSurface (src/readers/aslVTKFormatReaders.h:66-78) hands the
// triangles of an STL file to VTK's vtkImplicitModeller (surfaceToData,
// src/readers/aslVTKFormatReaders.cxx:132-217): per grid node the unsigned
// distance to the nearest triangle, capped by CapValue (also the initial
// distance of every node), then ASL rescales it to its map convention,
// map = (d - shift) / div1 / div2 in the map's float type.  VTK is a
// third-party dependency absent here (any version >= 7.1, unpinned), so this
// restates the published point-triangle distance (C. Ericson, Real-Time
// Collision Detection, 2005, sec. 5.1.5): the closest point by Voronoi
// regions of vertices, edges and face; zero-area triangles use their edges.
// Each node keeps the minimum over the triangles whose node box (bounding
// box grown by the cap, in node indices) holds it of the squared distance
// evaluated in double and rounded to float; nodes outside every box keep
// cap^2 (they are farther than the cap from every triangle).
#include <algorithm>
#include <cmath>
#include <vector>

#include "lbgk_oracle.h"

namespace {

double dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

double segment_d2(const double *p, const double *a, const double *b) {
  double ab[3], ap[3], r[3];
  for (int d = 0; d < 3; ++d) {
    ab[d] = b[d] - a[d];
    ap[d] = p[d] - a[d];
  }
  const double l2 = dot(ab, ab);
  double t = l2 > 0. ? dot(ap, ab) / l2 : 0.;
  t = std::min(1., std::max(0., t));
  for (int d = 0; d < 3; ++d) r[d] = p[d] - (a[d] + t * ab[d]);
  return dot(r, r);
}

double triangle_d2(const double *p, const double *t) {
  const double *a = t, *b = t + 3, *c = t + 6;
  double ab[3], ac[3], ap[3], bp[3], cp[3];
  for (int d = 0; d < 3; ++d) {
    ab[d] = b[d] - a[d];
    ac[d] = c[d] - a[d];
    ap[d] = p[d] - a[d];
    bp[d] = p[d] - b[d];
    cp[d] = p[d] - c[d];
  }
  // (numerically) zero area, sin^2 of the angle at a below 1e-20: the
  // region tests lose their meaning and the triangle is its edges (the
  // distance changes by at most 1e-10 of the edge length)
  const double n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
  if (n[0] * n[0] + n[1] * n[1] + n[2] * n[2] <= 1e-20 * (dot(ab, ab) * dot(ac, ac)))
    return std::min(std::min(segment_d2(p, a, b), segment_d2(p, b, c)), segment_d2(p, c, a));
  const double d1 = dot(ab, ap), d2 = dot(ac, ap), d3 = dot(ab, bp), d4 = dot(ac, bp), d5 = dot(ab, cp),
               d6 = dot(ac, cp);
  const double vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
  double q[3];
  if (d1 <= 0. && d2 <= 0.) {
    std::copy(a, a + 3, q);
  } else if (d3 >= 0. && d4 <= d3) {
    std::copy(b, b + 3, q);
  } else if (vc <= 0. && d1 >= 0. && d3 <= 0.) {
    const double v = d1 / (d1 - d3);
    for (int d = 0; d < 3; ++d) q[d] = a[d] + v * ab[d];
  } else if (d6 >= 0. && d5 <= d6) {
    std::copy(c, c + 3, q);
  } else if (vb <= 0. && d2 >= 0. && d6 <= 0.) {
    const double w = d2 / (d2 - d6);
    for (int d = 0; d < 3; ++d) q[d] = a[d] + w * ac[d];
  } else if (va <= 0. && d4 - d3 >= 0. && d5 - d6 >= 0.) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int d = 0; d < 3; ++d) q[d] = b[d] + w * (c[d] - b[d]);
  } else {
    const double den = 1. / (va + vb + vc);
    const double v = vb * den, w = vc * den;
    for (int d = 0; d < 3; ++d) q[d] = a[d] + ab[d] * v + ac[d] * w;
  }
  double r[3];
  for (int d = 0; d < 3; ++d) r[d] = p[d] - q[d];
  return dot(r, r);
}

// Node index range [lo, hi] of a triangle per axis; false when empty.
bool box(const double *t, const int *size, const double *pos, const double *h, double cap, int *lo, int *hi) {
  for (int d = 0; d < 3; ++d) {
    const double mn = std::fmin(std::fmin(t[d], t[3 + d]), t[6 + d]);
    const double mx = std::fmax(std::fmax(t[d], t[3 + d]), t[6 + d]);
    const double l = std::fmax(std::ceil((mn - cap - pos[d]) / h[d]), 0.);
    const double u = std::fmin(std::floor((mx + cap - pos[d]) / h[d]), double(size[d] - 1));
    if (!(l <= u)) return false;
    lo[d] = int(l);
    hi[d] = int(u);
  }
  return true;
}

}  // namespace

extern "C" {

double orc_triangle_d2(const double *p, const double *t) { return triangle_d2(p, t); }

void orc_surface_distance_map(const double *tri, long long ntri, const int *size, const double *pos, const double *h,
                              double cap, double shift, double div1, double div2, float *map) {
  const long long n = (long long)size[0] * size[1] * size[2];
  std::vector<float> d2(n, float(cap * cap));
  for (long long k = 0; k < ntri; ++k) {
    const double *t = tri + 9 * k;
    int lo[3], hi[3];
    if (!box(t, size, pos, h, cap, lo, hi)) continue;
    for (int i = lo[0]; i <= hi[0]; ++i)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int l = lo[2]; l <= hi[2]; ++l) {
          const double p[3] = {pos[0] + h[0] * double(i), pos[1] + h[1] * double(j), pos[2] + h[2] * double(l)};
          const float f = float(triangle_d2(p, t));
          float &m = d2[((long long)i * size[1] + j) * size[2] + l];
          if (f < m) m = f;
        }
  }
  const float s = float(shift), a = float(div1), b = float(div2);
  for (long long i = 0; i < n; ++i) map[i] = ((std::sqrt(d2[i]) - s) / a) / b;
}

}  // extern "C"
