// readSurface over the GPU distance map (src/readers/aslVTKFormatReaders.cxx:
// 132-217 and 266-367).  File coordinates (x, y, z) become ASL's axes
// (2, 1, 0): the modeller's image has x fastest and makeData reads it with
// the last ASL index fastest (src/utilities/aslVTKCasters.cxx:509-531).
#include <readers/aslVTKFormatReaders.h>

#include <algorithm>
#include <cmath>

#include <aslhip.h>

namespace asl {
namespace detail {
void check_abi(int r);
}

namespace {

using detail::check_abi;

std::vector<double> load_surface(const std::string &fileName) {
  const size_t dot = fileName.rfind('.');
  const std::string ext = dot == std::string::npos ? "" : fileName.substr(dot);
  if (ext != ".stl" && ext != ".vtp") errorMessage("Reader: file format not supported");
  const bool vtp = ext == ".vtp";
  long long n = 0;
  check_abi(vtp ? aslhip_vtp_count(fileName.c_str(), &n) : aslhip_stl_count(fileName.c_str(), &n));
  std::vector<double> tri(9 * n);
  if (n) check_abi(vtp ? aslhip_vtp_read(fileName.c_str(), tri.data(), n) : aslhip_stl_read(fileName.c_str(), tri.data(), n));
  for (long long k = 0; k < 3 * n; ++k) std::swap(tri[3 * k], tri[3 * k + 2]);
  return tri;
}

// The map on the grid of size nodes at pos + h * index (ASL axis order).
SPDataWithGhostNodesACLData sample(const std::vector<double> &tri, const int size[3], const double pos[3],
                                   const double h[3], double blockDx, double cap, double shift, double div1,
                                   double div2) {
  for (int d = 0; d < 3; ++d)
    if (size[d] < 3) errorMessage("readSurface: the sampled grid needs at least 3 nodes per axis");
  const Block grid(makeAVec(size[0], size[1], size[2]), blockDx, makeAVec(pos[0], pos[1], pos[2]));
  auto data = std::make_shared<DataWithGhostNodesACLData>(offset(grid, -1), 1, acl::TYPE_FLOAT, 1);
  std::vector<float> m(static_cast<size_t>(size[0]) * size[1] * size[2]);
  check_abi(aslhip_surface_distance_map(tri.data(), static_cast<long long>(tri.size() / 9), size, pos, h, cap, shift,
                                        div1, div2, m.data()));
  std::copy(m.begin(), m.end(), data->values_[0].begin());
  return data;
}

// vtkImplicitModeller::ComputeModelBounds with AdjustBounds on (its default):
// the points' box grown by 0.0125 of its largest side.
void model_bounds(const std::vector<double> &tri, double lo[3], double hi[3]) {
  if (tri.empty()) errorMessage("readSurface: the surface has no triangles");
  for (int d = 0; d < 3; ++d) {
    lo[d] = hi[d] = tri[d];
    for (size_t k = d; k < tri.size(); k += 3) {
      lo[d] = std::min(lo[d], tri[k]);
      hi[d] = std::max(hi[d], tri[k]);
    }
  }
  double m = 0.;
  for (int d = 0; d < 3; ++d) m = std::max(m, hi[d] - lo[d]);
  for (int d = 0; d < 3; ++d) {
    lo[d] -= m * 0.0125;
    hi[d] += m * 0.0125;
  }
}

// Sample counts (bounds / dx, truncated) and spacings of a box; the block's
// dx is the spacing along the file's x (ASL axis 2).
SPDataWithGhostNodesACLData sample_box(const std::vector<double> &tri, const double lo[3], const double hi[3],
                                       double dx) {
  int size[3];
  double h[3];
  for (int d = 0; d < 3; ++d) {
    size[d] = static_cast<int>((hi[d] - lo[d]) / dx);
    if (size[d] < 3) errorMessage("readSurface: the surface spans fewer than 3 nodes along an axis");
    h[d] = (hi[d] - lo[d]) / (size[d] - 1);
  }
  return sample(tri, size, lo, h, h[2], 3.6 * dx, 1.8 * dx, dx, 1.8);
}

}  // namespace

SPDataWithGhostNodesACLData readSurface(const std::string &fileName, double dx, acl::CommandQueue) {
  const auto tri = load_surface(fileName);
  double lo[3], hi[3];
  model_bounds(tri, lo, hi);
  return sample_box(tri, lo, hi, dx);
}

SPDataWithGhostNodesACLData readSurface(const std::string &fileName, Block &b, acl::CommandQueue) {
  if (nD(b) != 3) errorMessage("readSurface: a 3D block is expected");
  const auto tri = load_surface(fileName);
  const auto bp = b.getBPosition();
  int size[3];
  double pos[3], h[3];
  for (int d = 0; d < 3; ++d) {
    size[d] = b.getSize()[d];
    pos[d] = b.position[d];
    h[d] = (bp[d] - pos[d]) / (size[d] - 1);
  }
  return sample(tri, size, pos, h, b.dx, 3.6 * b.dx, 1.2 * b.dx, b.dx, 1.8);
}

SPDataWithGhostNodesACLData readSurface(const std::string &fileName, double dx, double offset_X0, double offset_XE,
                                        double offset_Y0, double offset_YE, double offset_Z0, double offset_ZE,
                                        acl::CommandQueue) {
  const auto tri = load_surface(fileName);
  double lo[3], hi[3];
  model_bounds(tri, lo, hi);
  const double o0[3] = {offset_X0, offset_Y0, offset_Z0}, o1[3] = {offset_XE, offset_YE, offset_ZE};
  double l[3], u[3];
  for (int d = 0; d < 3; ++d) {
    l[d] = lo[d] - (hi[d] - lo[d]) * o0[d];
    u[d] = hi[d] + (hi[d] - lo[d]) * o1[d];
  }
  return sample_box(tri, l, u, dx);
}

}  // namespace asl
