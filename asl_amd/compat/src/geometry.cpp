// Level-set geometry of the compat layer: ASL's distance functions
// (src/math/aslDistanceFunction.cxx:55-164) and generators
// (src/aslGeomInc.cxx:33-221).  Maps are built on the GPU from the tree's
// postfix program (aslhip_df_map); getDistance is the same expression at one
// point on the host (constants rounded to the element type, as in ASL's
// generated kernels).
#include <algorithm>
#include <cmath>

#include <aslGeomInc.h>
#include <aslhip.h>

namespace asl {
namespace detail {
void check_abi(int r);
}
namespace {

template <class Impl>
class DFBase : public DistanceFunction {
 public:
  float getDistance(const float *p) const override { return static_cast<const Impl *>(this)->template eval<float>(p); }
  double getDistance(const double *p) const override {
    return static_cast<const Impl *>(this)->template eval<double>(p);
  }
};

class Sphere : public DFBase<Sphere> {
 public:
  Sphere(double r, const AVec<double> &c) : r_(r), c_{c[0], c[1], c[2]} {}
  template <class T>
  T eval(const T *p) const {
    const T d0 = p[0] - T(c_[0]), d1 = p[1] - T(c_[1]), d2 = p[2] - T(c_[2]);
    const T l2 = (d0 * d0 + d1 * d1) + d2 * d2;
    return std::sqrt(l2) - T(r_);
  }
  void emit(std::vector<double> &p) const override { p.insert(p.end(), {ASLHIP_DF_SPHERE, c_[0], c_[1], c_[2], r_}); }

 private:
  double r_, c_[3];
};

// DistFCylinder / DistFCone: distance from the axis through c along o.
class Cylinder : public DFBase<Cylinder> {
 public:
  Cylinder(double r, const AVec<double> &l, const AVec<double> &c, bool cone, double tanth)
      : r_(r), cone_(cone), tan_(tanth), c_{c[0], c[1], c[2]} {
    const AVec<double> o = normalize(l);
    for (int d = 0; d < 3; ++d) o_[d] = o[d];
  }
  template <class T>
  T eval(const T *p) const {
    const T d0 = p[0] - T(c_[0]), d1 = p[1] - T(c_[1]), d2 = p[2] - T(c_[2]);
    const T proj = (d0 * T(o_[0]) + d1 * T(o_[1])) + d2 * T(o_[2]);
    const T e0 = d0 - proj * T(o_[0]), e1 = d1 - proj * T(o_[1]), e2 = d2 - proj * T(o_[2]);
    const T rad = std::sqrt((e0 * e0 + e1 * e1) + e2 * e2);
    return cone_ ? rad - T(tan_) * proj : rad - T(r_);
  }
  void emit(std::vector<double> &p) const override {
    p.insert(p.end(), {double(cone_ ? ASLHIP_DF_CONE : ASLHIP_DF_CYLINDER), c_[0], c_[1], c_[2], o_[0], o_[1], o_[2],
                       cone_ ? tan_ : r_});
  }

 private:
  double r_;
  bool cone_;
  double tan_, c_[3], o_[3];
};

class Plane : public DFBase<Plane> {
 public:
  Plane(const AVec<double> &n, const AVec<double> &p0) {
    const AVec<double> nn = normalize(n);
    for (int d = 0; d < 3; ++d) n_[d] = nn[d];
    b_ = (p0 * n) / std::sqrt(l2(n));
  }
  template <class T>
  T eval(const T *p) const {
    return ((p[0] * T(n_[0]) + p[1] * T(n_[1])) + p[2] * T(n_[2])) - T(b_);
  }
  void emit(std::vector<double> &p) const override { p.insert(p.end(), {ASLHIP_DF_PLANE, n_[0], n_[1], n_[2], b_}); }

 private:
  double n_[3], b_;
};

class Binary : public DFBase<Binary> {
 public:
  Binary(SPDistanceFunction a, SPDistanceFunction b, bool is_and) : a_(a), b_(b), and_(is_and) {}
  template <class T>
  T eval(const T *p) const {
    const T x = a_->getDistance(p), y = b_->getDistance(p);
    return and_ ? std::max(x, y) : std::min(x, y);
  }
  void emit(std::vector<double> &p) const override {
    a_->emit(p);
    b_->emit(p);
    p.push_back(and_ ? ASLHIP_DF_AND : ASLHIP_DF_OR);
  }

 private:
  SPDistanceFunction a_, b_;
  bool and_;
};

class Inversion : public DFBase<Inversion> {
 public:
  explicit Inversion(SPDistanceFunction a) : a_(a) {}
  template <class T>
  T eval(const T *p) const { return -a_->getDistance(p); }
  void emit(std::vector<double> &p) const override {
    a_->emit(p);
    p.push_back(ASLHIP_DF_NEG);
  }

 private:
  SPDistanceFunction a_;
};

// DistFNormalization: clamp(d / (1.8 dx), -1, 1), factor rounded to T.
class Normalization : public DFBase<Normalization> {
 public:
  Normalization(SPDistanceFunction a, double dx) : a_(a), f_(1. / dx / 1.8) {}
  template <class T>
  T eval(const T *p) const {
    const T dist = T(f_) * a_->getDistance(p);
    const T top = dist > T(1.) ? T(1.) : dist;
    return dist < T(-1.) ? T(-1.) : top;
  }
  void emit(std::vector<double> &p) const override {
    a_->emit(p);
    p.insert(p.end(), {ASLHIP_DF_NORMALIZE, f_});
  }

 private:
  SPDistanceFunction a_;
  double f_;
};

// The map of df on the container's ghost-inclusive block, on the GPU.
void sample(AbstractData &d, const DistanceFunction &df) {
  const Block &b = d.getBlock();
  if (nD(b) != 3) errorMessage("initData: a 3D block is expected");
  std::vector<double> prog;
  df.emit(prog);
  const int size[3] = {b.getSize()[0], b.getSize()[1], b.getSize()[2]};
  const double pos[3] = {b.position[0], b.position[1], b.position[2]};
  std::vector<double> &out = d.values_.at(0);
  if (d.getTypeID() == acl::TYPE_DOUBLE) {
    detail::check_abi(aslhip_df_map(prog.data(), int(prog.size()), size, pos, b.dx, 1, out.data()));
  } else {
    std::vector<float> m(out.size());
    detail::check_abi(aslhip_df_map(prog.data(), int(prog.size()), size, pos, b.dx, 0, m.data()));
    std::copy(m.begin(), m.end(), out.begin());
  }
}

}  // namespace

SPDistanceFunction generateDFSphere(double r, const AVec<double> &c) { return std::make_shared<Sphere>(r, c); }

SPDistanceFunction generateDFCylinderInf(double r, const AVec<double> &l, const AVec<double> &c) {
  return std::make_shared<Cylinder>(r, l, c, false, 0.);
}

SPDistanceFunction generateDFCylinder(double r, const AVec<double> &l, const AVec<double> &c) {
  return generateDFCylinderInf(r, l, c) & generateDFPlane(l, c + l * .5) & generateDFPlane(-l, c - l * .5);
}

SPDistanceFunction generateDFCone(double r, const AVec<double> &l, const AVec<double> &a) {
  // DistFCone(r/|l|, l, a) with tanTheta = tan(r/|l|), cut by the planes at
  // the apex and the base (src/aslGeomInc.cxx:108-116,
  // src/math/aslDistanceFunction.cxx:118-131)
  const double th = r / std::sqrt(l2(l));
  return std::make_shared<Cylinder>(0., l, a, true, std::tan(th)) & generateDFPlane(l, a + l) &
         generateDFPlane(-l, a);
}

SPDistanceFunction generateDFPlane(const AVec<double> &n, const AVec<double> &p0) {
  return std::make_shared<Plane>(n, p0);
}

SPDistanceFunction generateDFConvexPolygonPrism(std::vector<AVec<double>> points) {
  AVec<double> center(points[0]);
  for (unsigned int i = 1; i < points.size(); ++i) center += points[i];
  center = center / double(points.size());
  const AVec<double> axis = crossProduct(points[0] - center, points[1] - center);
  SPDistanceFunction res = generateDFPlane(crossProduct(points[1] - points[0], axis), points[0]);
  for (unsigned int i = 1; i < points.size(); ++i)
    res = res & generateDFPlane(crossProduct(points[(i + 1) % points.size()] - points[i], axis), points[i]);
  return res;
}

SPDistanceFunction generateDFConvexPolygonPyramid(std::vector<AVec<double>> points, AVec<double> a) {
  AVec<double> center(points[0]);
  for (unsigned int i = 1; i < points.size(); ++i) center += points[i];
  center = center / double(points.size());
  const AVec<double> axis = crossProduct(points[0] - center, points[1] - center);
  SPDistanceFunction res = generateDFPlane(crossProduct(points[1] - points[0], a - points[0]), points[0]);
  for (unsigned int i = 1; i < points.size(); ++i)
    res = res & generateDFPlane(crossProduct(points[(i + 1) % points.size()] - points[i], a - points[i]), points[i]);
  return res & generateDFPlane(-axis, center);
}

SPDistanceFunction generateDFInBlock(const Block &b, unsigned int nG) {
  const unsigned int nd = nD(b);
  AVec<> n(nd, 0.), nc(n);
  nc[0] = 1.;
  const AVec<> offs(nd, (double(nG) - .5) * b.dx);
  const AVec<> p0(b.position + offs), pE(b.getBPosition() - offs);
  SPDistanceFunction res = generateDFPlane(nc, p0) | generateDFPlane(-nc, pE);
  for (unsigned int d = 1; d < nd; ++d) {
    nc = n;
    nc[d] = 1.;
    res = res | generateDFPlane(nc, p0) | generateDFPlane(-nc, pE);
  }
  return res;
}

SPDistanceFunction operator&(SPDistanceFunction a, SPDistanceFunction b) {
  if (!a && !b) errorMessage("DistanceFunction::operator& - both operands are not initialized");
  if (!a) return b;
  if (!b) return a;
  return std::make_shared<Binary>(a, b, true);
}

SPDistanceFunction operator|(SPDistanceFunction a, SPDistanceFunction b) {
  if (!a && !b) errorMessage("DistanceFunction::operator| - both operands are not initialized");
  if (!a) return b;
  if (!b) return a;
  return std::make_shared<Binary>(a, b, false);
}

SPDistanceFunction operator-(SPDistanceFunction a) {
  if (!a) errorMessage("DistanceFunction::operator- - operand is not initialized");
  return std::make_shared<Inversion>(a);
}

SPDistanceFunction normalize(SPDistanceFunction a, double dx) {
  if (!a) errorMessage("DistanceFunction::normalize - argument is not initialized");
  return std::make_shared<Normalization>(a, dx);
}

void initData(SPAbstractDataWithGhostNodes d, SPDistanceFunction f) {
  if (!f) errorMessage("DistanceFunction::initData - argument is not initialized");
  sample(*d, *f);
}

SPAbstractDataWithGhostNodes generateDataContainer_SP(const Block &b, SPDistanceFunction df, unsigned int gN,
                                                      acl::TypeID t) {
  auto d = std::make_shared<DataWithGhostNodesACLData>(b, 1, t, int(gN));
  initData(d, df);
  return d;
}

}  // namespace asl
