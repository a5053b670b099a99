// Wall-velocity fields of the compat layer (src/math/aslPositionFunction.cxx).
#include <math/aslPositionFunction.h>

namespace asl {

SPPositionFunction generatePFConstant(const AVec<double> &v) {
  auto p = std::make_shared<PositionFunction>();
  p->constant = v;
  return p;
}

SPPositionFunction generatePFRotationField(const AVec<double> &axis, const AVec<double> &c) {
  auto p = std::make_shared<PositionFunction>();
  p->rotation = true;
  p->axis = axis;
  p->centre = c;
  p->constant = AVec<double>(3, 0.);
  return p;
}

}  // namespace asl
