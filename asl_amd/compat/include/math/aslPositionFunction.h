// aslPositionFunction.h — wall velocity fields of velocity map BCs
// (src/math/aslPositionFunction.h:48-170).  The GPU path supports constant
// fields and rotation fields omega x (r - c).
#ifndef ASLPOSITIONFUNCTION_H
#define ASLPOSITIONFUNCTION_H

#include <memory>

#include "aslVectors.h"

namespace asl {

class PositionFunction {
 public:
  virtual ~PositionFunction() {}
  AVec<double> constant;        // PFConstant value
  AVec<double> axis, centre;    // PFRotationField
  bool rotation = false;
};
typedef std::shared_ptr<PositionFunction> SPPositionFunction;

SPPositionFunction generatePFConstant(const AVec<double> &v);
SPPositionFunction generatePFRotationField(const AVec<double> &axis, const AVec<double> &c);

}  // namespace asl

#endif
