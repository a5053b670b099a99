// aslDistanceFunction.h — level-set geometry (src/math/aslDistanceFunction.h,
// src/math/aslDistanceFunction.cxx:55-164), evaluated on the host in the
// element type of the map (float for FP32 maps), with ASL's expression order.
#ifndef ASLDISTANCEFUNCTION_H
#define ASLDISTANCEFUNCTION_H

#include <memory>

#include "aslVectors.h"

namespace asl {

class DistanceFunction {
 public:
  virtual ~DistanceFunction() {}
  virtual float getDistance(const float *pos) const = 0;
  virtual double getDistance(const double *pos) const = 0;
};
typedef std::shared_ptr<DistanceFunction> SPDistanceFunction;

}  // namespace asl

#endif
