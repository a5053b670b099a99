// aslDistanceFunction.h — level-set geometry (src/math/aslDistanceFunction.h,
// src/math/aslDistanceFunction.cxx:55-164).  A tree emits itself as a postfix
// program (emit) that aslhip_df_map evaluates on the GPU at every node of a
// map; getDistance evaluates the same expression at one point on the host,
// in the element type of the map, with ASL's expression order.
#ifndef ASLDISTANCEFUNCTION_H
#define ASLDISTANCEFUNCTION_H

#include <memory>
#include <vector>

#include "aslVectors.h"

namespace asl {

class DistanceFunction {
 public:
  virtual ~DistanceFunction() {}
  virtual float getDistance(const float *pos) const = 0;
  virtual double getDistance(const double *pos) const = 0;
  // appends the postfix program of this expression (ASLHIP_DF_* op codes)
  virtual void emit(std::vector<double> &prog) const = 0;
};
typedef std::shared_ptr<DistanceFunction> SPDistanceFunction;

}  // namespace asl

#endif
