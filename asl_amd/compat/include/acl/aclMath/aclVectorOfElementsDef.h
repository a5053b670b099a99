// aclVectorOfElementsDef.h — acl::VectorOfElements as the LBGK interface
// uses it: parameters built by generateVEConstant (src/acl/aclGenerators.h:36-48).
// Here it holds host constants (doubles) with the element type they were
// made from; that type selects FP32 or FP64 lattices like ASL's kernels.
#ifndef ACLVECTOROFELEMENTSDEF_H
#define ACLVECTOROFELEMENTSDEF_H

#include <vector>

#include "../aclTypes.h"

namespace acl {

class VectorOfElements {
 public:
  VectorOfElements() {}
  explicit VectorOfElements(unsigned int n) : values(n, 0.) {}
  unsigned int size() const { return static_cast<unsigned int>(values.size()); }
  double operator[](unsigned int i) const { return values[i]; }
  std::vector<double> values;
  TypeID type = TYPE_DOUBLE;
};

typedef VectorOfElements VectorOfElementsData;

}  // namespace acl

// ASL's header opens namespace std at global scope
// (src/acl/aclMath/aclVectorOfElementsDef.h:31); drivers rely on it.
using namespace std;

#endif
