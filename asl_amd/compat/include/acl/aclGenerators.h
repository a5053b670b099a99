// aclGenerators.h — acl::generateVEConstant / generateVEConstantN
// (src/acl/aclGenerators.h:36-60), host constants.
#ifndef ACLGENERATORS_H
#define ACLGENERATORS_H

#include "../math/aslVectors.h"
#include "aclMath/aclVectorOfElementsDef.h"

namespace acl {

template <typename T>
inline VectorOfElements generateVEConstant(T a) {
  VectorOfElements v(1);
  v.values[0] = double(a);
  v.type = typeToTypeID<T>();
  return v;
}
template <typename T>
inline VectorOfElements generateVEConstant(T a, T b) {
  VectorOfElements v(2);
  v.values = {double(a), double(b)};
  v.type = typeToTypeID<T>();
  return v;
}
template <typename T>
inline VectorOfElements generateVEConstant(T a, T b, T c) {
  VectorOfElements v(3);
  v.values = {double(a), double(b), double(c)};
  v.type = typeToTypeID<T>();
  return v;
}
template <typename T>
inline VectorOfElements generateVEConstant(const asl::AVec<T> &a) {
  VectorOfElements v(a.getSize());
  for (unsigned int i = 0; i < a.getSize(); ++i) v.values[i] = double(a[i]);
  v.type = typeToTypeID<T>();
  return v;
}
template <typename T>
inline VectorOfElements generateVEConstant(const std::vector<T> &a) {
  VectorOfElements v(static_cast<unsigned int>(a.size()));
  for (unsigned int i = 0; i < a.size(); ++i) v.values[i] = double(a[i]);
  v.type = typeToTypeID<T>();
  return v;
}
template <typename T>
inline VectorOfElements generateVEConstantN(unsigned int n, T a) {
  VectorOfElements v(n);
  for (auto &x : v.values) x = double(a);
  v.type = typeToTypeID<T>();
  return v;
}

}  // namespace acl

#endif
