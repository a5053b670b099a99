// aslUValue.h — asl::UValue<T> (src/utilities/aslUValue.h:34-45).
#ifndef ASLUVALUE_H
#define ASLUVALUE_H

#include <memory>

namespace asl {

template <typename T>
class UValue {
 public:
  std::shared_ptr<T> p;
  UValue() : p(new T) {}
  explicit UValue(const T &a) : p(new T(a)) {}
  const T &operator=(const T &a) { return (*p = a); }
  const T &v() const { return *p; }
  T &v() { return *p; }
};

}  // namespace asl

#endif
