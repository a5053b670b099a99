// aslVectors.h — asl::AVec, the small dynamic-length vector of ASL's API
// (src/math/aslVectorsDynamicLength.h): the element-wise and dot-product
// operators the drivers and the LBGK interface use.
#ifndef ASLVECTORS_H
#define ASLVECTORS_H

#include <cmath>
#include <initializer_list>
#include <vector>

#include "../aslUtilities.h"

namespace asl {

template <typename T = double>
class AVec {
 public:
  AVec() : v_(1, T(0)) {}
  explicit AVec(unsigned int n) : v_(n, T(0)) {}
  AVec(unsigned int n, T a) : v_(n, a) {}
  template <typename U>
  explicit AVec(const AVec<U> &a) : v_(a.getSize()) {
    for (unsigned int i = 0; i < a.getSize(); ++i) v_[i] = T(a[i]);
  }
  unsigned int getSize() const { return static_cast<unsigned int>(v_.size()); }
  void resize(unsigned int n) { v_.resize(n); }
  T &operator[](unsigned int i) { return v_[i]; }
  const T &operator[](unsigned int i) const { return v_[i]; }
  const T *getPointer() const { return v_.data(); }
  AVec &operator+=(const AVec &a) {
    check(a);
    for (unsigned int i = 0; i < getSize(); ++i) v_[i] += a[i];
    return *this;
  }
  AVec &operator-=(const AVec &a) {
    check(a);
    for (unsigned int i = 0; i < getSize(); ++i) v_[i] -= a[i];
    return *this;
  }
  AVec &operator*=(T a) {
    for (auto &x : v_) x *= a;
    return *this;
  }
  AVec &operator/=(T a) {
    for (auto &x : v_) x /= a;
    return *this;
  }
  void check(const AVec &a) const {
    if (a.getSize() != getSize()) errorMessage("AVec: operands have different sizes");
  }

 private:
  std::vector<T> v_;
};

template <typename T>
inline unsigned int nD(const AVec<T> &a) { return a.getSize(); }

template <typename T>
AVec<T> makeAVec(T a) { AVec<T> r(1); r[0] = a; return r; }
template <typename T>
AVec<T> makeAVec(T a, T b) { AVec<T> r(2); r[0] = a; r[1] = b; return r; }
template <typename T>
AVec<T> makeAVec(T a, T b, T c) { AVec<T> r(3); r[0] = a; r[1] = b; r[2] = c; return r; }

template <typename T>
AVec<T> operator+(AVec<T> a, const AVec<T> &b) { return a += b; }
template <typename T>
AVec<T> operator-(AVec<T> a, const AVec<T> &b) { return a -= b; }
template <typename T>
AVec<T> operator-(AVec<T> a) { return a *= T(-1); }
template <typename T>
T operator*(const AVec<T> &a, const AVec<T> &b) {  // dot product
  a.check(b);
  T s = a[0] * b[0];
  for (unsigned int i = 1; i < a.getSize(); ++i) s = s + a[i] * b[i];
  return s;
}
template <typename T>
AVec<T> operator*(AVec<T> a, T b) { return a *= b; }
template <typename T>
AVec<T> operator*(T b, AVec<T> a) { return a *= b; }
template <typename T>
AVec<T> operator/(AVec<T> a, T b) { return a /= b; }
template <typename T>
bool operator==(const AVec<T> &a, const AVec<T> &b) {
  if (a.getSize() != b.getSize()) return false;
  for (unsigned int i = 0; i < a.getSize(); ++i)
    if (!(a[i] == b[i])) return false;
  return true;
}
template <typename T>
bool operator!=(const AVec<T> &a, const AVec<T> &b) { return !(a == b); }

template <typename T>
T l2(const AVec<T> &a) { return a * a; }
template <typename T>
AVec<T> normalize(const AVec<T> &a) { return a / T(std::sqrt(double(l2(a)))); }
template <typename T>
T productOfElements(const AVec<T> &a) {
  T p = a[0];
  for (unsigned int i = 1; i < a.getSize(); ++i) p *= a[i];
  return p;
}
template <typename T>
T minComponent(const AVec<T> &a) {
  T m = a[0];
  for (unsigned int i = 1; i < a.getSize(); ++i) m = a[i] < m ? a[i] : m;
  return m;
}
template <typename T>
T maxComponent(const AVec<T> &a) {
  T m = a[0];
  for (unsigned int i = 1; i < a.getSize(); ++i) m = a[i] > m ? a[i] : m;
  return m;
}
template <typename T>
AVec<T> crossProduct(const AVec<T> &a, const AVec<T> &b) {
  AVec<T> c(3);
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
  return c;
}

}  // namespace asl

#endif
