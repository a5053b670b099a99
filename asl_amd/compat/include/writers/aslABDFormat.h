// aslABDFormat.h — the ASL Binary Dump stream (src/writers/aslABDFormat.h:
// 36-135, src/writers/aslABDFormat.cxx:35-81): values in native byte order
// with no padding or tags; an AVec or string is its unsigned length followed
// by its elements; a Block is dx, size, position; a data container is its
// first component as size doubles (the block written separately).
#ifndef ASLABDFORMAT_H
#define ASLABDFORMAT_H

#include <fstream>
#include <memory>
#include <string>
#include <utility>

#include <iostream>

#include "../data/aslBlocks.h"
#include "../data/aslDataWithGhostNodes.h"

// ASL's header opens namespace std to its includers, and code written
// against it (its own ABD test) relies on that.
using namespace std;

namespace asl {

class ABDFileIn : public std::ifstream {
 public:
  ABDFileIn() : std::ifstream() {}
  explicit ABDFileIn(std::string name) : std::ifstream(name, std::ios::in | std::ios::binary) {}
};

class ABDFileOut : public std::ofstream {
 public:
  ABDFileOut() : std::ofstream() {}
  explicit ABDFileOut(std::string name) : std::ofstream(name, std::ios::out | std::ios::binary) {}
};

namespace abd_detail {
template <typename T>
inline ABDFileOut &put(ABDFileOut &f, const T &a) {
  f.write(reinterpret_cast<const char *>(&a), sizeof(T));
  return f;
}
template <typename T>
inline ABDFileIn &get(ABDFileIn &f, T &a) {
  f.read(reinterpret_cast<char *>(&a), sizeof(T));
  return f;
}
}  // namespace abd_detail

inline ABDFileOut &operator<<(ABDFileOut &f, const int a) { return abd_detail::put(f, a); }
inline ABDFileIn &operator>>(ABDFileIn &f, int &a) { return abd_detail::get(f, a); }
inline ABDFileOut &operator<<(ABDFileOut &f, const unsigned int a) { return abd_detail::put(f, a); }
inline ABDFileIn &operator>>(ABDFileIn &f, unsigned int &a) { return abd_detail::get(f, a); }
inline ABDFileOut &operator<<(ABDFileOut &f, const float a) { return abd_detail::put(f, a); }
inline ABDFileIn &operator>>(ABDFileIn &f, float &a) { return abd_detail::get(f, a); }
inline ABDFileOut &operator<<(ABDFileOut &f, const double a) { return abd_detail::put(f, a); }
inline ABDFileIn &operator>>(ABDFileIn &f, double &a) { return abd_detail::get(f, a); }

template <typename T>
inline ABDFileOut &operator<<(ABDFileOut &f, std::pair<T *, unsigned int> a) {
  f.write(reinterpret_cast<const char *>(a.first), sizeof(T) * a.second);
  return f;
}
template <typename T>
inline ABDFileIn &operator>>(ABDFileIn &f, std::pair<T *, unsigned int> a) {
  f.read(reinterpret_cast<char *>(a.first), sizeof(T) * a.second);
  return f;
}

inline ABDFileOut &operator<<(ABDFileOut &f, const std::string &a) {
  const unsigned int n = static_cast<unsigned int>(a.size());
  f << n;
  f.write(a.data(), n);
  return f;
}
inline ABDFileIn &operator>>(ABDFileIn &f, std::string &a) {
  unsigned int n = 0;
  f >> n;
  a.resize(n);
  if (n) f.read(&a[0], n);
  return f;
}

template <typename T>
inline ABDFileOut &operator<<(ABDFileOut &f, const AVec<T> &a) {
  const unsigned int n = a.getSize();
  f << n;
  for (unsigned int i = 0; i < n; ++i) abd_detail::put(f, a[i]);
  return f;
}
template <typename T>
inline ABDFileIn &operator>>(ABDFileIn &f, AVec<T> &a) {
  unsigned int n = 0;
  f >> n;
  a.resize(n);
  for (unsigned int i = 0; i < n; ++i) abd_detail::get(f, a[i]);
  return f;
}

ABDFileOut &operator<<(ABDFileOut &f, const Block &a);
ABDFileIn &operator>>(ABDFileIn &f, Block &a);
/// the first component of a, as doubles (the block is written separately)
ABDFileOut &operator<<(ABDFileOut &f, const AbstractData &a);
/// reads the first component of a; a must already have the right size
ABDFileIn &operator>>(ABDFileIn &f, AbstractData &a);
/// as operator>>, staged through d (allocated when empty)
ABDFileIn &get(ABDFileIn &f, AbstractData &a, std::shared_ptr<double> d);
/// Declared by ASL without a definition (src/writers/aslABDFormat.h:132-135);
/// here: name, block, number of components, then every component as doubles.
void writeABD(const std::string &fileName, const AbstractData &data, const std::string &name);

}  // namespace asl

#endif
