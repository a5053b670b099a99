// aslUtilities.h — ASL's error and string helpers (src/aslUtilities.h).
// Errors throw std::logic_error("ASL ERROR: ...") like ASL
// (src/aslUtilities.cxx:68-95).
#ifndef ASLUTILITIES_H
#define ASLUTILITIES_H

#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

namespace asl {

inline void errorMessage(const std::string &msg) { throw std::logic_error("ASL ERROR: " + msg); }

// ASL's test helper: prints the outcome of a check (src/aslUtilities.cxx:98-104).
inline void errorMessage(bool status) { std::cout << (status ? "Test passed" : "Test FAILED") << std::endl; }

inline void warningMessage(const std::string &msg) { std::cerr << "ASL WARNING: " << msg << std::endl; }

template <typename T>
std::string numToStr(const T &v) {
  std::ostringstream s;
  s << v;
  return s.str();
}

inline std::string numToStr(int v, unsigned int digits) {
  std::ostringstream s;
  s.width(digits);
  s.fill('0');
  s << v;
  return s.str();
}

}  // namespace asl

using std::cout;
using std::endl;

#endif
