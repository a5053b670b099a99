// aslParametersManager.h — ASL's parameter API without Boost
// (src/utilities/aslParametersManager.h:44-175).  Parameters register with
// the current manager; load(argc, argv) reads "--key value", "--key=value"
// and "key = value" lines of a parameters file given as the first
// positional argument.  "--platform"/"--device" are accepted (one GPU here).
#ifndef ASLPARAMETERSMANAGER_H
#define ASLPARAMETERSMANAGER_H

#include <functional>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../math/aslVectors.h"
#include "aslUValue.h"

namespace asl {

namespace detail {
template <typename T>
inline void parse_value(const std::string &s, T &v) {
  std::istringstream is(s);
  if (!(is >> v)) errorMessage("ParametersManager: cannot parse \"" + s + "\"");
}
inline void parse_value(const std::string &s, std::string &v) { v = s; }
template <typename T>
inline void parse_value(const std::string &s, AVec<T> &v) {
  std::istringstream is(s);
  std::vector<T> x;
  T a;
  while (is >> a) x.push_back(a);
  if (x.empty()) errorMessage("ParametersManager: cannot parse vector \"" + s + "\"");
  v = AVec<T>(static_cast<unsigned int>(x.size()));
  for (unsigned int i = 0; i < x.size(); ++i) v[i] = x[i];
}
}  // namespace detail

class ParametersManager {
 public:
  struct Entry {
    std::string description, units;
    bool required = false, set = false;
    std::function<void(const std::string &)> assign;
  };
  ParametersManager() { current = this; }
  ~ParametersManager() {
    if (current == this) current = nullptr;
  }
  void enable() { current = this; }
  template <typename T>
  void add(UValue<T> parameter, const char *key, const char *description, const char *units) {
    Entry e{description, units, true, false, [parameter](const std::string &s) mutable {
              detail::parse_value(s, parameter.v());
            }};
    entries_[key] = e;
  }
  template <typename T>
  void add(UValue<T> parameter, T defaultValue, const char *key, const char *description, const char *units) {
    parameter.v() = defaultValue;
    add(parameter, key, description, units);
    entries_[key].required = false;
  }
  void load(std::string paramFile);
  void set(const std::string &key, const std::string &value);
  std::string getDir() { return dir_; }
  static ParametersManager *current;

 protected:
  std::map<std::string, Entry> entries_;
  std::string dir_ = "./";
  void check_required();
  void print_help(const std::string &name) const;
};

template <typename T>
class Parameter {
 public:
  Parameter(const char *key_, const char *description_, const char *units_ = "")
      : key(key_), description(description_), units(units_) {
    if (!ParametersManager::current) errorMessage("Parameter \"" + key + "\" declared before its ParametersManager");
    ParametersManager::current->add(parameter, key_, description_, units_);
  }
  Parameter(T defaultValue, const char *key_, const char *description_, const char *units_ = "")
      : key(key_), description(description_), units(units_) {
    if (!ParametersManager::current) errorMessage("Parameter \"" + key + "\" declared before its ParametersManager");
    ParametersManager::current->add(parameter, defaultValue, key_, description_, units_);
  }
  const T &v() const { return parameter.v(); }
  T &v() { return parameter.v(); }
  std::shared_ptr<T> p() { return parameter.p; }

 private:
  UValue<T> parameter;
  const std::string key, description, units;
};

class ApplicationParametersManager : public ParametersManager {
 public:
  ApplicationParametersManager(const char *applicationName_, const char *applicationVersion_);
  void load(int argc, char *argv[]);

 private:
  UValue<std::string> platform, device;
  std::string applicationName, applicationVersion;
};

}  // namespace asl

#endif
