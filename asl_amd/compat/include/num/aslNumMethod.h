// aslNumMethod.h — the plugin interface (src/num/aslNumMethod.h:34-77).
#ifndef ASLNUMMETHOD_H
#define ASLNUMMETHOD_H

#include <memory>
#include <vector>

namespace asl {

class NumMethod {
 public:
  virtual void execute() = 0;
  virtual void init() = 0;
  virtual ~NumMethod() {}
};
typedef std::shared_ptr<NumMethod> SPNumMethod;

template <class T>
inline void initAll(std::vector<T *> &v) {
  for (auto *m : v) m->init();
}
template <class T>
inline void initAll(std::vector<std::shared_ptr<T>> &v) {
  for (auto &m : v) m->init();
}
template <class T>
inline void executeAll(std::vector<T *> &v) {
  for (auto *m : v) m->execute();
}
template <class T>
inline void executeAll(std::vector<std::shared_ptr<T>> &v) {
  for (auto &m : v) m->execute();
}

}  // namespace asl

#endif
