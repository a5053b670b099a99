// aslWriter.h — output writers (src/writers/aslWriter.h:44-72).
#ifndef ASLWRITER_H
#define ASLWRITER_H

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../data/aslDataWithGhostNodes.h"

namespace asl {

class Writer {
 public:
  explicit Writer(const std::string &file_) : file(file_), numOfWrites(0) { current = this; }
  virtual ~Writer() {
    if (current == this) current = nullptr;
  }
  virtual void write() = 0;
  void enable() { current = this; }
  void addScalars(std::string name, AbstractData &data) { scalarFields.push_back({name, &data}); }
  void addVector(std::string name, AbstractData &data) { vectorFields.push_back({name, &data}); }
  static Writer *current;

 protected:
  std::vector<std::pair<std::string, AbstractData *>> scalarFields, vectorFields;
  std::string file;
  unsigned int numOfWrites;
};

}  // namespace asl

#endif
