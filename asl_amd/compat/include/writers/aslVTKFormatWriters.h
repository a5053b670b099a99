// aslVTKFormatWriters.h — WriterVTKXML without VTK (src/writers/aslVTKFormatWriters.h:41-57):
// writes <file>_<n>.vti (VTK XML ImageData, appended raw binary) over the
// internal block of the first registered field.
#ifndef ASLVTKFORMATWRITERS_H
#define ASLVTKFORMATWRITERS_H

#include "aslWriter.h"

namespace asl {

class WriterVTKXML : public Writer {
 public:
  explicit WriterVTKXML(const std::string &file_) : Writer(file_) {}
  virtual void write();
};

}  // namespace asl

#endif
