// aslTemplates.h — asl::VectorTemplate for the LBGK lattices
// (src/math/aslTemplates.h:87-105, src/math/aslTemplates.cxx:146-200):
// vectors, laplasCoefs (lc), quasiparticlesCoefs (lc/6, +1 on rest) and the
// inverse-direction table.  d3q27 carries the textbook weights (ASL's has none).
#ifndef ASLTEMPLATES_H
#define ASLTEMPLATES_H

#include <vector>

#include "aslVectors.h"

namespace asl {

class VectorTemplate {
 public:
  std::vector<AVec<int>> vectors;
  std::vector<double> laplasCoefs;
  std::vector<double> quasiparticlesCoefs;
  std::vector<unsigned int> invertVectors;
};

const VectorTemplate &d3q15();
const VectorTemplate &d3q19();
const VectorTemplate &d3q27();

inline unsigned int nD(const VectorTemplate &t) { return t.vectors[0].getSize(); }

}  // namespace asl

#endif
