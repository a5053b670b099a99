// aslBCond.h — boundary-condition base classes and slice helpers
// (src/num/aslBCond.h:58-321).
#ifndef ASLBCOND_H
#define ASLBCOND_H

#include <memory>
#include <vector>

#include "../data/aslBlocks.h"
#include "../math/aslTemplates.h"
#include "aslNumMethod.h"

namespace asl {

/// names of the walls of a box (src/num/aslBCond.h:309)
enum SlicesNames { X0, XE, Y0, YE, Z0, ZE };

class BCond : public NumMethod {
 public:
  explicit BCond(const Block &b) : block(b), templ(nullptr) {}
  BCond(const Block &b, const VectorTemplate *const t) : block(b), templ(t) {}
  const Block &getBlock() { return block; }
  const VectorTemplate *getVT() { return templ; }
  /// faces this BC covers (bit 1 << SlicesNames)
  unsigned int faces() const { return faces_; }
  void addFace(SlicesNames s) { faces_ |= 1u << unsigned(s); }

 protected:
  Block block;
  const VectorTemplate *templ;
  unsigned int faces_ = 0;
};
typedef std::shared_ptr<BCond> SPBCond;

void addSliceX0(BCond &);
void addSliceXE(BCond &);
void addSliceY0(BCond &);
void addSliceYE(BCond &);
void addSliceZ0(BCond &);
void addSliceZE(BCond &);
void addSlices(BCond &, const std::vector<SlicesNames> &);

}  // namespace asl

#endif
