// aslBlocks.h — asl::Block, the grid geometry (src/data/aslBlocks.h:56-196):
// size, position, dx and the linear index x*Sy*Sz + y*Sz + z (last index
// fastest).
#ifndef ASLBLOCKS_H
#define ASLBLOCKS_H

#include "../math/aslVectors.h"

namespace asl {

class Block {
 public:
  typedef AVec<int> DV;
  typedef AVec<> V;
  V position;
  double dx = 1.;
  DV c2iTransformVector;

  Block() : position(1, 0.), c2iTransformVector(1, 1), size_(1, 1) {}
  explicit Block(unsigned int nd) : position(nd, 0.), c2iTransformVector(nd, 1), size_(nd, 0) {}
  explicit Block(const DV &s, double d = 1.) : position(s.getSize(), 0.), dx(d), size_(s) { update(); }
  Block(const DV &s, double d, const V &p) : position(p), dx(d), size_(s) {
    if (p.getSize() != s.getSize()) errorMessage("Block::Block() Size and Position dimensionalities are different");
    update();
  }
  int c2i(const DV &c) const {
    if (c.getSize() != c2iTransformVector.getSize())
      errorMessage("Block::c2i() - The input vector size does not correspond to the block dimensionality");
    return c * c2iTransformVector;
  }
  void setSize(const DV &s) {
    size_ = s;
    update();
  }
  const DV &getSize() const { return size_; }
  const V getBPosition() const { return position + AVec<>(size_ - DV(size_.getSize(), 1)) * dx; }

 private:
  DV size_;
  void update() {
    const unsigned int n = size_.getSize();
    c2iTransformVector = DV(n, 1);
    int a = 1;
    for (unsigned int i = 0; i + 1 < n; ++i) {
      a *= size_[n - 1 - i];
      c2iTransformVector[n - 2 - i] = a;
    }
  }
};

inline const unsigned int nD(const Block &b) { return b.getSize().getSize(); }

// offset(b, a): the block grown by a cells on every side (src/data/aslBlocks.cxx:28-37).
inline const Block offset(const Block &bl, int a = 1) {
  if (minComponent(bl.getSize()) < -a * 2) errorMessage("The inset is larger than the block size");
  const unsigned int n = nD(bl);
  Block b(bl.getSize() + Block::DV(n, 2 * a), bl.dx, bl.position - Block::V(n, a * bl.dx));
  return b;
}

}  // namespace asl

#endif
