// aslDataWithGhostNodes.h — ASL's data containers with one ghost layer
// (src/data/aslDataWithGhostNodes.h:57-145, src/data/aslDataWrapper.h).
// getBlock() is the ghost-inclusive block; values are host arrays in ASL's
// layout, one per component.  Containers that view a lattice field (rho, u,
// f) fetch it from the GPU when read.
#ifndef ASLDATAWITHGHOSTNODES_H
#define ASLDATAWITHGHOSTNODES_H

#include <functional>
#include <memory>
#include <vector>

#include "../acl/aclTypes.h"
#include "aslBlocks.h"

namespace asl {

class AbstractData {
 public:
  explicit AbstractData(const Block &b) : block(b) {}
  virtual ~AbstractData() {}
  const Block &getBlock() const { return block; }
  unsigned int getNComponents() const { return static_cast<unsigned int>(values_.size()); }
  acl::TypeID getTypeID() const { return type_; }
  // Host values of component c (ghost-inclusive layout), fetched first if
  // the container views a lattice field.
  std::vector<double> &component(unsigned int c) {
    refresh();
    return values_.at(c);
  }
  // Values as the element type of the container (float maps stay exact).
  std::vector<float> componentAsFloat(unsigned int c);
  void refresh() {
    if (fetch_) fetch_(*this);
  }
  // After values_ were written by a reader (ABD): hands them to the lattice
  // field the container views, if any.
  void commit() {
    if (store_) store_(*this);
  }
  // Device address of the values of a lattice field (getRho(),
  // getVelocity(): component-major, ghost-inclusive ASL layout, the
  // lattice's element type) for a solver coupled on the GPU, as ASL's device
  // containers are handed to one; nullptr for host-only containers.  stream:
  // the consumer's hipStream_t, ordered as aslhip_macro_device orders it.
  const void *deviceData(void *stream = nullptr) { return device_ ? device_(stream) : nullptr; }
  // internal: storage and the fetch / store / device hooks
  std::vector<std::vector<double>> values_;
  acl::TypeID type_ = acl::TYPE_DOUBLE;
  std::function<void(AbstractData &)> fetch_;
  std::function<void(AbstractData &)> store_;
  std::function<const void *(void *)> device_;

 protected:
  Block block;
};

class AbstractDataWithGhostNodes : public AbstractData {
 public:
  explicit AbstractDataWithGhostNodes(const Block &b, int nGN = 1)
      : AbstractData(offset(b, nGN)), internalBlock(b), ghostBorder(nGN) {}
  const Block &getInternalBlock() const { return internalBlock; }
  const unsigned int getGhostBorder() const { return ghostBorder; }

 protected:
  Block internalBlock;
  unsigned int ghostBorder;
};

class DataWithGhostNodesACLData : public AbstractDataWithGhostNodes {
 public:
  DataWithGhostNodesACLData(const Block &b, unsigned int n, acl::TypeID t, int nGN = 1)
      : AbstractDataWithGhostNodes(b, nGN) {
    const int len = productOfElements(getBlock().getSize());
    values_.assign(n, std::vector<double>(len, 0.));
    type_ = t;
  }
};
typedef DataWithGhostNodesACLData DataWithGhostNodesACL;

typedef std::shared_ptr<AbstractData> SPAbstractData;
typedef std::shared_ptr<AbstractDataWithGhostNodes> SPAbstractDataWithGhostNodes;
typedef std::shared_ptr<DataWithGhostNodesACLData> SPDataWithGhostNodesACLData;
typedef std::shared_ptr<DataWithGhostNodesACL> SPDataWithGhostNodesACL;

// generateDataContainerACL_SP<T>(block, n, gN) (src/aslGenerators.cxx:108-155)
template <typename T>
SPDataWithGhostNodesACLData generateDataContainerACL_SP(const Block &b, unsigned int n, unsigned int gN) {
  return std::make_shared<DataWithGhostNodesACLData>(b, n, acl::typeToTypeID<T>(), int(gN));
}
SPDataWithGhostNodesACLData generateDataContainerACL_SP(const Block &b, acl::TypeID t, unsigned int n,
                                                        unsigned int gN);

}  // namespace asl

#endif
