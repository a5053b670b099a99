// aslLBGK.h — asl::LBGK on the gfx950 library (src/num/aslLBGK.h:77-191).
//
// The constructor creates one lattice handle (include/aslhip.h) for the
// internal block; the element type of the viscosity parameter picks FP32 or
// FP64 populations, as the type of ASL's generated kernel does.  execute()
// requests one fused step (stream + collide + every executed BC); the step
// stays pending until the host next reads the lattice or requests another
// step, so that rho and u are written only by the step a host read observes.
#ifndef ASLLBGK_H
#define ASLLBGK_H

#include <memory>

#include "../acl/aclGenerators.h"
#include <aslhip.h>
#include "../data/aslDataWithGhostNodes.h"
#include "../math/aslTemplates.h"
#include "aslSingleKernelNM.h"

namespace asl {

class LBGK : public SingleKernelNM {
 public:
  typedef SPDataWithGhostNodesACL Data;
  typedef SPDataWithGhostNodesACLData DataD;
  typedef acl::VectorOfElements Param;

  const VectorTemplate *vectorTemplate;

  LBGK(DataD v, Param nu, const VectorTemplate *vT);
  LBGK(Block b, Param nu, const VectorTemplate *vT, bool compVel = true, bool compRho = true,
       acl::CommandQueue queue = acl::hardware.defaultQueue);
  ~LBGK();
  virtual void init();
  virtual void execute();
  void setViscosity(Param nu);
  double getViscosity(unsigned int i = 0);
  /// Coriolis term of a rotating frame (src/num/aslLBGK.cxx:207-208); an
  /// empty Param switches it off.
  void setOmega(Param w);
  void setVectorTemplate(const VectorTemplate *vT);
  const VectorTemplate *getVectorTemplate() const { return vectorTemplate; }
  Data getF();
  DataD getRho();
  DataD getVelocity();
  void setCompressible(bool flag = true);
  const bool &getCompressible() const { return compressible_; }

  // the library handle (for the BC layer and for callers of the C ABI)
  aslhip_lattice *handle() const { return h_; }
  acl::TypeID typeID() const { return type_; }
  // the internal block the lattice was built on
  const Block &block() const { return block_; }
  // number of state changes so far (cached host views compare against it)
  unsigned long long version() const { return version_; }
  void execute_count_bump() { ++version_; }

 protected:
  void create(const Block &b, acl::TypeID t);
  aslhip_lattice *h_ = nullptr;
  Block block_;
  acl::TypeID type_ = acl::TYPE_FLOAT;
  bool compressible_ = true;
  unsigned long long version_ = 0;
  double nu_ = 0.;
  DataD rho_, vel_;
  Data f_;
};
typedef std::shared_ptr<LBGK> SPLBGK;

class LBGKUtilities {
 public:
  typedef acl::VectorOfElements Param;
  explicit LBGKUtilities(SPLBGK lbgk) : num(lbgk) {}
  /// initF(rho, vel) with constants or a per-cell field container
  /// (src/num/aslLBGK.cxx:262-280)
  void initF(Param rho, Param vel);
  void initF(Param vel);
  void initF(SPAbstractDataWithGhostNodes rho, SPAbstractDataWithGhostNodes vel);

 private:
  SPLBGK num;
};
typedef std::shared_ptr<LBGKUtilities> SPLBGKUtilities;

/// LBGKTurbulence (src/num/aslLBGK.h:163-171, src/num/aslLBGK.cxx:298-355):
/// LBGK whose relaxation rate is limited per cell by the distance of the
/// populations from equilibrium.  ASL evaluates that distance with fEq read
/// before it is assigned (:315-317, an uninitialised private variable of the
/// generated kernel); here fEq is the equilibrium of the cell's own density
/// and momentum.
class LBGKTurbulence : public LBGK {
 public:
  LBGKTurbulence(DataD v, Param nu, const VectorTemplate *vT);
  LBGKTurbulence(Block b, Param nu, const VectorTemplate *vT, bool compVel = true, bool compRho = true,
                 acl::CommandQueue queue = acl::hardware.defaultQueue);
};
typedef std::shared_ptr<LBGKTurbulence> SPLBGKTurbulence;

}  // namespace asl

#endif
