// aslLBGKBC.h — LBGK boundary conditions (src/num/aslLBGKBC.h:45-336).
//
// init() registers the BC with the lattice (include/aslhip.h); execute()
// records the driver's execute order, which decides which BC fills a ghost
// link that several BCs claim (the last one executed).  The fused step
// applies every BC itself.
#ifndef ASLLBGKBC_H
#define ASLLBGKBC_H

#include <memory>
#include <vector>

#include "../acl/aclGenerators.h"
#include "../data/aslDataWithGhostNodes.h"
#include "../math/aslPositionFunction.h"
#include "aslBCond.h"
#include "aslLBGK.h"

namespace asl {

class BCLBGKCommon : public BCond {
 protected:
  SPLBGK num;
  int kind_;
  double p_ = 0., u_[3] = {0., 0., 0.};
  int id_ = -1;

 public:
  BCLBGKCommon(SPLBGK nm, int kind);
  virtual void init();
  virtual void execute();
};

class BCNoSlip : public BCLBGKCommon {
 public:
  explicit BCNoSlip(SPLBGK nm);
};

class BCConstantPressure : public BCLBGKCommon {
 public:
  BCConstantPressure(SPLBGK nm, const acl::VectorOfElements &p);
};

class BCConstantVelocity : public BCLBGKCommon {
 public:
  BCConstantVelocity(SPLBGK nm, const acl::VectorOfElements &v);
};

class BCConstantPressureVelocity : public BCLBGKCommon {
 public:
  BCConstantPressureVelocity(SPLBGK nm, const acl::VectorOfElements &p, const acl::VectorOfElements &v);
};

class BCondWithMap : public NumMethod {
 protected:
  SPLBGK num;
  SPAbstractDataWithGhostNodes map, computationalDomain;
  int kind_;
  double p_ = 0., u_[3] = {0., 0., 0.};
  bool has_rot_ = false;
  double rot_[6] = {0., 0., 0., 0., 0., 0.};
  int id_ = -1;

 public:
  BCondWithMap(SPLBGK nm, int kind, SPAbstractDataWithGhostNodes m, SPAbstractDataWithGhostNodes cd = nullptr);
  virtual void init();
  virtual void execute();
};

class BCNoSlipMap : public BCondWithMap {
 public:
  BCNoSlipMap(SPLBGK nm, SPAbstractDataWithGhostNodes map);
};

class BCVelocityMap : public BCondWithMap {
 public:
  BCVelocityMap(SPLBGK nm, SPPositionFunction v, SPAbstractDataWithGhostNodes map);
  BCVelocityMap(SPLBGK nm, SPPositionFunction v, SPAbstractDataWithGhostNodes map,
                SPAbstractDataWithGhostNodes computationalDomain);
};

class BCConstantPressureVelocityMap : public BCondWithMap {
 public:
  BCConstantPressureVelocityMap(SPLBGK nm, acl::VectorOfElements p, SPAbstractDataWithGhostNodes map);
  BCConstantPressureVelocityMap(SPLBGK nm, acl::VectorOfElements p, acl::VectorOfElements v,
                                SPAbstractDataWithGhostNodes map);
};

/// Visualisation fills of an LBGK's rho or velocity at the boundary nodes of
/// a map (map <= 0 and > -0.9999), on the GPU in place (aslhip_vis_add /
/// aslhip_vis_execute), behind ASL's generateBCNoSlipVel, generateBCNoSlipRho
/// and generateBCVelocityVel (src/num/aslLBGKBC.cxx:808-852): the kernels of
/// BCConstantValueMiddlePointMap, BCConstantGradientMap and BCValuePFMap
/// (src/num/aslBasicBC.cxx:124-145, 167-227, 283-351).  init() copies the
/// map to the device; execute() fills and marks the host views stale.
class BCMacroFillMap : public NumMethod {
 public:
  BCMacroFillMap(SPLBGK nm, int kind, SPAbstractDataWithGhostNodes map, SPPositionFunction v = nullptr);
  virtual void init();
  virtual void execute();

 private:
  SPLBGK num;
  int kind_;
  SPAbstractDataWithGhostNodes map;
  SPPositionFunction value;
  int id_ = -1;
};

// ComputeSurfaceFluxMap / ComputeSurfaceForceMap (src/num/aslLBGKBC.h:247-279):
// execute() fills the field with the mass / momentum exchanged across the
// links between the ghost and computation nodes of map (aslhip_surface_flux /
// aslhip_surface_force).
class ComputeSurfaceFluxMap : public NumMethod {
 public:
  ComputeSurfaceFluxMap(SPLBGK nm, SPDataWithGhostNodesACLData fF, SPAbstractDataWithGhostNodes map);
  virtual void init() {}
  virtual void execute();

 protected:
  SPLBGK num;
  SPDataWithGhostNodesACLData field;
  SPAbstractDataWithGhostNodes map;
};

class ComputeSurfaceForceMap : public ComputeSurfaceFluxMap {
 public:
  ComputeSurfaceForceMap(SPLBGK nm, SPDataWithGhostNodesACLData fF, SPAbstractDataWithGhostNodes map)
      : ComputeSurfaceFluxMap(nm, fF, map) {}
  virtual void execute();
};

SPNumMethod generateComputeSurfaceFlux(SPLBGK nm, SPDataWithGhostNodesACLData fF, SPAbstractDataWithGhostNodes map);
SPNumMethod generateComputeSurfaceForce(SPLBGK nm, SPDataWithGhostNodesACLData fF, SPAbstractDataWithGhostNodes map);

SPBCond generateBCNoSlip(SPLBGK nm, const std::vector<SlicesNames> &sl);
SPBCond generateBCConstantVelocity(SPLBGK nm, AVec<> v, const std::vector<SlicesNames> &sl);
SPBCond generateBCConstantPressure(SPLBGK nm, double p, const std::vector<SlicesNames> &sl);
SPBCond generateBCConstantPressureVelocity(SPLBGK nm, double p, AVec<> v, const std::vector<SlicesNames> &sl);
SPNumMethod generateBCConstantPressure(SPLBGK nm, double p, SPAbstractDataWithGhostNodes map);
SPNumMethod generateBCConstantPressureVelocity(SPLBGK nm, double p, AVec<> v, SPAbstractDataWithGhostNodes map);
SPNumMethod generateBCNoSlip(SPLBGK nm, SPAbstractDataWithGhostNodes map);
SPNumMethod generateBCNoSlipVel(SPLBGK nmU, SPAbstractDataWithGhostNodes map);
SPNumMethod generateBCNoSlipRho(SPLBGK nmU, SPAbstractDataWithGhostNodes map);
SPNumMethod generateBCVelocityVel(SPLBGK nm, SPPositionFunction v, SPAbstractDataWithGhostNodes map);
SPNumMethod generateBCVelocity(SPLBGK nm, SPPositionFunction v, SPAbstractDataWithGhostNodes map);
SPNumMethod generateBCVelocity(SPLBGK nm, SPPositionFunction v, SPAbstractDataWithGhostNodes map,
                               SPAbstractDataWithGhostNodes computationalDomain);

}  // namespace asl

#endif
