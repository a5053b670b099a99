// LBGK boundary conditions of the compat layer over the C ABI.
// init() registers a BC with the lattice (aslhip_bc_add_slice / _map);
// execute() records the driver's execute order (aslhip_bc_execute).
#include <algorithm>

#include <num/aslLBGKBC.h>

namespace asl {
namespace detail {
void check_abi(int r);
}
using detail::check_abi;

// addSliceX0 ... addSliceZE: the faces a slice BC covers
// (src/num/aslBCond.cxx:373-510).
void addSliceX0(BCond &a) { a.addFace(X0); }
void addSliceXE(BCond &a) { a.addFace(XE); }
void addSliceY0(BCond &a) { a.addFace(Y0); }
void addSliceYE(BCond &a) { a.addFace(YE); }
void addSliceZ0(BCond &a) { a.addFace(Z0); }
void addSliceZE(BCond &a) { a.addFace(ZE); }
void addSlices(BCond &bc, const std::vector<SlicesNames> &sl) {
  for (SlicesNames s : sl) bc.addFace(s);
}

namespace {

std::vector<float> map_array(LBGK &num, SPAbstractDataWithGhostNodes m, const char *what) {
  if (!m) errorMessage(std::string(what) + ": map is not initialized");
  const Block &lb = num.getRho()->getBlock();
  if (m->getBlock().getSize() != lb.getSize())
    errorMessage(std::string(what) + ": the map's block differs from the lattice's");
  return m->componentAsFloat(0);
}

void vec3(const acl::VectorOfElements &v, double *u, const char *what) {
  if (v.size() != 3) errorMessage(std::string(what) + ": velocity must have 3 components");
  for (int d = 0; d < 3; ++d) u[d] = v[d];
}

}  // namespace

BCLBGKCommon::BCLBGKCommon(SPLBGK nm, int kind)
    : BCond(nm->getRho()->getBlock(), nm->getVectorTemplate()), num(nm), kind_(kind) {}

void BCLBGKCommon::init() {
  if (id_ >= 0) return;
  if (!faces_) errorMessage("BCLBGKCommon::init: the boundary condition has no slices");
  check_abi(aslhip_bc_add_slice(num->handle(), kind_, faces_, p_, u_, &id_));
}

void BCLBGKCommon::execute() {
  if (id_ < 0) init();
  check_abi(aslhip_bc_execute(num->handle(), id_));
}

BCNoSlip::BCNoSlip(SPLBGK nm) : BCLBGKCommon(nm, ASLHIP_BC_NOSLIP) {}

BCConstantPressure::BCConstantPressure(SPLBGK nm, const acl::VectorOfElements &p)
    : BCLBGKCommon(nm, ASLHIP_BC_CONSTANT_PRESSURE) {
  if (p.size() != 1) errorMessage("BCConstantPressure: pressure must be a scalar");
  p_ = p[0];
}

BCConstantVelocity::BCConstantVelocity(SPLBGK nm, const acl::VectorOfElements &v)
    : BCLBGKCommon(nm, ASLHIP_BC_CONSTANT_VELOCITY) {
  vec3(v, u_, "BCConstantVelocity");
}

BCConstantPressureVelocity::BCConstantPressureVelocity(SPLBGK nm, const acl::VectorOfElements &p,
                                                       const acl::VectorOfElements &v)
    : BCLBGKCommon(nm, ASLHIP_BC_CONSTANT_PRESSURE_VELOCITY) {
  if (p.size() != 1) errorMessage("BCConstantPressureVelocity: pressure must be a scalar");
  p_ = p[0];
  vec3(v, u_, "BCConstantPressureVelocity");
}

BCondWithMap::BCondWithMap(SPLBGK nm, int kind, SPAbstractDataWithGhostNodes m, SPAbstractDataWithGhostNodes cd)
    : num(nm), map(m), computationalDomain(cd), kind_(kind) {}

void BCondWithMap::init() {
  if (id_ >= 0) return;
  const std::vector<float> m = map_array(*num, map, "BCondWithMap");
  std::vector<float> cd;
  if (computationalDomain) cd = map_array(*num, computationalDomain, "BCondWithMap (computational domain)");
  check_abi(aslhip_bc_add_map(num->handle(), kind_, m.data(), cd.empty() ? nullptr : cd.data(), p_, u_,
                              has_rot_ ? rot_ : nullptr, &id_));
}

void BCondWithMap::execute() {
  if (id_ < 0) init();
  check_abi(aslhip_bc_execute(num->handle(), id_));
}

BCNoSlipMap::BCNoSlipMap(SPLBGK nm, SPAbstractDataWithGhostNodes map)
    : BCondWithMap(nm, ASLHIP_BC_NOSLIP_MAP, map) {}

BCVelocityMap::BCVelocityMap(SPLBGK nm, SPPositionFunction v, SPAbstractDataWithGhostNodes map)
    : BCVelocityMap(nm, v, map, nullptr) {}

BCVelocityMap::BCVelocityMap(SPLBGK nm, SPPositionFunction v, SPAbstractDataWithGhostNodes map,
                             SPAbstractDataWithGhostNodes cd)
    : BCondWithMap(nm, ASLHIP_BC_VELOCITY_MAP, map, cd) {
  if (!v) errorMessage("BCVelocityMap: velocity function is not initialized");
  if (v->rotation) {
    has_rot_ = true;
    for (int d = 0; d < 3; ++d) {
      rot_[d] = v->axis[d];
      rot_[3 + d] = v->centre[d];
    }
  } else {
    if (v->constant.getSize() != 3) errorMessage("BCVelocityMap: velocity must have 3 components");
    for (int d = 0; d < 3; ++d) u_[d] = v->constant[d];
  }
}

BCConstantPressureVelocityMap::BCConstantPressureVelocityMap(SPLBGK nm, acl::VectorOfElements p,
                                                             SPAbstractDataWithGhostNodes map)
    : BCondWithMap(nm, ASLHIP_BC_PRESSURE_VELOCITY_MAP, map) {
  if (p.size() != 1) errorMessage("BCConstantPressureVelocityMap: pressure must be a scalar");
  p_ = p[0];
}

BCConstantPressureVelocityMap::BCConstantPressureVelocityMap(SPLBGK nm, acl::VectorOfElements p,
                                                             acl::VectorOfElements v,
                                                             SPAbstractDataWithGhostNodes map)
    : BCConstantPressureVelocityMap(nm, p, map) {
  vec3(v, u_, "BCConstantPressureVelocityMap");
}

BCMacroFillMap::BCMacroFillMap(SPLBGK nm, int kind, SPAbstractDataWithGhostNodes m, SPPositionFunction v)
    : num(nm), kind_(kind), map(m), value(v) {
  if (kind == ASLHIP_VIS_VELOCITY && !value) errorMessage("BCValuePFMap: no position function");
}

void BCMacroFillMap::init() {
  if (id_ >= 0) return;
  const std::vector<float> m = map_array(*num, map, "BCMacroFillMap");
  double u[3] = {0., 0., 0.}, rot[6] = {0., 0., 0., 0., 0., 0.};
  bool has_rot = false;
  if (kind_ == ASLHIP_VIS_VELOCITY) {
    if (value->rotation) {
      has_rot = true;
      for (int d = 0; d < 3; ++d) {
        rot[d] = value->axis[d];
        rot[3 + d] = value->centre[d];
      }
    } else {
      if (value->constant.getSize() != 3) errorMessage("BCValuePFMap: velocity must have 3 components");
      for (int d = 0; d < 3; ++d) u[d] = value->constant[d];
    }
  }
  check_abi(aslhip_vis_add(num->handle(), kind_, m.data(), u, has_rot ? rot : nullptr, &id_));
}

void BCMacroFillMap::execute() {
  if (id_ < 0) init();
  check_abi(aslhip_vis_execute(num->handle(), id_));
  num->execute_count_bump();  // the host views of rho and u fetch the filled values
}

ComputeSurfaceFluxMap::ComputeSurfaceFluxMap(SPLBGK nm, SPDataWithGhostNodesACLData fF,
                                             SPAbstractDataWithGhostNodes m)
    : num(nm), field(fF), map(m) {}

namespace {
// Runs the surface kernel and stores its result (ncomp components) in field.
void surface_into(LBGK &num, AbstractData &field, const std::vector<float> &m, unsigned int ncomp, bool flux) {
  if (field.getNComponents() != ncomp)
    errorMessage(flux ? "ComputeSurfaceFluxMap: the flux field needs 1 component"
                      : "ComputeSurfaceForceMap: the force field needs 3 components");
  const size_t len = m.size();
  if (field.getBlock().getSize() != num.getRho()->getBlock().getSize())
    errorMessage("ComputeSurfaceForceMap: the field's block differs from the lattice's");
  if (num.typeID() == acl::TYPE_DOUBLE) {
    std::vector<double> out(ncomp * len);
    check_abi(flux ? aslhip_surface_flux(num.handle(), m.data(), out.data())
                   : aslhip_surface_force(num.handle(), m.data(), out.data()));
    for (unsigned int c = 0; c < ncomp; ++c) std::copy(out.begin() + c * len, out.begin() + (c + 1) * len, field.values_[c].begin());
  } else {
    std::vector<float> out(ncomp * len);
    check_abi(flux ? aslhip_surface_flux(num.handle(), m.data(), out.data())
                   : aslhip_surface_force(num.handle(), m.data(), out.data()));
    for (unsigned int c = 0; c < ncomp; ++c)
      for (size_t i = 0; i < len; ++i) field.values_[c][i] = out[c * len + i];
  }
}
}  // namespace

void ComputeSurfaceFluxMap::execute() {
  surface_into(*num, *field, map_array(*num, map, "ComputeSurfaceFluxMap"), 1, true);
}

void ComputeSurfaceForceMap::execute() {
  surface_into(*num, *field, map_array(*num, map, "ComputeSurfaceForceMap"), 3, false);
}

SPNumMethod generateComputeSurfaceFlux(SPLBGK nm, SPDataWithGhostNodesACLData fF, SPAbstractDataWithGhostNodes map) {
  return std::make_shared<ComputeSurfaceFluxMap>(nm, fF, map);
}

SPNumMethod generateComputeSurfaceForce(SPLBGK nm, SPDataWithGhostNodesACLData fF, SPAbstractDataWithGhostNodes map) {
  return std::make_shared<ComputeSurfaceForceMap>(nm, fF, map);
}

SPBCond generateBCNoSlip(SPLBGK nm, const std::vector<SlicesNames> &sl) {
  auto bc = std::make_shared<BCNoSlip>(nm);
  addSlices(*bc, sl);
  return bc;
}

SPBCond generateBCConstantVelocity(SPLBGK nm, AVec<> v, const std::vector<SlicesNames> &sl) {
  auto bc = std::make_shared<BCConstantVelocity>(nm, acl::generateVEConstant(v));
  addSlices(*bc, sl);
  return bc;
}

SPBCond generateBCConstantPressure(SPLBGK nm, double p, const std::vector<SlicesNames> &sl) {
  auto bc = std::make_shared<BCConstantPressure>(nm, acl::generateVEConstant(p));
  addSlices(*bc, sl);
  return bc;
}

SPBCond generateBCConstantPressureVelocity(SPLBGK nm, double p, AVec<> v, const std::vector<SlicesNames> &sl) {
  auto bc = std::make_shared<BCConstantPressureVelocity>(nm, acl::generateVEConstant(p), acl::generateVEConstant(v));
  addSlices(*bc, sl);
  return bc;
}

SPNumMethod generateBCConstantPressure(SPLBGK nm, double p, SPAbstractDataWithGhostNodes map) {
  return std::make_shared<BCConstantPressureVelocityMap>(nm, acl::generateVEConstant(p), map);
}

SPNumMethod generateBCConstantPressureVelocity(SPLBGK nm, double p, AVec<> v, SPAbstractDataWithGhostNodes map) {
  return std::make_shared<BCConstantPressureVelocityMap>(nm, acl::generateVEConstant(p), acl::generateVEConstant(v),
                                                         map);
}

SPNumMethod generateBCNoSlip(SPLBGK nm, SPAbstractDataWithGhostNodes map) {
  return std::make_shared<BCNoSlipMap>(nm, map);
}

SPNumMethod generateBCNoSlipVel(SPLBGK nmU, SPAbstractDataWithGhostNodes map) {
  return std::make_shared<BCMacroFillMap>(nmU, ASLHIP_VIS_NOSLIP_VEL, map);
}

SPNumMethod generateBCNoSlipRho(SPLBGK nmU, SPAbstractDataWithGhostNodes map) {
  return std::make_shared<BCMacroFillMap>(nmU, ASLHIP_VIS_NOSLIP_RHO, map);
}

SPNumMethod generateBCVelocityVel(SPLBGK nm, SPPositionFunction v, SPAbstractDataWithGhostNodes map) {
  return std::make_shared<BCMacroFillMap>(nm, ASLHIP_VIS_VELOCITY, map, v);
}

SPNumMethod generateBCVelocity(SPLBGK nm, SPPositionFunction v, SPAbstractDataWithGhostNodes map) {
  return std::make_shared<BCVelocityMap>(nm, v, map);
}

SPNumMethod generateBCVelocity(SPLBGK nm, SPPositionFunction v, SPAbstractDataWithGhostNodes map,
                               SPAbstractDataWithGhostNodes computationalDomain) {
  return std::make_shared<BCVelocityMap>(nm, v, map, computationalDomain);
}

}  // namespace asl
