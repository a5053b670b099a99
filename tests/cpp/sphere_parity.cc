// ASL-style driver test: channel flow past a level-set sphere (the
// 256x128x128 benchmark flow at reduced size, or full size with arguments)
// through the asl:: compatibility layer: map from generateDFSphere +
// normalize, velocity inlet X0, pressure outlet XE, no-slip side walls,
// no-slip map.  Compared with the CPU oracle fed the same map; also checks
// that the lattice's cell types match the map's ghost / computation masks.
#include <cmath>
#include <cstdlib>
#include <iostream>

#include <aslDataInc.h>
#include <aslGeomInc.h>
#include <acl/aclGenerators.h>
#include <math/aslTemplates.h>
#include <num/aslLBGK.h>
#include <num/aslLBGKBC.h>

#include "../../oracle/lbgk_oracle.h"

typedef float FlT;

int main(int argc, char *argv[]) {
  const int nx = argc > 1 ? std::atoi(argv[1]) : 64, ny = argc > 2 ? std::atoi(argv[2]) : 32;
  const int nz = ny, steps = argc > 3 ? std::atoi(argv[3]) : 300;
  const double r = nx / 16., nu = (0.56 - 0.5) / 3., uin = 0.04;
  asl::Block block(asl::makeAVec(nx, ny, nz), 1.);
  auto ball = asl::normalize(asl::generateDFSphere(r, asl::makeAVec(nx / 4., ny / 2., nz / 2.)), 1.);
  auto map = asl::generateDataContainer_SP(block, ball, 1u, acl::typeToTypeID<FlT>());
  asl::SPLBGK lbgk(new asl::LBGK(block, acl::generateVEConstant(FlT(nu)), &asl::d3q19()));
  lbgk->init();
  asl::SPLBGKUtilities(new asl::LBGKUtilities(lbgk))->initF(acl::generateVEConstant(uin, .0, .0));
  std::vector<asl::SPNumMethod> bc;
  bc.push_back(asl::generateBCConstantVelocity(lbgk, asl::makeAVec(uin, 0., 0.), {asl::X0}));
  bc.push_back(asl::generateBCConstantPressure(lbgk, 1., {asl::XE}));
  bc.push_back(asl::generateBCNoSlip(lbgk, {asl::Y0, asl::YE, asl::Z0, asl::ZE}));
  bc.push_back(asl::generateBCNoSlip(lbgk, map));
  asl::initAll(bc);
  asl::executeAll(bc);
  for (int i = 0; i < steps; ++i) {
    lbgk->execute();
    asl::executeAll(bc);
  }
  auto &rho = lbgk->getRho()->component(0);
  auto vel = lbgk->getVelocity();
  const long S0 = nx + 2, S1 = ny + 2, S2 = nz + 2, len = S0 * S1 * S2;
  std::vector<unsigned char> types(len);
  aslhip_read_cell_types(lbgk->handle(), types.data());

  const std::vector<float> m = map->componentAsFloat(0);
  void *O = orc_create(nx, ny, nz, 19, 0, 0);
  orc_set_viscosity(O, double(FlT(nu)));
  const double zero[3] = {0., 0., 0.}, u[3] = {uin, 0., 0.};
  orc_init_equilibrium(O, 1., u);
  const int ids[4] = {orc_bc_add_slice(O, 3, 1u, 0., u), orc_bc_add_slice(O, 2, 2u, 1., zero),
                      orc_bc_add_slice(O, 1, 4u | 8u | 16u | 32u, 0., zero),
                      orc_bc_add_map(O, 5, m.data(), nullptr, 0., zero, nullptr)};
  for (int b : ids) orc_bc_execute(O, b);
  orc_run(O, steps, ids, 4);
  std::vector<double> orho(len), ou(3 * len);
  std::vector<uint8_t> masks(len);
  orc_read_rho(O, orho.data());
  orc_read_velocity(O, ou.data());
  orc_map_masks(O, ids[3], masks.data());
  orc_destroy(O);

  double dr = 0., du = 0., umax = 0.;
  bool bitwise = true, types_ok = true;
  long fluid = 0;
  for (int x = 0; x < S0; ++x)
    for (int y = 0; y < S1; ++y)
      for (int z = 0; z < S2; ++z) {
        const long j = (x * S1 + y) * S2 + z;
        const bool interior = x >= 1 && x <= nx && y >= 1 && y <= ny && z >= 1 && z <= nz;
        const bool comp = interior && (masks[j] & 4);
        // cell types: computation nodes are fluid (0/1), every other cell is
        // not computed (2) with a normalised map
        types_ok = types_ok && (comp ? types[j] <= 1 : types[j] == 2);
        if (!comp) continue;
        ++fluid;
        dr = std::max(dr, std::fabs(rho[j] - orho[j]));
        bitwise = bitwise && float(rho[j]) == float(orho[j]);
        for (int d = 0; d < 3; ++d) {
          const double g = vel->component(d)[j], o = ou[d * len + j];
          du = std::max(du, std::fabs(g - o));
          umax = std::max(umax, std::fabs(o));
          bitwise = bitwise && float(g) == float(o);
        }
      }
  std::cout << "sphere " << nx << "x" << ny << "x" << nz << " x " << steps << " FP32: " << fluid
            << " computation nodes, max|drho| " << dr << ", max|du| " << du << " (max|u| " << umax
            << "), bit-identical: " << (bitwise ? "yes" : "no") << ", cell types match masks: "
            << (types_ok ? "yes" : "no") << std::endl;
  return (dr <= 1e-5 && du <= 1e-5 * umax && types_ok) ? EXIT_SUCCESS : EXIT_FAILURE;
}
