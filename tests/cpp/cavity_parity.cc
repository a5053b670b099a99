// ASL-style driver test: the 64^3 lid-driven cavity (tau = 0.6, lid 0.05 on
// ZE, no-slip elsewhere, lid executed first as in examples/flow/flow2.cc)
// through the asl:: compatibility layer on the GPU, 1000 FP32 steps, against
// the CPU oracle.  Returns EXIT_SUCCESS if rho and u agree within 1e-5 (the
// port's contract) and reports whether they are bit-identical.  Also reads
// the velocity through getVelocity()->deviceData() (the device copy a coupled
// solver on the GPU consumes) and compares it with the host view.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdlib>
#include <iostream>

#include <aslDataInc.h>
#include <acl/aclGenerators.h>
#include <math/aslTemplates.h>
#include <num/aslLBGK.h>
#include <num/aslLBGKBC.h>

#include "../../oracle/lbgk_oracle.h"

typedef float FlT;

int main(int argc, char *argv[]) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 64;
  const int steps = argc > 2 ? std::atoi(argv[2]) : 1000;
  const double nu = (0.6 - 0.5) / 3.;
  asl::Block block(asl::makeAVec(n, n, n), 1.);
  asl::SPLBGK lbgk(new asl::LBGK(block, acl::generateVEConstant(FlT(nu)), &asl::d3q19()));
  lbgk->init();
  asl::SPLBGKUtilities util(new asl::LBGKUtilities(lbgk));
  util->initF(acl::generateVEConstant(.0, .0, .0));
  auto lid = asl::generateBCConstantVelocity(lbgk, asl::makeAVec(0.05, 0., 0.), {asl::ZE});
  auto walls = asl::generateBCNoSlip(lbgk, {asl::X0, asl::XE, asl::Y0, asl::YE, asl::Z0});
  lid->init();
  walls->init();
  lid->execute();
  walls->execute();
  for (int i = 0; i < steps; ++i) {
    lbgk->execute();
    lid->execute();
    walls->execute();
  }
  auto &rho = lbgk->getRho()->component(0);
  auto vel = lbgk->getVelocity();

  void *O = orc_create(n, n, n, 19, 0, 0);
  orc_set_viscosity(O, double(FlT(nu)));
  const double zero[3] = {0., 0., 0.}, ulid[3] = {0.05, 0., 0.};
  orc_init_equilibrium(O, 1., zero);
  const int order[2] = {orc_bc_add_slice(O, 3, 32u, 0., ulid), orc_bc_add_slice(O, 1, 31u, 0., zero)};
  orc_bc_execute(O, order[0]);
  orc_bc_execute(O, order[1]);
  orc_run(O, steps, order, 2);
  const long S = n + 2, len = S * S * S;
  std::vector<double> orho(len), ou(3 * len);
  orc_read_rho(O, orho.data());
  orc_read_velocity(O, ou.data());
  orc_destroy(O);

  double dr = 0., du = 0., umax = 0.;
  bool bitwise = true;
  for (int x = 1; x <= n; ++x)
    for (int y = 1; y <= n; ++y)
      for (int z = 1; z <= n; ++z) {
        const long j = (x * S + y) * S + z;
        dr = std::max(dr, std::fabs(rho[j] - orho[j]));
        bitwise = bitwise && float(rho[j]) == float(orho[j]);
        for (int d = 0; d < 3; ++d) {
          const double g = vel->component(d)[j], o = ou[d * len + j];
          du = std::max(du, std::fabs(g - o));
          umax = std::max(umax, std::fabs(o));
          bitwise = bitwise && float(g) == float(o);
        }
      }
  std::cout << "cavity " << n << "^3 x " << steps << " FP32 via asl::LBGK: max|drho| " << dr << ", max|du| " << du
            << " (max|u| " << umax << "), bit-identical: " << (bitwise ? "yes" : "no") << std::endl;
  // the device view: component-major, ghost-inclusive, FP32 on an FP32 lattice
  const void *dev = vel->deviceData(nullptr);
  std::vector<float> du_dev(3 * len);
  bool same_dev = dev != nullptr &&
                  hipMemcpy(du_dev.data(), dev, du_dev.size() * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
  for (int x = 1; x <= n && same_dev; ++x)
    for (int y = 1; y <= n; ++y)
      for (int z = 1; z <= n; ++z) {
        const long j = (x * S + y) * S + z;
        for (int d = 0; d < 3; ++d) same_dev = same_dev && du_dev[d * len + j] == float(vel->component(d)[j]);
      }
  std::cout << "device velocity view equals host view: " << (same_dev ? "yes" : "no") << std::endl;
  const bool ok = dr <= 1e-5 && du <= 1e-5 * umax && umax > 1e-3 && same_dev;
  return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}
