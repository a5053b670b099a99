// ASL-style driver test in the shape of the locomotive drivers
// (examples/flow/locomotive.cc): a body read from an STL file with
// readSurface(file, Block), a tunnel from a level-set cylinder cut by a ground
// plane, LBGKTurbulence on D3Q15, a velocity map BC on the tunnel, a no-slip
// map on the body, pressure-velocity slices on X0/XE, the visualisation BCs,
// the surface force on the body and a .vti output.  Compared with the CPU
// oracle fed the same maps (bit for bit on computation nodes), and the body
// map with the oracle's distance map of the same triangles.
//
// usage: stl_body_parity file.stl|file.vtp [steps] [output-prefix]
#include <cmath>
#include <cstdlib>
#include <iostream>

#include <aslDataInc.h>
#include <aslGeomInc.h>
#include <acl/aclGenerators.h>
#include <math/aslPositionFunction.h>
#include <math/aslTemplates.h>
#include <num/aslLBGK.h>
#include <num/aslLBGKBC.h>
#include <readers/aslVTKFormatReaders.h>
#include <writers/aslVTKFormatWriters.h>

#include "../../oracle/lbgk_oracle.h"

typedef float FlT;

static asl::SPDistanceFunction tunnel_df(const asl::Block &b) {
  const double r = (b.getBPosition()[2] - b.position[2]) / 2.1;
  const double l = b.getBPosition()[0] - b.position[0] + b.dx;
  asl::AVec<> c(.5 * (b.getBPosition() + b.position));
  c[1] = b.position[1] + .25 * r;
  asl::AVec<> g(c);
  g[1] = b.position[1];
  return -(asl::generateDFCylinder(r, asl::makeAVec(l, 0., 0.), c) &
           asl::generateDFPlane(asl::makeAVec(0., -1., 0.), g));
}

int main(int argc, char *argv[]) {
  if (argc < 2) {
    std::cerr << "usage: stl_body_parity file.stl [steps] [output-prefix]" << std::endl;
    return EXIT_FAILURE;
  }
  const std::string stl = argv[1];
  const int steps = argc > 2 ? std::atoi(argv[2]) : 200;
  const double nu = 0.002, u0 = 0.1;

  // physical units as in the locomotive drivers: the tunnel's level set spans
  // several cells per unit, so its guard band shields the fluid from the
  // cells beyond it on the D3Q15 corner links too
  const double dx = 0.25;
  asl::Block bl(asl::makeAVec(48, 20, 24), dx, asl::makeAVec(-dx, -dx, -dx));
  auto body = asl::readSurface(stl, bl);
  asl::Block block(body->getInternalBlock());
  auto tunnel = asl::generateDataContainerACL_SP<FlT>(block, 1, 1u);
  asl::initData(tunnel, tunnel_df(block));
  auto force = asl::generateDataContainerACL_SP<FlT>(block, 3, 1u);
  asl::initData(force, asl::makeAVec(0., 0., 0.));

  asl::SPLBGK lbgk(new asl::LBGKTurbulence(block, acl::generateVEConstant(FlT(nu)), &asl::d3q15()));
  lbgk->init();
  asl::SPLBGKUtilities util(new asl::LBGKUtilities(lbgk));
  util->initF(acl::generateVEConstant(u0, .0, .0));
  auto vf = asl::generatePFConstant(asl::makeAVec(u0, 0., 0.));
  std::vector<asl::SPNumMethod> bc, bcV;
  bc.push_back(asl::generateBCVelocity(lbgk, vf, tunnel));
  bc.push_back(asl::generateBCNoSlip(lbgk, body));
  bc.push_back(asl::generateBCConstantPressureVelocity(lbgk, 1., asl::makeAVec(u0, 0., 0.), {asl::X0, asl::XE}));
  bcV.push_back(asl::generateBCVelocityVel(lbgk, vf, tunnel));
  bcV.push_back(asl::generateBCNoSlipRho(lbgk, tunnel));
  bcV.push_back(asl::generateBCNoSlipVel(lbgk, body));
  bcV.push_back(asl::generateBCNoSlipRho(lbgk, body));
  asl::initAll(bc);
  asl::initAll(bcV);
  auto computeForce = asl::generateComputeSurfaceForce(lbgk, force, body);
  computeForce->init();
  asl::executeAll(bc);
  for (int i = 0; i < steps; ++i) {
    lbgk->execute();
    asl::executeAll(bc);
  }
  computeForce->execute();

  const int nx = block.getSize()[0], ny = block.getSize()[1], nz = block.getSize()[2];
  const long S0 = nx + 2, S1 = ny + 2, S2 = nz + 2, len = S0 * S1 * S2;
  std::vector<double> rho = lbgk->getRho()->component(0);
  std::vector<std::vector<double>> vel(3), fg(3);
  for (int d = 0; d < 3; ++d) {
    vel[d] = lbgk->getVelocity()->component(d);
    fg[d] = force->component(d);
  }
  std::vector<unsigned char> types(len);
  aslhip_read_cell_types(lbgk->handle(), types.data());

  // the body map against the oracle's map of the same triangles
  long long ntri = 0;
  const bool vtp = stl.size() > 4 && stl.compare(stl.size() - 4, 4, ".vtp") == 0;
  if (vtp)
    aslhip_vtp_count(stl.c_str(), &ntri);
  else
    aslhip_stl_count(stl.c_str(), &ntri);
  std::vector<double> tri(9 * ntri);
  if (vtp)
    aslhip_vtp_read(stl.c_str(), tri.data(), ntri);
  else
    aslhip_stl_read(stl.c_str(), tri.data(), ntri);
  for (long long k = 0; k < 3 * ntri; ++k) std::swap(tri[3 * k], tri[3 * k + 2]);
  const int gsize[3] = {int(S0), int(S1), int(S2)};
  double gpos[3], gh[3];
  for (int d = 0; d < 3; ++d) {
    gpos[d] = bl.position[d];
    gh[d] = (bl.getBPosition()[d] - bl.position[d]) / (gsize[d] - 1);
  }
  std::vector<float> omap(len);
  orc_surface_distance_map(tri.data(), ntri, gsize, gpos, gh, 3.6 * dx, 1.2 * dx, dx, 1.8, omap.data());
  const std::vector<float> bm = body->componentAsFloat(0), tm = tunnel->componentAsFloat(0);
  bool map_ok = true;
  long shell = 0;
  for (long j = 0; j < len; ++j) {
    map_ok = map_ok && bm[j] == omap[j];
    shell += bm[j] <= 0.f;
  }

  void *O = orc_create(nx, ny, nz, 15, 0, 0);
  orc_set_viscosity(O, double(FlT(nu)));
  orc_set_turbulence(O, 1);
  const double zero[3] = {0., 0., 0.}, u[3] = {u0, 0., 0.};
  orc_init_equilibrium(O, 1., u);
  const int ids[3] = {orc_bc_add_map(O, 7, tm.data(), nullptr, 0., u, nullptr),
                      orc_bc_add_map(O, 5, bm.data(), nullptr, 0., zero, nullptr),
                      orc_bc_add_slice(O, 4, 1u | 2u, 1., u)};
  for (int b : ids) orc_bc_execute(O, b);
  orc_run(O, steps, ids, 3);
  std::vector<double> orho(len), ou(3 * len), of(3 * len);
  orc_read_rho(O, orho.data());
  orc_read_velocity(O, ou.data());
  orc_surface(O, bm.data(), 0, of.data());
  orc_destroy(O);

  bool bitwise = true;
  long fluid = 0;
  double umax = 0., fmax = 0.;
  for (int x = 1; x <= nx; ++x)
    for (int y = 1; y <= ny; ++y)
      for (int z = 1; z <= nz; ++z) {
        const long j = (x * S1 + y) * S2 + z;
        for (int d = 0; d < 3; ++d) {
          bitwise = bitwise && float(fg[d][j]) == float(of[d * len + j]);
          fmax = std::max(fmax, std::fabs(of[d * len + j]));
        }
        if (types[j] > 1) continue;
        ++fluid;
        bitwise = bitwise && float(rho[j]) == float(orho[j]);
        for (int d = 0; d < 3; ++d) {
          bitwise = bitwise && float(vel[d][j]) == float(ou[d * len + j]);
          umax = std::max(umax, std::fabs(ou[d * len + j]));
        }
      }

  // the visualisation BCs fill the ghost side of the maps on the host views
  asl::executeAll(bcV);
  if (argc > 3) {
    asl::WriterVTKXML writer(argv[3]);
    writer.addScalars("map", *body);
    writer.addScalars("tunnel", *tunnel);
    writer.addScalars("rho", *lbgk->getRho());
    writer.addVector("v", *lbgk->getVelocity());
    writer.addVector("force", *force);
    writer.write();
  }
  std::cout << "STL body " << ntri << " triangles, " << shell << " shell nodes; LBGKTurbulence D3Q15 " << nx << "x"
            << ny << "x" << nz << " x " << steps << ": " << fluid << " computation nodes, max|u| " << umax
            << ", max|force| " << fmax << "; map bit-identical: " << (map_ok ? "yes" : "no")
            << ", bit-identical: " << (bitwise ? "yes" : "no") << std::endl;
  return map_ok && bitwise && fluid > 0 && shell > 0 ? EXIT_SUCCESS : EXIT_FAILURE;
}
