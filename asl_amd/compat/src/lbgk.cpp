// asl::LBGK and asl::LBGKUtilities over the C ABI (include/aslhip.h).
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include <num/aslLBGK.h>

namespace asl {
namespace detail {

void check_abi(int r) {
  if (r == 0) return;
  char buf[1024];
  aslhip_last_error(buf, sizeof buf);
  throw std::logic_error(std::strncmp(buf, "ASL ERROR", 9) == 0 ? std::string(buf) : "ASL ERROR: " + std::string(buf));
}

}  // namespace detail

namespace {

using detail::check_abi;

int q_of(const VectorTemplate *vT) {
  if (!vT) errorMessage("LBGK: no vector template");
  const int q = static_cast<int>(vT->vectors.size());
  const VectorTemplate *known = q == 15 ? &d3q15() : (q == 19 ? &d3q19() : (q == 27 ? &d3q27() : nullptr));
  if (!known || nD(*vT) != 3) errorMessage("LBGK: only d3q15(), d3q19() and d3q27() are supported");
  for (int i = 0; i < q; ++i)
    if (vT->vectors[i] != known->vectors[i]) errorMessage("LBGK: the vector template's direction order is not ASL's");
  return q;
}

// Ghost-inclusive cell count of the lattice on internal block b.
size_t lattice_len(const Block &b) { return static_cast<size_t>(productOfElements(offset(b, 1).getSize())); }

// A container that is to hold (or provide) one of the lattice's fields must
// have the lattice's ghost-inclusive size, or the library would read or write
// past its host buffers.
void check_len(const AbstractData &d, const Block &internal, const char *what) {
  const size_t len = lattice_len(internal);
  for (unsigned int c = 0; c < d.getNComponents(); ++c)
    if (d.values_[c].size() != len)
      errorMessage(std::string(what) + ": the container's block (ghost nodes included) is not the lattice's");
}

// A host view of a lattice field, fetched when the lattice has advanced.
// field 0: rho, 1: u, 2: f.  A reader that writes the view (ABD) commits it:
// populations go back to the lattice; rho and u are outputs of the next
// step and refuse.
void bind_view(AbstractData &d, aslhip_lattice *h, acl::TypeID t, int field, int ncomp, LBGK *owner) {
  unsigned long long seen = ~0ULL;
  d.store_ = [h, t, field, ncomp, owner](AbstractData &self) {
    if (field != 2) errorMessage("LBGK: getRho() and getVelocity() are outputs of the method; restore getF() instead");
    const size_t len = self.values_[0].size();
    if (t == acl::TYPE_DOUBLE) {
      std::vector<double> buf(len * ncomp);
      for (int c = 0; c < ncomp; ++c) std::copy(self.values_[c].begin(), self.values_[c].end(), buf.begin() + c * len);
      check_abi(aslhip_write_f(h, buf.data()));
    } else {
      std::vector<float> buf(len * ncomp);
      for (int c = 0; c < ncomp; ++c)
        for (size_t j = 0; j < len; ++j) buf[c * len + j] = static_cast<float>(self.values_[c][j]);
      check_abi(aslhip_write_f(h, buf.data()));
    }
    owner->execute_count_bump();
  };
  if (field != 2)
    d.device_ = [h, field](void *stream) -> const void * {
      const void *r = nullptr, *u = nullptr;
      check_abi(aslhip_macro_device(h, stream, &r, &u));
      return field == 0 ? r : u;
    };
  d.fetch_ = [h, t, field, ncomp, owner, seen](AbstractData &self) mutable {
    if (seen == owner->version()) return;
    const size_t len = self.values_[0].size();
    if (t == acl::TYPE_DOUBLE) {
      std::vector<double> buf(len * ncomp);
      check_abi(field == 0 ? aslhip_read_rho(h, buf.data())
                           : (field == 1 ? aslhip_read_velocity(h, buf.data()) : aslhip_read_f(h, buf.data())));
      for (int c = 0; c < ncomp; ++c) std::copy(buf.begin() + c * len, buf.begin() + (c + 1) * len, self.values_[c].begin());
    } else {
      std::vector<float> buf(len * ncomp);
      check_abi(field == 0 ? aslhip_read_rho(h, buf.data())
                           : (field == 1 ? aslhip_read_velocity(h, buf.data()) : aslhip_read_f(h, buf.data())));
      for (int c = 0; c < ncomp; ++c)
        for (size_t j = 0; j < len; ++j) self.values_[c][j] = double(buf[c * len + j]);
    }
    seen = owner->version();
  };
}

}  // namespace

LBGK::LBGK(Block b, Param nu, const VectorTemplate *vT, bool compVel, bool compRho, acl::CommandQueue)
    : vectorTemplate(vT) {
  if (nu.size() < 1) errorMessage("LBGK: empty viscosity parameter");
  create(b, nu.type == acl::TYPE_DOUBLE ? acl::TYPE_DOUBLE : acl::TYPE_FLOAT);
  nu_ = nu[0];
  check_abi(aslhip_set_viscosity(h_, nu_));
  check_abi(aslhip_set_compute_macro(h_, compRho ? 1 : 0, compVel ? 1 : 0));
}

LBGK::LBGK(DataD v, Param nu, const VectorTemplate *vT) : LBGK(v->getInternalBlock(), nu, vT) {
  if (v->getNComponents() != 3) errorMessage("LBGK: the velocity container must have 3 components");
  check_len(*v, v->getInternalBlock(), "LBGK");
  vel_ = v;
  bind_view(*vel_, h_, type_, 1, 3, this);
}

LBGKTurbulence::LBGKTurbulence(Block b, Param nu, const VectorTemplate *vT, bool compVel, bool compRho,
                               acl::CommandQueue queue)
    : LBGK(b, nu, vT, compVel, compRho, queue) {
  check_abi(aslhip_set_turbulence(h_, 1));
}

LBGKTurbulence::LBGKTurbulence(DataD v, Param nu, const VectorTemplate *vT) : LBGK(v, nu, vT) {
  check_abi(aslhip_set_turbulence(h_, 1));
}

void LBGK::create(const Block &b, acl::TypeID t) {
  if (nD(b) != 3) errorMessage("LBGK: only 3D blocks are supported");
  const int q = q_of(vectorTemplate);
  block_ = b;
  type_ = t;
  check_abi(aslhip_lattice_create(&h_, b.getSize()[0], b.getSize()[1], b.getSize()[2], q,
                                  t == acl::TYPE_DOUBLE ? ASLHIP_F64 : ASLHIP_F32, 0));
  // positions of the rotating velocity-map BCs: the block's dx and position
  const double pos[3] = {b.position[0], b.position[1], b.position[2]};
  check_abi(aslhip_set_geometry(h_, b.dx, pos));
  // opt-in defined value for velocity-map averages over outer-layer cells
  if (const char *e = std::getenv("ASLHIP_VMAP_SKIP_OUTER"))
    if (std::string(e) == "1") check_abi(aslhip_set_vmap_skip_outer(h_, 1));
  rho_ = std::make_shared<DataWithGhostNodesACLData>(b, 1, t);
  vel_ = std::make_shared<DataWithGhostNodesACLData>(b, 3, t);
  f_ = std::make_shared<DataWithGhostNodesACLData>(b, q, t);
  bind_view(*rho_, h_, t, 0, 1, this);
  bind_view(*vel_, h_, t, 1, 3, this);
  bind_view(*f_, h_, t, 2, q, this);
}

LBGK::~LBGK() {
  if (h_) aslhip_lattice_destroy(h_);
  h_ = nullptr;
}

void LBGK::init() {}

void LBGK::execute() {
  check_abi(aslhip_step(h_, 1, ASLHIP_STEP_DEFER));
  ++version_;
}

void LBGK::setViscosity(Param nu) {
  if (nu.size() < 1) errorMessage("LBGK::setViscosity: empty parameter");
  nu_ = nu[0];
  check_abi(aslhip_set_viscosity(h_, nu_));
}

double LBGK::getViscosity(unsigned int) { return nu_; }

// src/num/aslLBGK.cxx:252-255; the term itself: :207-208 (crossProduct
// needs a 3-component omega in 3D, aclVectorOfElementsOperations.cxx:514-535).
void LBGK::setOmega(Param w) {
  if (w.size() == 0) {
    check_abi(aslhip_set_omega(h_, nullptr));
    return;
  }
  if (w.size() != 3) errorMessage("crossProduct - two VectorOfElements have different sizes");
  const double om[3] = {w[0], w[1], w[2]};
  check_abi(aslhip_set_omega(h_, om));
}

void LBGK::setVectorTemplate(const VectorTemplate *vT) {
  if (q_of(vT) != q_of(vectorTemplate)) errorMessage("LBGK::setVectorTemplate: the lattice has another template");
  vectorTemplate = vT;
}

LBGK::Data LBGK::getF() { return f_; }
LBGK::DataD LBGK::getRho() { return rho_; }
LBGK::DataD LBGK::getVelocity() { return vel_; }

void LBGK::setCompressible(bool flag) {
  compressible_ = flag;
  check_abi(aslhip_set_compressible(h_, flag ? 1 : 0));
}

void LBGKUtilities::initF(Param rho, Param vel) {
  if (rho.size() != 1 || vel.size() != 3) errorMessage("LBGKUtilities::initF: expects a scalar rho and a 3-vector");
  const double u[3] = {vel[0], vel[1], vel[2]};
  check_abi(aslhip_init_equilibrium(num->handle(), rho[0], u));
  num->execute_count_bump();
}

void LBGKUtilities::initF(Param vel) { initF(acl::generateVEConstant(1.), vel); }

void LBGKUtilities::initF(SPAbstractDataWithGhostNodes rho, SPAbstractDataWithGhostNodes vel) {
  if (!vel || vel->getNComponents() != 3) errorMessage("LBGKUtilities::initF: velocity field needs 3 components");
  vel->refresh();
  check_len(*vel, num->block(), "LBGKUtilities::initF");
  if (rho) {
    rho->refresh();
    check_len(*rho, num->block(), "LBGKUtilities::initF");
  }
  const bool dbl = num->typeID() == acl::TYPE_DOUBLE;
  const size_t len = vel->component(0).size();
  std::vector<double> ud(3 * len), rd;
  for (int c = 0; c < 3; ++c) std::copy(vel->component(c).begin(), vel->component(c).end(), ud.begin() + c * len);
  if (rho) rd = rho->component(0);
  if (dbl) {
    check_abi(aslhip_init_equilibrium_field(num->handle(), rho ? rd.data() : nullptr, ud.data()));
  } else {
    std::vector<float> uf(ud.begin(), ud.end()), rf(rd.begin(), rd.end());
    check_abi(aslhip_init_equilibrium_field(num->handle(), rho ? rf.data() : nullptr, uf.data()));
  }
  num->execute_count_bump();
}

}  // namespace asl
