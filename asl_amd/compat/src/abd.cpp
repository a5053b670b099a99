// ABD operators on blocks and data containers (src/writers/aslABDFormat.cxx:35-81).
#include <writers/aslABDFormat.h>

#include <vector>

namespace asl {

ABDFileOut &operator<<(ABDFileOut &f, const Block &b) {
  f << b.dx << b.getSize() << b.position;
  return f;
}

ABDFileIn &operator>>(ABDFileIn &f, Block &b) {
  double dx = 0.;
  Block::V pos;
  Block::DV size;
  f >> dx >> size >> pos;
  b.dx = dx;
  b.position = pos;
  b.setSize(size);
  return f;
}

ABDFileOut &operator<<(ABDFileOut &f, const AbstractData &a) {
  // const view: the values of a lattice field are fetched by component()
  std::vector<double> v = const_cast<AbstractData &>(a).component(0);
  f << std::make_pair(v.data(), static_cast<unsigned int>(v.size()));
  return f;
}

// Like ASL's, these read component 0 only.  On a lattice view (getF()) the
// other components keep their current values and the populations are written
// back to the lattice; the rho and u views refuse the read (commit()).
ABDFileIn &operator>>(ABDFileIn &f, AbstractData &a) {
  a.refresh();
  std::vector<double> &v = a.values_.at(0);
  f >> std::make_pair(v.data(), static_cast<unsigned int>(v.size()));
  a.commit();
  return f;
}

ABDFileIn &get(ABDFileIn &f, AbstractData &a, std::shared_ptr<double> d) {
  a.refresh();
  const unsigned int s = static_cast<unsigned int>(a.values_.at(0).size());
  if (!d) d.reset(new double[s], std::default_delete<double[]>());
  f >> std::make_pair(d.get(), s);
  std::copy(d.get(), d.get() + s, a.values_[0].begin());
  a.commit();
  return f;
}

void writeABD(const std::string &fileName, const AbstractData &data, const std::string &name) {
  ABDFileOut f(fileName);
  if (!f) errorMessage("writeABD: cannot open " + fileName);
  AbstractData &a = const_cast<AbstractData &>(data);
  f << name << a.getBlock() << a.getNComponents();
  for (unsigned int c = 0; c < a.getNComponents(); ++c) {
    std::vector<double> v = a.component(c);
    f << std::make_pair(v.data(), static_cast<unsigned int>(v.size()));
  }
  if (!f) errorMessage("writeABD: cannot write " + fileName);
}

}  // namespace asl
