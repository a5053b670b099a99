// WriterVTKXML without VTK: <file>_<n>.vti, VTK XML ImageData with the data
// appended raw (little endian, UInt64 block headers), internal cells only,
// x index fastest as VTK expects (ASL's layout has z fastest).
#include <cstdint>
#include <fstream>

#include <writers/aslVTKFormatWriters.h>

namespace asl {

void WriterVTKXML::write() {
  if (scalarFields.empty() && vectorFields.empty()) errorMessage("WriterVTKXML::write: no fields");
  AbstractData *first = scalarFields.empty() ? vectorFields[0].second : scalarFields[0].second;
  const Block &gb = first->getBlock();
  const AVec<int> S = gb.getSize();
  const int g = dynamic_cast<AbstractDataWithGhostNodes *>(first) ? 1 : 0;
  const int n[3] = {S[0] - 2 * g, S[1] - 2 * g, S[2] - 2 * g};
  const double origin[3] = {gb.position[0] + g * gb.dx, gb.position[1] + g * gb.dx, gb.position[2] + g * gb.dx};
  const std::string name = file + "_" + numToStr(numOfWrites) + ".vti";
  std::ofstream out(name, std::ios::binary);
  if (!out) errorMessage("WriterVTKXML::write: cannot open " + name);
  struct Arr {
    std::string name;
    AbstractData *d;
    int nc;
  };
  std::vector<Arr> arrs;
  for (auto &s : scalarFields) arrs.push_back({s.first, s.second, 1});
  for (auto &v : vectorFields) arrs.push_back({v.first, v.second, 3});
  const uint64_t cells = uint64_t(n[0]) * n[1] * n[2];
  out << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
      << "<ImageData WholeExtent=\"0 " << n[0] - 1 << " 0 " << n[1] - 1 << " 0 " << n[2] - 1 << "\" Origin=\""
      << origin[0] << " " << origin[1] << " " << origin[2] << "\" Spacing=\"" << gb.dx << " " << gb.dx << " " << gb.dx
      << "\">\n<Piece Extent=\"0 " << n[0] - 1 << " 0 " << n[1] - 1 << " 0 " << n[2] - 1 << "\">\n<PointData>\n";
  uint64_t offset = 0;
  for (auto &a : arrs) {
    out << "<DataArray type=\"Float32\" Name=\"" << a.name << "\" NumberOfComponents=\"" << a.nc
        << "\" format=\"appended\" offset=\"" << offset << "\"/>\n";
    offset += 8 + cells * a.nc * 4;
  }
  out << "</PointData>\n</Piece>\n</ImageData>\n<AppendedData encoding=\"raw\">\n_";
  for (auto &a : arrs) {
    const uint64_t bytes = cells * a.nc * 4;
    out.write(reinterpret_cast<const char *>(&bytes), 8);
    std::vector<const std::vector<double> *> comp(a.nc);
    for (int c = 0; c < a.nc; ++c) comp[c] = &a.d->component(c);
    std::vector<float> buf;
    buf.reserve(cells * a.nc);
    for (int z = 0; z < n[2]; ++z)
      for (int y = 0; y < n[1]; ++y)
        for (int x = 0; x < n[0]; ++x) {
          const long j = (long(x + g) * S[1] + (y + g)) * S[2] + (z + g);
          for (int c = 0; c < a.nc; ++c) buf.push_back(float((*comp[c])[j]));
        }
    out.write(reinterpret_cast<const char *>(buf.data()), static_cast<std::streamsize>(buf.size() * 4));
  }
  out << "\n</AppendedData>\n</VTKFile>\n";
  ++numOfWrites;
}

}  // namespace asl
