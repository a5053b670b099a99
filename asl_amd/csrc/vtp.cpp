// VTK XML PolyData (.vtp) surfaces for asl::readSurface: ASL reads them with
// vtkXMLPolyDataReader (src/readers/aslVTKFormatReaders.cxx:219-265) and hands
// the polydata to the same surfaceToData as STL surfaces.  This reader keeps
// what surfaceToData uses: the points and the polygon and triangle-strip
// cells of every piece, as triangles (a polygon as the fan from its first
// point, a strip as its consecutive triples), piece by piece, polygons before
// strips (VTK's cell order).  Vertex and line cells are not surfaces and are
// skipped.
//
// File layout handled (the VTK XML format): DataArray elements in ascii,
// binary (base64 inline) or appended (raw or base64) form; header_type
// UInt32 or UInt64; uncompressed data or vtkZLibDataCompressor blocks; either
// byte order.  Binary data is [header][bytes] with header = byte count, or
// [nblocks, block size, last block size, compressed sizes...] when
// compressed; VTK base64-encodes header and data as separate units, which the
// group-wise decoder below reads either way.
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aslhip.h"
#include "status.hpp"

namespace aslhip {
namespace {

struct VtpError {
  std::string msg;
};

[[noreturn]] void bad(const std::string &m) { throw VtpError{m}; }

// One XML start tag: its name, attributes and the text up to the matching
// end tag (for elements that hold data).
struct Tag {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attr;
  size_t begin = 0, end = 0;  // the tag itself: [begin, end)
  bool self_closing = false;
  const std::string *get(const char *k) const {
    for (auto &a : attr)
      if (a.first == k) return &a.second;
    return nullptr;
  }
};

// The next start tag at or after pos (comments, declarations and end tags are
// skipped); false at the end of the text or of the XML part (AppendedData's
// '_' marker is never reached: callers stop at that element).
bool next_tag(const std::string &s, size_t pos, size_t stop, Tag &t) {
  while (true) {
    const size_t lt = s.find('<', pos);
    if (lt == std::string::npos || lt >= stop) return false;
    if (s.compare(lt, 4, "<!--") == 0) {
      const size_t e = s.find("-->", lt);
      if (e == std::string::npos) bad("unterminated comment");
      pos = e + 3;
      continue;
    }
    if (lt + 1 < s.size() && (s[lt + 1] == '/' || s[lt + 1] == '?' || s[lt + 1] == '!')) {
      pos = lt + 1;
      continue;
    }
    const size_t gt = s.find('>', lt);
    if (gt == std::string::npos) bad("unterminated tag");
    t = Tag();
    t.begin = lt;
    t.end = gt + 1;
    t.self_closing = s[gt - 1] == '/';
    size_t i = lt + 1;
    while (i < gt && !isspace(static_cast<unsigned char>(s[i])) && s[i] != '/') t.name += s[i++];
    while (i < gt) {
      while (i < gt && (isspace(static_cast<unsigned char>(s[i])) || s[i] == '/')) ++i;
      if (i >= gt) break;
      std::string k;
      while (i < gt && s[i] != '=' && !isspace(static_cast<unsigned char>(s[i]))) k += s[i++];
      while (i < gt && s[i] != '"' && s[i] != '\'') ++i;
      if (i >= gt) bad("attribute without a quoted value in <" + t.name + ">");
      const char q = s[i++];
      const size_t e = s.find(q, i);
      if (e == std::string::npos || e > gt) bad("unterminated attribute in <" + t.name + ">");
      t.attr.emplace_back(k, s.substr(i, e - i));
      i = e + 1;
    }
    return true;
  }
}

// End of the element opened by t (position of its "</name").
size_t element_end(const std::string &s, const Tag &t) {
  if (t.self_closing) return t.end;
  const size_t e = s.find("</" + t.name, t.end);
  if (e == std::string::npos) bad("<" + t.name + "> is not closed");
  return e;
}

int b64_value(char c) {
  if (c >= 'A' && c <= 'Z') return c - 'A';
  if (c >= 'a' && c <= 'z') return c - 'a' + 26;
  if (c >= '0' && c <= '9') return c - '0' + 52;
  if (c == '+') return 62;
  if (c == '/') return 63;
  return -1;
}

// Base64 text read four characters at a time; a group with '=' padding
// yields fewer bytes, and decoding goes on with the next group (header and
// data encoded as separate units).  Whitespace is skipped.
struct B64 {
  const std::string &s;
  size_t pos, stop;
  std::vector<unsigned char> out;
  B64(const std::string &str, size_t p, size_t e) : s(str), pos(p), stop(e) {}
  // decodes until out holds at least n bytes
  void need(size_t n) {
    while (out.size() < n) {
      int v[4], got = 0, pad = 0;
      while (got < 4) {
        if (pos >= stop) bad("base64 data ends early");
        const char c = s[pos++];
        if (isspace(static_cast<unsigned char>(c))) continue;
        if (c == '=') {
          v[got++] = 0;
          ++pad;
          continue;
        }
        const int x = b64_value(c);
        if (x < 0 || pad) bad("invalid base64 data");
        v[got++] = x;
      }
      const unsigned w = (v[0] << 18) | (v[1] << 12) | (v[2] << 6) | v[3];
      const unsigned char b[3] = {static_cast<unsigned char>(w >> 16), static_cast<unsigned char>(w >> 8),
                                  static_cast<unsigned char>(w)};
      if (pad > 2) bad("invalid base64 padding");
      out.insert(out.end(), b, b + 3 - pad);
    }
  }
};

struct Format {
  bool big_endian = false;
  int header_bytes = 4;  // UInt32 or UInt64
  bool zlib = false;
  size_t appended = std::string::npos;  // first byte after '_'
  bool appended_raw = true;
  size_t appended_stop = 0;
};

uint64_t read_uint(const unsigned char *p, int n, bool big) {
  uint64_t v = 0;
  for (int i = 0; i < n; ++i) v |= static_cast<uint64_t>(p[big ? n - 1 - i : i]) << (8 * i);
  return v;
}

// The bytes of one binary DataArray (inline base64 or appended), header read
// and blocks inflated.  src: the source of raw or decoded bytes on demand.
template <class Need>
std::vector<unsigned char> binary_payload(const Format &f, Need need) {
  const int H = f.header_bytes;
  if (!f.zlib) {
    const unsigned char *h = need(H);
    const uint64_t n = read_uint(h, H, f.big_endian);
    const unsigned char *d = need(H + n);
    return std::vector<unsigned char>(d + H, d + H + n);
  }
  const unsigned char *h = need(3 * H);
  const uint64_t nb = read_uint(h, H, f.big_endian);
  const uint64_t bs = read_uint(h + H, H, f.big_endian);
  const uint64_t last = read_uint(h + 2 * H, H, f.big_endian);
  if (nb > (1ull << 32)) bad("implausible block count");
  h = need((3 + nb) * H);
  std::vector<uint64_t> cs(nb);
  uint64_t total = 0;
  for (uint64_t b = 0; b < nb; ++b) total += cs[b] = read_uint(h + (3 + b) * H, H, f.big_endian);
  const size_t hdr = (3 + nb) * H;
  const unsigned char *d = need(hdr + total);
  std::vector<unsigned char> out;
  uint64_t off = hdr;
  for (uint64_t b = 0; b < nb; ++b) {
    uLongf usize = static_cast<uLongf>(b + 1 == nb && last ? last : bs);
    const size_t at = out.size();
    out.resize(at + usize);
    if (uncompress(out.data() + at, &usize, d + off, static_cast<uLong>(cs[b])) != Z_OK)
      bad("corrupt zlib block");
    out.resize(at + usize);
    off += cs[b];
  }
  return out;
}

// Numbers of one DataArray as doubles (coordinates) or 64-bit integers
// (connectivity, offsets).
template <class T>
std::vector<T> data_array(const std::string &s, const Tag &t, const Format &f) {
  const std::string *type = t.get("type");
  const std::string *fmt = t.get("format");
  if (!type) bad("<DataArray> without a type");
  const std::string form = fmt ? *fmt : "ascii";
  int size = 0;
  bool is_float = false, is_signed = true;
  if (*type == "Float32") size = 4, is_float = true;
  else if (*type == "Float64") size = 8, is_float = true;
  else if (*type == "Int8" || *type == "UInt8") size = 1, is_signed = *type == "Int8";
  else if (*type == "Int16" || *type == "UInt16") size = 2, is_signed = *type == "Int16";
  else if (*type == "Int32" || *type == "UInt32") size = 4, is_signed = *type == "Int32";
  else if (*type == "Int64" || *type == "UInt64") size = 8, is_signed = *type == "Int64";
  else bad("unsupported DataArray type " + *type);
  std::vector<T> v;
  if (form == "ascii") {
    const size_t e = element_end(s, t);
    const std::string text = s.substr(t.end, e - t.end);
    const char *p = text.c_str();
    char *q = nullptr;
    while (true) {
      while (*p && isspace(static_cast<unsigned char>(*p))) ++p;
      if (!*p) break;
      const double x = strtod(p, &q);
      if (q == p) bad("invalid number in an ascii DataArray");
      if (is_float && size == 4) v.push_back(static_cast<T>(static_cast<float>(x)));
      else v.push_back(static_cast<T>(x));
      p = q;
    }
    return v;
  }
  std::vector<unsigned char> bytes;
  if (form == "binary") {
    B64 b(s, t.end, element_end(s, t));
    bytes = binary_payload(f, [&](size_t n) {
      b.need(n);
      return b.out.data();
    });
  } else if (form == "appended") {
    const std::string *o = t.get("offset");
    if (!o || f.appended == std::string::npos) bad("appended DataArray without AppendedData");
    const size_t at = f.appended + static_cast<size_t>(strtoull(o->c_str(), nullptr, 10));
    if (f.appended_raw) {
      bytes = binary_payload(f, [&](size_t n) {
        if (at + n > s.size()) bad("appended data ends early");
        return reinterpret_cast<const unsigned char *>(s.data() + at);
      });
    } else {
      B64 b(s, at, f.appended_stop);
      bytes = binary_payload(f, [&](size_t n) {
        b.need(n);
        return b.out.data();
      });
    }
  } else {
    bad("unsupported DataArray format " + form);
  }
  if (bytes.size() % size) bad("DataArray byte count is not a multiple of its element size");
  const size_t n = bytes.size() / size;
  v.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const uint64_t u = read_uint(bytes.data() + i * size, size, f.big_endian);
    if (is_float) {
      if (size == 4) {
        uint32_t w = static_cast<uint32_t>(u);
        float x;
        std::memcpy(&x, &w, 4);
        v[i] = static_cast<T>(x);
      } else {
        double x;
        std::memcpy(&x, &u, 8);
        v[i] = static_cast<T>(x);
      }
    } else if (is_signed) {
      const int sh = 64 - 8 * size;
      v[i] = static_cast<T>(static_cast<int64_t>(u << sh) >> sh);
    } else {
      v[i] = static_cast<T>(u);
    }
  }
  return v;
}

// The first DataArray inside [from, stop) with the given Name (any if name
// is null).
bool find_array(const std::string &s, size_t from, size_t stop, const char *name, Tag &out) {
  Tag t;
  size_t pos = from;
  while (next_tag(s, pos, stop, t)) {
    pos = t.end;
    if (t.name != "DataArray") continue;
    const std::string *nm = t.get("Name");
    if (!name || (nm && *nm == name)) {
      out = t;
      return true;
    }
  }
  return false;
}

// All triangles of the file (9 doubles each, the file's x y z).
std::vector<double> vtp_triangles(const char *path) {
  FILE *fp = std::fopen(path, "rb");
  if (!fp) bad(std::string("cannot open ") + path);
  std::string s;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof buf, fp)) > 0) s.append(buf, got);
  std::fclose(fp);

  Format f;
  Tag t;
  size_t pos = 0;
  // <VTKFile type="PolyData" ...>
  bool found = false;
  while (next_tag(s, pos, s.size(), t)) {
    pos = t.end;
    if (t.name == "VTKFile") {
      found = true;
      break;
    }
  }
  if (!found) bad("not a VTK XML file");
  const std::string *type = t.get("type");
  if (!type || *type != "PolyData") bad("not a PolyData file");
  if (const std::string *bo = t.get("byte_order")) f.big_endian = *bo == "BigEndian";
  if (const std::string *ht = t.get("header_type")) {
    if (*ht == "UInt64") f.header_bytes = 8;
    else if (*ht != "UInt32") bad("unsupported header_type " + *ht);
  }
  if (const std::string *c = t.get("compressor")) {
    if (*c == "vtkZLibDataCompressor") f.zlib = true;
    else if (!c->empty()) bad("unsupported compressor " + *c);
  }
  // the XML part ends where AppendedData's raw bytes begin
  size_t xml_end = s.size();
  const size_t ad = s.find("<AppendedData", pos);
  if (ad != std::string::npos) {
    Tag a;
    if (!next_tag(s, ad, s.size(), a)) bad("broken <AppendedData>");
    const std::string *enc = a.get("encoding");
    f.appended_raw = !enc || *enc == "raw";
    const size_t us = s.find('_', a.end);
    if (us == std::string::npos) bad("<AppendedData> without its '_' marker");
    f.appended = us + 1;
    f.appended_stop = f.appended_raw ? s.size() : s.find("</AppendedData", f.appended);
    if (f.appended_stop == std::string::npos) f.appended_stop = s.size();
    xml_end = ad;
  }

  std::vector<double> tri;
  while (next_tag(s, pos, xml_end, t)) {
    pos = t.end;
    if (t.name != "Piece") continue;
    const size_t piece_end = element_end(s, t);
    std::vector<double> pts;
    std::vector<double> ptri[2];  // polygons, strips (VTK's cell order within a piece)
    Tag e;
    size_t p2 = t.end;
    while (next_tag(s, p2, piece_end, e)) {
      const size_t ee = element_end(s, e);
      p2 = e.self_closing ? e.end : ee;
      if (e.name == "Points") {
        Tag arr;
        if (!find_array(s, e.end, ee, nullptr, arr)) bad("<Points> without a DataArray");
        const std::string *nc = arr.get("NumberOfComponents");
        if (nc && std::atoi(nc->c_str()) != 3) bad("points need 3 components");
        pts = data_array<double>(s, arr, f);
        if (pts.size() % 3) bad("point coordinates are not triples");
      } else if (e.name == "Polys" || e.name == "Strips") {
        Tag ca, oa;
        if (!find_array(s, e.end, ee, "connectivity", ca) || !find_array(s, e.end, ee, "offsets", oa))
          bad("<" + e.name + "> needs connectivity and offsets");
        const std::vector<long long> con = data_array<long long>(s, ca, f);
        const std::vector<long long> offs = data_array<long long>(s, oa, f);
        const long long np = static_cast<long long>(pts.size() / 3);
        if (pts.empty() && !offs.empty()) bad("cells before the points of their piece");
        long long first = 0;
        for (long long end : offs) {
          if (end < first || end > static_cast<long long>(con.size())) bad("cell offsets out of order");
          auto vert = [&](long long i) -> const double * {
            const long long id = con[i];
            if (id < 0 || id >= np) bad("cell point index out of range");
            return &pts[3 * id];
          };
          const long long m = end - first;
          std::vector<double> &dst = ptri[e.name == "Polys" ? 0 : 1];
          for (long long k = 0; k + 2 < m; ++k) {
            const double *a = e.name == "Polys" ? vert(first) : vert(first + k);
            const double *b = vert(first + k + 1);
            const double *c = vert(first + k + 2);
            dst.insert(dst.end(), a, a + 3);
            dst.insert(dst.end(), b, b + 3);
            dst.insert(dst.end(), c, c + 3);
          }
          first = end;
        }
      }
    }
    for (auto &pt : ptri) tri.insert(tri.end(), pt.begin(), pt.end());
    pos = piece_end;
  }
  return tri;
}

int vtp_parse(const char *path, double *tri, long long cap, long long *ntri) {
  if (!path) return fail(ASLHIP_EINVAL, "ASL ERROR: readSurface: no file name");
  try {
    const std::vector<double> t = vtp_triangles(path);
    const long long n = static_cast<long long>(t.size() / 9);
    if (ntri) *ntri = n;
    if (tri) {
      if (cap != n) return fail(ASLHIP_EINVAL, "ASL ERROR: readSurface: %s holds %lld triangles, not %lld", path, n, cap);
      std::memcpy(tri, t.data(), t.size() * sizeof(double));
    }
    return 0;
  } catch (const VtpError &e) {
    return fail(ASLHIP_EINVAL, "ASL ERROR: readSurface: %s: %s", path, e.msg.c_str());
  }
}

}  // namespace
}  // namespace aslhip

extern "C" {

int aslhip_vtp_count(const char *path, long long *ntri) {
  if (!ntri) return aslhip::fail(ASLHIP_EINVAL, "ASL ERROR: aslhip_vtp_count needs an output");
  return aslhip::vtp_parse(path, nullptr, 0, ntri);
}

int aslhip_vtp_read(const char *path, double *tri, long long ntri) {
  if (!tri && ntri) return aslhip::fail(ASLHIP_EINVAL, "ASL ERROR: aslhip_vtp_read needs an output");
  return aslhip::vtp_parse(path, tri, ntri, nullptr);
}

}  // extern "C"
