// lattice_consts.hpp — D3Q15 / D3Q19 / D3Q27 direction sets as compile-time
// tables, usable from host and device code.
//
// Direction order and weights follow ASL's VectorTemplate tables
// (src/math/aslTemplates.cxx:146-200): rest, +x, +y, +z, -x, -y, -z, then the
// edge directions, then (D3Q27) the corners.  Weights are ASL's
// quasiparticlesCoefs = lc_i / 6 with +1 on the rest direction
// (src/math/aslTemplates.cxx:38-70), evaluated in double exactly as there.
// D3Q27 has no weights in ASL (:187-200); the textbook 8/27, 2/27, 1/54,
// 1/216 are used (extension).  D3Q15 and D3Q19 also carry ASL's laplacian
// and gradient stencil coefficients (D3Q27 has none in ASL).
#pragma once

#if defined(__HIPCC__)
#define ASLHIP_HD __host__ __device__
#else
#define ASLHIP_HD
#endif

namespace aslhip {

template <int Q>
struct Lat;

template <>
struct Lat<15> {
  static constexpr int Q = 15;
  static constexpr int c[15][3] = {{0, 0, 0},  {1, 0, 0},   {0, 1, 0},   {0, 0, 1},   {-1, 0, 0},
                                   {0, -1, 0}, {0, 0, -1},  {1, 1, 1},   {1, 1, -1},  {1, -1, 1},
                                   {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
  static constexpr ASLHIP_HD double w(int i) {
    return i == 0 ? (-14. / 3.) / 6. + 1. : (i < 7 ? (2. / 3.) / 6. : (1. / 12.) / 6.);
  }
  // laplasCoefs and gradientCoefs (src/math/aslTemplates.cxx:154-163)
  static constexpr ASLHIP_HD double lc(int i) { return i == 0 ? -14. / 3. : (i < 7 ? 2. / 3. : 1. / 12.); }
  static constexpr ASLHIP_HD double gc(int i) { return i == 0 ? 0. : (i < 7 ? 1. / 3. : 1. / 24.); }
};

template <>
struct Lat<19> {
  static constexpr int Q = 19;
  static constexpr int c[19][3] = {{0, 0, 0},  {1, 0, 0},   {0, 1, 0},  {0, 0, 1},   {-1, 0, 0},
                                   {0, -1, 0}, {0, 0, -1},  {1, 1, 0},  {1, -1, 0},  {-1, -1, 0},
                                   {-1, 1, 0}, {0, 1, 1},   {0, 1, -1}, {0, -1, -1}, {0, -1, 1},
                                   {1, 0, 1},  {1, 0, -1},  {-1, 0, -1}, {-1, 0, 1}};
  static constexpr ASLHIP_HD double w(int i) {
    return i == 0 ? -4. / 6. + 1. : (i < 7 ? (1. / 3.) / 6. : (1. / 6.) / 6.);
  }
  // laplasCoefs and gradientCoefs (src/math/aslTemplates.cxx:177-182)
  static constexpr ASLHIP_HD double lc(int i) { return i == 0 ? -4. : (i < 7 ? 1. / 3. : 1. / 6.); }
  static constexpr ASLHIP_HD double gc(int i) { return i == 0 ? 0. : (i < 7 ? 1. / 6. : 1. / 12.); }
};

template <>
struct Lat<27> {
  static constexpr int Q = 27;
  static constexpr int c[27][3] = {
      {0, 0, 0},   {1, 0, 0},   {0, 1, 0},  {0, 0, 1},   {-1, 0, 0},  {0, -1, 0},  {0, 0, -1},
      {1, 1, 0},   {1, -1, 0},  {-1, -1, 0}, {-1, 1, 0}, {0, 1, 1},   {0, 1, -1},  {0, -1, -1},
      {0, -1, 1},  {1, 0, 1},   {1, 0, -1}, {-1, 0, -1}, {-1, 0, 1},  {1, 1, 1},   {1, 1, -1},
      {1, -1, 1},  {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
  static constexpr ASLHIP_HD double w(int i) {
    return i == 0 ? 8. / 27. : (i < 7 ? 2. / 27. : (i < 19 ? 1. / 54. : 1. / 216.));
  }
};

// Index of the opposite direction (ASL's invertVectors, aslTemplates.cxx:72-83).
template <int Q>
constexpr ASLHIP_HD int inv_dir(int i) {
  for (int j = 0; j < Q; ++j)
    if (Lat<Q>::c[j][0] == -Lat<Q>::c[i][0] && Lat<Q>::c[j][1] == -Lat<Q>::c[i][1] &&
        Lat<Q>::c[j][2] == -Lat<Q>::c[i][2])
      return j;
  return -1;
}

// Packed cell types (one byte per ghost-inclusive cell).
enum : unsigned char {
  CODE_FLUID = 0,     // computation node, every incoming link streams normally
  CODE_FLUID_BC = 1,  // computation node with BC-resolved links (see LinkRec)
  CODE_SKIP = 2,      // not computed (outer ghost layer, solid map cells)
  CODE_GHOST = 3      // computed but not a computation node (ASL collides it)
};

// How one incoming link is resolved.
enum : unsigned char {
  MODE_NONE = 0,    // normal streaming
  MODE_BOUNCE = 1,  // the receiving cell's own opposite population + addend
  MODE_CONST = 2,   // a constant population
  MODE_LINK = 3     // written between steps by the boundary-link kernel
};

// Per-pattern record: BC slot of each incoming link (0 = none) and the mask
// of outgoing links this cell must not write (their receivers resolve them).
struct LinkRec {
  unsigned char slot[28];
  unsigned int skip;
};
static_assert(sizeof(LinkRec) == 32, "LinkRec must be 32 bytes");

}  // namespace aslhip
