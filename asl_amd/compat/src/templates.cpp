// Lattice templates of ASL's VectorTemplate (src/math/aslTemplates.cxx:38-200):
// quasiparticlesCoefs = lc/6 with +1 on the rest direction, evaluated in
// double as there; invertVectors by pairwise search.
#include <math/aslTemplates.h>

namespace asl {
namespace {

VectorTemplate build(const int (*c)[3], const double *lc, int n, const double *w_exact) {
  VectorTemplate t;
  for (int i = 0; i < n; ++i) {
    t.vectors.push_back(makeAVec(c[i][0], c[i][1], c[i][2]));
    t.laplasCoefs.push_back(lc[i]);
    t.quasiparticlesCoefs.push_back(w_exact ? w_exact[i] : lc[i] / 6.);
  }
  if (!w_exact) t.quasiparticlesCoefs[0] += 1.;
  t.invertVectors.assign(n, 0);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      if (t.vectors[i] == -t.vectors[j]) {
        t.invertVectors[i] = j;
        t.invertVectors[j] = i;
      }
  return t;
}

const int C27[27][3] = {{0, 0, 0},   {1, 0, 0},   {0, 1, 0},  {0, 0, 1},   {-1, 0, 0},  {0, -1, 0},  {0, 0, -1},
                        {1, 1, 0},   {1, -1, 0},  {-1, -1, 0}, {-1, 1, 0}, {0, 1, 1},   {0, 1, -1},  {0, -1, -1},
                        {0, -1, 1},  {1, 0, 1},   {1, 0, -1}, {-1, 0, -1}, {-1, 0, 1},  {1, 1, 1},   {1, 1, -1},
                        {1, -1, 1},  {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};

}  // namespace

const VectorTemplate &d3q15() {
  static const int c[15][3] = {{0, 0, 0},  {1, 0, 0},   {0, 1, 0},   {0, 0, 1},   {-1, 0, 0},
                               {0, -1, 0}, {0, 0, -1},  {1, 1, 1},   {1, 1, -1},  {1, -1, 1},
                               {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
  static const double lc[15] = {-14. / 3., 2. / 3.,  2. / 3.,  2. / 3.,  2. / 3.,  2. / 3.,  2. / 3., 1. / 12.,
                                1. / 12.,  1. / 12., 1. / 12., 1. / 12., 1. / 12., 1. / 12., 1. / 12.};
  static const VectorTemplate t = build(c, lc, 15, nullptr);
  return t;
}

const VectorTemplate &d3q19() {
  static const double lc[19] = {-4.,     1. / 3., 1. / 3., 1. / 3., 1. / 3., 1. / 3., 1. / 3.,
                                1. / 6., 1. / 6., 1. / 6., 1. / 6., 1. / 6., 1. / 6., 1. / 6.,
                                1. / 6., 1. / 6., 1. / 6., 1. / 6., 1. / 6.};
  static const VectorTemplate t = build(C27, lc, 19, nullptr);
  return t;
}

const VectorTemplate &d3q27() {
  // ASL builds d3q27() without weights; the textbook D3Q27 set is used.
  static double w[27], lc[27];
  static bool done = false;
  if (!done) {
    for (int i = 0; i < 27; ++i) {
      w[i] = i == 0 ? 8. / 27. : (i < 7 ? 2. / 27. : (i < 19 ? 1. / 54. : 1. / 216.));
      lc[i] = i == 0 ? 6. * (w[i] - 1.) : 6. * w[i];
    }
    done = true;
  }
  static const VectorTemplate t = build(C27, lc, 27, w);
  return t;
}

}  // namespace asl
