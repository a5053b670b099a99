// CPU restatement of the distance-function maps (test infrastructure only).
//
// ASL samples a DistanceFunction tree at every node of a block in the map's
// element type (initData / generateDataContainer_SP, src/aslGeomInc.cxx:
// 187-221; node positions from Index2PositionACL,
// src/math/aslIndex2Position.cxx:64-108).  The formulas
// (src/math/aslDistanceFunction.cxx:55-164):
//   sphere      sqrt(l2(p - c)) - r                               (:98-101)
//   cylinder    sqrt(l2(d - (d.o) o)) - r,  d = p - c             (:103-116)
//   cone        sqrt(l2(d - (d.o) o)) - tan(th) (d.o),  d = p - a (:118-131)
//   plane       p.n - b                                           (:133-143)
//   and / or    max / min                                         (:64-79)
//   inversion   -x                                                (:86-89)
//   normalize   clamp(x / dx / 1.8, -1, 1)                        (:145-164)
// evaluated here from the same postfix program the GPU reads, constants
// rounded to T where used, sums left to right, one rounding per operation.
#include <cmath>
#include <vector>

#include "lbgk_oracle.h"

namespace {

int nparam(int op) { return op == 1 || op == 4 ? 4 : (op == 2 || op == 3) ? 7 : op == 8 ? 1 : 0; }

template <class T>
T eval(const double *prog, int nwords, const T *p) {
  std::vector<T> st;
  for (int pc = 0; pc < nwords; pc += 1 + nparam(int(prog[pc]))) {
    const int op = int(prog[pc]);
    const double *a = prog + pc + 1;
    if (op == 1) {
      const T d0 = p[0] - T(a[0]), d1 = p[1] - T(a[1]), d2 = p[2] - T(a[2]);
      st.push_back(std::sqrt((d0 * d0 + d1 * d1) + d2 * d2) - T(a[3]));
    } else if (op == 2 || op == 3) {
      const T d[3] = {p[0] - T(a[0]), p[1] - T(a[1]), p[2] - T(a[2])};
      const T o[3] = {T(a[3]), T(a[4]), T(a[5])};
      const T proj = (d[0] * o[0] + d[1] * o[1]) + d[2] * o[2];
      T e[3];
      for (int k = 0; k < 3; ++k) e[k] = d[k] - proj * o[k];
      const T rad = std::sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]);
      st.push_back(op == 3 ? rad - T(a[6]) * proj : rad - T(a[6]));
    } else if (op == 4) {
      st.push_back(((p[0] * T(a[0]) + p[1] * T(a[1])) + p[2] * T(a[2])) - T(a[3]));
    } else if (op == 5 || op == 6) {
      const T y = st.back();
      st.pop_back();
      const T x = st.back();
      st.back() = op == 5 ? (x < y ? y : x) : (y < x ? y : x);
    } else if (op == 7) {
      st.back() = -st.back();
    } else if (op == 8) {
      const T dist = T(a[0]) * st.back();
      const T top = dist > T(1) ? T(1) : dist;
      st.back() = dist < T(-1) ? T(-1) : top;
    }
  }
  return st.back();
}

template <class T>
void run(const double *prog, int nwords, const int *size, const double *pos, double dx, T *out) {
  const T p0[3] = {T(pos[0]), T(pos[1]), T(pos[2])}, h = T(dx);
  long long j = 0;
  for (int x = 0; x < size[0]; ++x)
    for (int y = 0; y < size[1]; ++y)
      for (int z = 0; z < size[2]; ++z, ++j) {
        const T p[3] = {p0[0] + h * T(x), p0[1] + h * T(y), p0[2] + h * T(z)};
        out[j] = eval<T>(prog, nwords, p);
      }
}

}  // namespace

extern "C" void orc_df_map(const double *prog, int nwords, const int *size, const double *pos, double dx, int fp64,
                           void *out) {
  if (fp64)
    run<double>(prog, nwords, size, pos, dx, static_cast<double *>(out));
  else
    run<float>(prog, nwords, size, pos, dx, static_cast<float *>(out));
}
