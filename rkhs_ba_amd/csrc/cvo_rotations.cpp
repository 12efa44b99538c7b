// The 60 initial rotations of the global initialisation
// (gen_icosahedron_init_rots, src/cvo/CvoGPU.cpp:279-342): the rotation group
// of the icosahedron in the orientation of that table (a 3-fold axis along x,
// a 2-fold axis along y), in the table's order.
//
// The group is generated here, not copied: closure of the standard
// icosahedron's 5-fold rotation about (0, 1, phi) and 3-fold rotation about
// (1, 1, 1) in FP64, conjugated by the frame [a3; a2; -(a3 x a2)] with
// a3 = (1, 1, 1)/sqrt(3) and the perpendicular 2-fold axis
// a2 = (-1/phi, phi, -1)/2 (the mirror-image frame selects the table's
// chirality).  The elements are put in a canonical order (trace descending,
// then the row-major entries ascending) and kReferenceOrder lists, for each
// position of the reference's table, the canonical element it holds.  The
// reference's literals are float roundings of these matrices (they agree to
// 8.2e-7); the table here is the FP64 group rounded to float once.

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

#include "cvo_c.h"

namespace {

using M3 = std::array<double, 9>;

M3 mul(const M3& a, const M3& b) {
  M3 c{};
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) {
      double s = 0.0;
      for (int j = 0; j < 3; ++j) s += a[r * 3 + j] * b[j * 3 + k];
      c[r * 3 + k] = s;
    }
  return c;
}

M3 transpose(const M3& a) {
  M3 t{};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) t[c * 3 + r] = a[r * 3 + c];
  return t;
}

// Rodrigues' rotation by `angle` about `axis`.
M3 axis_angle(double x, double y, double z, double angle) {
  const double n = std::sqrt(x * x + y * y + z * z);
  x /= n; y /= n; z /= n;
  const M3 K = {0.0, -z, y, z, 0.0, -x, -y, x, 0.0};
  const M3 K2 = mul(K, K);
  const double s = std::sin(angle), c1 = 1.0 - std::cos(angle);
  M3 R{};
  for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0 ? 1.0 : 0.0) + s * K[i] + c1 * K2[i];
  return R;
}

bool same(const M3& a, const M3& b) {
  for (int i = 0; i < 9; ++i)
    if (std::fabs(a[i] - b[i]) > 1e-9) return false;
  return true;
}

// Distinct entries of the group's matrices differ by more than 0.01, so a
// tolerance of 1e-6 decides every comparison the same way FP64 rounding would.
bool canonical_less(const M3& a, const M3& b) {
  const double ta = a[0] + a[4] + a[8], tb = b[0] + b[4] + b[8];
  if (std::fabs(ta - tb) > 1e-6) return ta > tb;
  for (int i = 0; i < 9; ++i)
    if (std::fabs(a[i] - b[i]) > 1e-6) return a[i] < b[i];
  return false;
}

// Position i of the reference's table holds canonical element kReferenceOrder[i].
constexpr int kReferenceOrder[60] = {40, 2,  26, 17, 18, 53, 20, 54, 19, 43, 29, 5,  58, 9,  10, 39, 25, 1,  48, 33,
                                     34, 38, 50, 37, 30, 44, 6,  31, 32, 0,  57, 8,  7,  51, 14, 13, 47, 46, 45, 56,
                                     24, 23, 59, 12, 11, 27, 41, 3,  15, 52, 16, 36, 35, 49, 55, 22, 21, 42, 28, 4};

std::vector<M3> build_table() {
  const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
  const double pi = std::acos(-1.0);
  const M3 gens[2] = {axis_angle(0.0, 1.0, phi, 2.0 * pi / 5.0), axis_angle(1.0, 1.0, 1.0, 2.0 * pi / 3.0)};
  std::vector<M3> group = {M3{1, 0, 0, 0, 1, 0, 0, 0, 1}};
  std::vector<M3> frontier = group;
  while (!frontier.empty()) {
    std::vector<M3> next;
    for (const M3& m : frontier)
      for (const M3& g : gens) {
        const M3 p = mul(g, m);
        bool seen = false;
        for (const M3& q : group) seen = seen || same(p, q);
        if (!seen) {
          group.push_back(p);
          next.push_back(p);
        }
      }
    frontier.swap(next);
  }
  if (group.size() != 60) return {};
  const double s3 = std::sqrt(3.0);
  const double u3[3] = {1.0 / s3, 1.0 / s3, 1.0 / s3};
  const double u2[3] = {-0.5 / phi, 0.5 * phi, -0.5};
  const double w[3] = {-(u3[1] * u2[2] - u3[2] * u2[1]), -(u3[2] * u2[0] - u3[0] * u2[2]),
                       -(u3[0] * u2[1] - u3[1] * u2[0])};
  const M3 Q = {u3[0], u3[1], u3[2], u2[0], u2[1], u2[2], w[0], w[1], w[2]};
  const M3 Qt = transpose(Q);
  for (M3& m : group) m = mul(mul(Q, m), Qt);
  std::sort(group.begin(), group.end(), canonical_less);
  std::vector<M3> table(60);
  for (int i = 0; i < 60; ++i) table[i] = group[kReferenceOrder[i]];
  return table;
}

}  // namespace

extern "C" cvo_status cvo_icosahedral_rotations(float out[540]) {
  if (!out) return CVO_ERR_INVALID_ARG;
  static const std::vector<M3> table = build_table();
  if (table.size() != 60) return CVO_ERR_INVALID_ARG;
  for (int i = 0; i < 60; ++i)
    for (int k = 0; k < 9; ++k) out[9 * i + k] = (float)table[i][k];
  return CVO_OK;
}
