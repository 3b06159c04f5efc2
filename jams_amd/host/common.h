// common.h -- small vector types, constants and errors shared by the host
// layer.  Constants restate helpers/consts.h:18-38 of the reference.
#pragma once

#include <array>
#include <cmath>
#include <stdexcept>
#include <string>

namespace jams {

using Vec3 = std::array<double, 3>;
using Vec3i = std::array<int, 3>;
using Mat3 = std::array<std::array<double, 3>, 3>;

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kElectronGFactor = 2.0023193043625;
constexpr double kHBarIU = 0.6582119569;          // meV ps
constexpr double kBohrMagnetonIU = 0.0578838181;  // meV / T
constexpr double kGyromagneticRatioIU = kElectronGFactor * kBohrMagnetonIU / kHBarIU;
constexpr double kBoltzmannIU = 0.0861733326;     // meV / K
constexpr double kJoule2meV = 6.24150907e21;
constexpr double kmRyd2meV = 13.605693123;
constexpr double kVacuumPermeabilityIU = 4 * kPi * 1E-7 / kJoule2meV;
constexpr double kLatticeTolerance = 1e-4;        // helpers/defaults.h:44

class RuntimeError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

inline Vec3 operator+(const Vec3 &a, const Vec3 &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline Vec3 operator-(const Vec3 &a, const Vec3 &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline Vec3 operator-(const Vec3 &a) { return {-a[0], -a[1], -a[2]}; }
inline Vec3 operator*(double s, const Vec3 &a) { return {s * a[0], s * a[1], s * a[2]}; }
inline double dot(const Vec3 &a, const Vec3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double norm(const Vec3 &a) { return std::sqrt(dot(a, a)); }
inline Vec3 cross(const Vec3 &a, const Vec3 &b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline Vec3 operator*(const Mat3 &m, const Vec3 &v) {
  return {m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2],
          m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2],
          m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2]};
}
inline Mat3 operator*(const Mat3 &a, const Mat3 &b) {
  Mat3 r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) r[i][j] += a[i][k] * b[k][j];
  return r;
}
inline Mat3 transpose(const Mat3 &a) {
  Mat3 r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i][j] = a[j][i];
  return r;
}
inline Mat3 inverse(const Mat3 &m) {
  const double a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2],
               g = m[2][0], h = m[2][1], i = m[2][2];
  const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  if (det == 0.0) throw RuntimeError("singular unit-cell matrix");
  return Mat3{{{(e * i - f * h) / det, (c * h - b * i) / det, (b * f - c * e) / det},
               {(f * g - d * i) / det, (a * i - c * g) / det, (c * d - a * f) / det},
               {(d * h - e * g) / det, (b * g - a * h) / det, (a * e - b * d) / det}}};
}
inline Mat3 identity3() { return Mat3{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}}; }

// helpers/maths.h comparisons
inline bool approximately_equal(double a, double b, double eps) {
  const double d = std::abs(a - b);
  if (d <= eps) return true;
  return d <= std::max(std::abs(a), std::abs(b)) * eps;
}
inline bool approximately_equal(const Vec3 &a, const Vec3 &b, double eps) {
  for (int n = 0; n < 3; ++n)
    if (!approximately_equal(a[n], b[n], eps)) return false;
  return true;
}
inline bool approximately_zero(double a, double eps) { return std::abs(a) <= eps; }
inline bool definately_greater_than(double a, double b, double eps) {
  return (a - b) > std::max(std::abs(a), std::abs(b)) * eps;
}
inline bool definately_less_than(double a, double b, double eps) {
  return (b - a) > std::max(std::abs(a), std::abs(b)) * eps;
}

std::string lowercase(std::string s);
std::string capitalize(std::string s);

}  // namespace jams
