// Potential types and impurities of the reference (potentialtypes.hpp:72-382,
// potentialtypes.cpp:56-500, noise.cpp:25-110), evaluated on the cell-centred
// grid (datalayout.cpp:21-36). Shared by the gfx950 kernels (values on the
// device) and the host (parameter resolution). The formulas define the
// potentials the reference's CLI builds, so they are kept as they are.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstring>
#include "line_ops.hpp"

namespace itp {

enum PotKind {
  kPotZero = 0, kPotHarmonic, kPotElliptic, kPotPrettyHardSquare, kPotSoftPentagon, kPotHenonHeiles,
  kPotGaussian, kPotQuartic, kPotSquareOscillator, kPotPower, kPotRing, kPotCosh, kPotSoftStadium,
  kPotPowerStadium, kPotNumKinds
};

enum ImpKind { kImpGaussian = 0, kImpCoulomb, kImpHemisphere, kImpDelta, kImpNumKinds };

constexpr int kPotMaxParams = 5;

// Resolved parameters of a potential kind (defaults filled in).
struct PotParams {
  int kind;
  double p[kPotMaxParams];
};

// Value at (x, y) (PotentialType::operator(), potentialtypes.hpp:72-382).
ITP_HD double pot_value(const PotParams& q, double x, double y) {
  const double* p = q.p;
  switch (q.kind) {
    case kPotHarmonic: {  // 0.5 A ((x-x0)^2 + (y-y0)^2)
      const double px = x - p[1], py = y - p[2];
      return 0.5 * p[0] * (px * px + py * py);
    }
    case kPotElliptic:  // 0.5 (Ax x^2 + Ay y^2)
      return 0.5 * (p[0] * x * x + p[1] * y * y);
    case kPotPrettyHardSquare: {  // max(|2x/pi|, |2y/pi|)^e
      const double ax = std::fabs(2 * x / M_PI), ay = std::fabs(2 * y / M_PI);
      return std::pow(ax > ay ? ax : ay, p[0]);
    }
    case kPotSoftPentagon: {  // r^2/2 (1 + sin(5 t)/2)
      const double r2 = x * x + y * y, t = std::atan2(y, x);
      return 0.5 * r2 * (1 + 0.5 * std::sin(5 * t));
    }
    case kPotHenonHeiles: {  // r^4 + a r^2 + b r^3 cos(3 t)
      const double r2 = x * x + y * y, r3 = std::pow(r2, 3.0 / 2.0), r4 = r2 * r2, t = std::atan2(y, x);
      return r4 + p[0] * r2 + p[1] * r3 * std::cos(3 * t);
    }
    case kPotGaussian: {  // A exp(-((x-x0)^2 + (y-y0)^2) / (2 w^2))
      const double xp = x - p[2], yp = y - p[3];
      return p[0] * std::exp(-(xp * xp + yp * yp) / (2 * p[1] * p[1]));
    }
    case kPotQuartic: {  // rotated by pi/4: x^2 y^2 / 2 + b (x^4 + y^4) / 4
      const double xr = (x - y) / std::sqrt(2.0), yr = (x + y) / std::sqrt(2.0);
      const double x2 = xr * xr, y2 = yr * yr;
      return (x2 * y2) / 2 + p[0] * (x2 * x2 + y2 * y2) / 4;
    }
    case kPotSquareOscillator:  // (|x|^a + |y|^a) / 2
      return 0.5 * (std::pow(std::fabs(x), p[0]) + std::pow(std::fabs(y), p[0]));
    case kPotPower:  // A r^a / 2
      return 0.5 * p[1] * std::pow(std::hypot(x, y), p[0]);
    case kPotRing: {  // |r - R|/w to the e, halved, plus an asymmetry spike at (R, 0)
      const double rp = std::hypot(x, y), z = std::fabs(rp - p[0]) / p[1];
      const double xp = x - p[0];
      const double G = p[3] * std::exp(-(xp * xp + y * y) / (2 * p[4] * p[4]));
      return 0.5 * std::pow(z, p[2]) + G;
    }
    case kPotCosh:  // A (cosh(r/L) - 1)
      return p[0] * (std::cosh(std::hypot(x, y) / p[1]) - 1);
    case kPotSoftStadium: {  // V / (1 + a exp(b (1 - q^2))), q the scaled distance to the spine
      const double R = p[0], halfL = p[1];
      double q;
      if (std::fabs(x) <= halfL) q = y / R;
      else if (x < -halfL) q = std::hypot(x + halfL, y) / R;
      else q = std::hypot(x - halfL, y) / R;
      return p[2] / (1 + p[3] * std::exp(p[4] * (1 - q * q)));
    }
    case kPotPowerStadium: {  // q^a / 2
      const double R = p[0], halfL = p[1];
      double q;
      if (std::fabs(x) <= halfL) q = std::fabs(y / R);
      else if (x < -halfL) q = std::hypot(x + halfL, y) / R;
      else q = std::hypot(x - halfL, y) / R;
      return 0.5 * std::pow(q, p[2]);
    }
    default:
      return 0.0;
  }
}

// Contribution of one impurity at (sx, sy) to the point (px, py)
// (ImpurityType::add_noise, noise.cpp:25-78): Gaussian d = {A, w},
// Coulomb g = {e/2, alpha}, hemisphere g = {A, r^2}. Delta impurities touch
// one grid point and are added on the host.
ITP_HD double imp_value(int kind, const double* d, const double* g, double sx, double sy, double px, double py) {
  const double rx = sx - px, ry = sy - py, r2 = rx * rx + ry * ry;
  switch (kind) {
    case kImpGaussian:
      return d[0] * std::exp(-0.5 * (r2 / (d[1] * d[1])));
    case kImpCoulomb:
      return g[1] * std::pow(r2, -g[0]);
    case kImpHemisphere:
      return r2 < g[1] ? g[0] * std::sqrt(1 - r2 / g[1]) : 0.0;
    default:
      return 0.0;
  }
}

// Cell-centred position of index i on an axis of s points (datalayout.cpp:29-35).
ITP_HD double grid_pos(long long i, long long s, double dx) { return (double)(2 * i + 1 - s) * 0.5 * dx; }

// Nearest grid index of coordinate v (DataLayout::nearest_index,
// datalayout.hpp:79-88, with round_to_int of itp2d_common.hpp:69-71);
// -1 when outside the grid.
inline long long grid_nearest(double v, long long s, double dx) {
  const double r = v / dx + ((double)s - 1) / 2;
  const long long i = (long long)(r > 0 ? std::floor(r + 0.5) : std::ceil(r - 0.5));
  return (i < 0 || i >= s) ? -1 : i;
}

// Parameters of a named potential (parse_potential_description and the
// constructors, potentialtypes.cpp:56-500): the accepted counts, the defaults
// and the validity checks. Returns nullptr and fills q, or an error message.
inline const char* pot_resolve(const char* name, const double* a, size_t n, PotParams& q) {
  auto is = [&](const char* s) { return std::strcmp(name, s) == 0; };
  for (double& v : q.p) v = 0;
  if (is("zero")) {
    q.kind = kPotZero;
    return n == 0 ? nullptr : "zero potential does not take parameters";
  }
  if (is("harmonic")) {
    q.kind = kPotHarmonic;
    if (n == 0) { q.p[0] = 1; q.p[1] = 0; q.p[2] = 0; }
    else if (n == 1) { q.p[0] = a[0]; q.p[1] = 0; q.p[2] = 0; }
    else if (n == 3) { q.p[0] = a[0]; q.p[1] = a[1]; q.p[2] = a[2]; }
    else return "harmonic oscillator potential takes either zero, one or three parameters";
    return q.p[0] < 0 ? "harmonic oscillator with negative prefactor" : nullptr;
  }
  if (is("elliptic")) {
    q.kind = kPotElliptic;
    if (n == 0) { q.p[0] = 1; q.p[1] = 1.6180339887498948482; }
    else if (n == 2) { q.p[0] = a[0]; q.p[1] = a[1]; }
    else return "elliptic oscillator potential takes either zero or two parameters";
    return (q.p[0] < 0 || q.p[1] < 0) ? "elliptic oscillator with negative prefactor" : nullptr;
  }
  if (is("prettyhardsquare")) {
    q.kind = kPotPrettyHardSquare;
    if (n == 0) q.p[0] = 8;
    else if (n == 1) q.p[0] = a[0];
    else return "pretty hard square potential only takes one parameter";
    return nullptr;
  }
  if (is("softpentagon")) {
    q.kind = kPotSoftPentagon;
    return n == 0 ? nullptr : "soft pentagon potential does not take parameters";
  }
  if (is("henonheiles") || is("henon")) {
    q.kind = kPotHenonHeiles;
    if (n == 0) { q.p[0] = 205.0 / 42.0; q.p[1] = -13.0 / 3.0; }
    else if (n == 2) { q.p[0] = a[0]; q.p[1] = a[1]; }
    else return "Henon Heiles potential takes either two parameters or none";
    return nullptr;
  }
  if (is("gaussian") || is("gaussianblob")) {
    q.kind = kPotGaussian;
    if (n == 0) { q.p[0] = 1; q.p[1] = 1; q.p[2] = 0; q.p[3] = 0; }
    else if (n == 2) { q.p[0] = a[0]; q.p[1] = a[1]; q.p[2] = 0; q.p[3] = 0; }
    else if (n == 4) { q.p[0] = a[0]; q.p[1] = a[1]; q.p[2] = a[2]; q.p[3] = a[3]; }
    else return "gaussian potential takes 0, 2 or 4 parameters";
    if (q.p[0] < 0) return "gaussian potential with negative amplitude";
    return q.p[1] == 0 ? "gaussian potential with zero width" : nullptr;
  }
  if (is("quartic") || is("quarticoscillator")) {
    q.kind = kPotQuartic;
    if (n == 0) q.p[0] = 0.01;
    else if (n == 1) q.p[0] = a[0];
    else return "quartic potential only takes one parameter";
    return q.p[0] < 0 ? "quartic potential with negative parameter" : nullptr;
  }
  if (is("squareoscillator")) {
    q.kind = kPotSquareOscillator;
    if (n == 0) q.p[0] = 8;
    else if (n == 1) q.p[0] = a[0];
    else return "square oscillator potential only takes one parameter";
    return nullptr;
  }
  if (is("power") || is("poweroscillator")) {
    q.kind = kPotPower;
    if (n == 0) { q.p[0] = 4; q.p[1] = 1; }
    else if (n == 1) { q.p[0] = a[0]; q.p[1] = 1; }
    else if (n == 2) { q.p[0] = a[0]; q.p[1] = a[1]; }
    else return "power oscillator potential takes either zero, one or two parameters";
    return q.p[1] < 0 ? "power oscillator with negative prefactor" : nullptr;
  }
  if (is("ring") || is("ringpotential")) {
    q.kind = kPotRing;
    if (n == 0) { q.p[0] = 3; q.p[1] = 1; q.p[2] = 2; q.p[3] = 0; q.p[4] = 1; }
    else if (n == 3) { q.p[0] = a[0]; q.p[1] = a[1]; q.p[2] = a[2]; q.p[3] = 0; q.p[4] = 1; }
    else if (n == 5) { for (int i = 0; i < 5; i++) q.p[i] = a[i]; }
    else return "ring potential takes either zero, three, or five parameters";
    if (q.p[0] < 0) return "ring oscillator with negative radius";
    return q.p[1] <= 0 ? "ring oscillator with non-positive width" : nullptr;
  }
  if (is("cosh") || is("coshpotential")) {
    q.kind = kPotCosh;
    if (n == 0) { q.p[0] = 1; q.p[1] = 1; }
    else if (n == 2) { q.p[0] = a[0]; q.p[1] = a[1]; }
    else return "cosh potential takes either zero or two parameters";
    if (q.p[0] <= 0) return "cosh potential with non-positive amplitude";
    return q.p[1] <= 0 ? "cosh potential with non-positive length scale" : nullptr;
  }
  if (is("softstadium")) {
    q.kind = kPotSoftStadium;
    if (n == 0) { q.p[0] = 1; q.p[1] = 1; q.p[2] = 100; q.p[3] = 1; q.p[4] = 10; }
    else if (n == 3) { q.p[0] = a[0]; q.p[1] = a[1] / 2; q.p[2] = a[2]; q.p[3] = 1; q.p[4] = 10; }
    else if (n == 5) { q.p[0] = a[0]; q.p[1] = a[1] / 2; q.p[2] = a[2]; q.p[3] = a[3]; q.p[4] = a[4]; }
    else return "soft stadium potential takes either zero, three or five parameters";
    if (q.p[0] <= 0) return "soft stadium with non-positive radius";
    if (q.p[1] <= 0) return "soft stadium with non-positive center length";
    if (q.p[2] <= 0) return "soft stadium with non-positive height";
    if (q.p[3] <= 0) return "soft stadium with non-positive a parameter";
    return q.p[4] <= 0 ? "soft stadium with non-positive b parameter" : nullptr;
  }
  if (is("powerstadium")) {
    q.kind = kPotPowerStadium;
    if (n == 0) { q.p[0] = 1; q.p[1] = 1; q.p[2] = 4; }
    else if (n == 1) { q.p[0] = 1; q.p[1] = 1; q.p[2] = a[0]; }
    else if (n == 3) { q.p[0] = a[0]; q.p[1] = a[1] / 2; q.p[2] = a[2]; }
    else return "power stadium potential takes either zero, one or three parameters";
    if (q.p[0] <= 0) return "power stadium with non-positive radius";
    return q.p[1] <= 0 ? "power stadium with non-positive center length" : nullptr;
  }
  q.kind = -1;
  return "unknown potential type";
}

}  // namespace itp
