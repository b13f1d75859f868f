// Spherical-harmonic form of the moment series of a triangle mesh.
//
// The moment series (moments.hip) needs, per direction d, the sums
//   H_n(d) = sum_f (d.S_f) h_n(d.(v_f0 - c), d.(v_f1 - c), d.(v_f2 - c)),
// a homogeneous polynomial of degree n + 1 in d.  On the unit sphere it is a
// combination of spherical harmonics of degree l <= n + 1 with l = n + 1
// (mod 2).  For a rigid transform x' = R x + t of the mesh (and the series
// centre moved with it), H'_n(d) = H_n(R^T d): the expansion of the mesh in
// its own frame, computed once, gives H_n for every rotation by evaluating
// the harmonics at R^T d.
//
// Harmonics: with x = cos(theta) and the orthonormal associated Legendre
// functions Pt_l^m (Pt_0^0 = 1/sqrt(4 pi),
// Pt_m^m = -sqrt((2m+1)/(2m)) sin(theta) Pt_{m-1}^{m-1},
// Pt_l^m = a_lm (x Pt_{l-1}^m - b_lm Pt_{l-2}^m),
// a_lm = sqrt((4l^2 - 1)/(l^2 - m^2)), b_lm = sqrt(((l-1)^2 - m^2)/(4(l-1)^2 - 1))),
//   H_n(d) = sum_{l, m} Pt_l^m(x) (a_nlm cos(m phi) + b_nlm sin(m phi)),
// a_nlm = (2 - [m = 0]) int H_n Pt_l^m cos(m phi) dOmega, b_nlm = 2 int H_n
// Pt_l^m sin(m phi) dOmega, computed by a quadrature exact for the degree
// 2 (n + 1) integrands (Gauss-Legendre in x, uniform in phi).
//
// Coefficient layout per mesh (pairs a, b): for m = 0..kL, for l = m..kL,
// for n = n0(l), n0(l) + 2, .. <= kSeriesOrder with n0(l) = max(l - 1, 0) of
// the parity of l - 1 (n0(0) = 1).
#pragma once

#include "device_common.hpp"

namespace pii {
namespace sh {

constexpr int kP = kSeriesOrder;       // largest series order
constexpr int kL = kP + 1;             // largest harmonic degree
constexpr int kNx = kL + 1;            // Gauss-Legendre nodes in x = cos(theta)
constexpr int kNphi = 2 * kL + 1;      // uniform nodes in phi
constexpr int kNq = kNx * kNphi;       // quadrature directions

constexpr int n0(int l) { return l == 0 ? 1 : l - 1; }
constexpr int count(int l) { return n0(l) > kP ? 0 : (kP - n0(l)) / 2 + 1; }

// C[l]: pairs of degrees below l in one m block (from l = 0); mstart[m]:
// first pair of the block of m.  Pair (m, l, n) is at
// mstart[m] + C[l] - C[m] + (n - n0(l)) / 2.
struct Layout {
  int C[kL + 2];
  int mstart[kL + 2];
  int total;
  constexpr Layout() : C(), mstart(), total(0) {
    C[0] = 0;
    for (int l = 0; l <= kL; ++l) C[l + 1] = C[l] + count(l);
    mstart[0] = 0;
    for (int m = 0; m <= kL; ++m) mstart[m + 1] = mstart[m] + (C[kL + 1] - C[m]);
    total = mstart[kL + 1];
  }
};

constexpr double csqrt(double v) {
  if (v <= 0.0) return 0.0;
  double x = v > 1.0 ? v : 1.0;
  for (int i = 0; i < 200; ++i) {
    const double y = 0.5 * (x + v / x);
    if (y == x) break;
    x = y;
  }
  return x;
}

constexpr double kInvSqrt4Pi = 0.28209479177387814;   // 1 / sqrt(4 pi)

// Recurrence factors a_lm, b_lm (l <= kL, m < l) and the sectoral seeds
// seed[m] = Pt_m^m / sin^m(theta) = (-1)^m prod_{k=1..m} sqrt((2k+1)/(2k)) / sqrt(4 pi).
// The recurrence in l does not involve sin(theta), so Q_l^m = Pt_l^m / sin^m
// obeys it from Q_m^m = seed[m]: a polynomial in x, with no division at the
// poles, and Pt_l^m cos(m phi) = Q_l^m Re((p_x + i p_y)^m) for the unit
// vector p.
struct Recurrence {
  double a[kL + 1][kL + 1];
  double b[kL + 1][kL + 1];
  double seed[kL + 1];
  constexpr Recurrence() : a(), b(), seed() {
    for (int l = 0; l <= kL; ++l)
      for (int m = 0; m < l; ++m) {
        a[l][m] = csqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
        b[l][m] = csqrt(((l - 1.0) * (l - 1.0) - (double)m * m) /
                        (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
      }
    double s = kInvSqrt4Pi;
    seed[0] = s;
    for (int m = 1; m <= kL; ++m) {
      s *= -csqrt((2.0 * m + 1.0) / (2.0 * m));
      seed[m] = s;
    }
  }
};

}  // namespace sh
}  // namespace pii
