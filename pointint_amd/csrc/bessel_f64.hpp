// FP64 J1(x)/x with its derivative, shared by the affine derivative kernels
// (dtheta.hip, jacobian.hip).  Each translation unit gets its own copy.
#pragma once

#include "device_common.hpp"

namespace pii {
namespace {

// J1(x)/x and its derivative -J2(x)/x for x >= 0 in FP64: the power series
// in y = x^2/4 below 12 (25 terms), Hankel asymptotics of J0 and J1 above
// (6 terms each of P and Q, coefficients from a_k(nu) = prod_{l<=k} (4 nu^2 -
// (2l-1)^2) / (k! 8^k)).
struct HankelPQ {
  double p0[6], q0[6], p1[6], q1[6];
  constexpr HankelPQ() : p0(), q0(), p1(), q1() {
    for (int nu = 0; nu <= 1; ++nu) {
      double a[13] = {};
      a[0] = 1.0;
      for (int k = 1; k <= 12; ++k)
        a[k] = a[k - 1] * (4.0 * nu * nu - (2.0 * k - 1) * (2.0 * k - 1)) / (8.0 * k);
      for (int k = 0; k < 6; ++k) {
        const double sp = (k % 2 == 0) ? 1.0 : -1.0;
        (nu == 0 ? p0 : p1)[k] = sp * a[2 * k];
        (nu == 0 ? q0 : q1)[k] = sp * a[2 * k + 1];
      }
    }
  }
};
__device__ constexpr HankelPQ kHankel{};

struct J1OxSeries {
  double c[25];
  constexpr J1OxSeries() : c() {
    double t = 0.5;
    for (int m = 0; m < 25; ++m) {
      c[m] = t;
      t = -t / ((double)(m + 1) * (double)(m + 2));
    }
  }
};
__device__ constexpr J1OxSeries kJ1Ox{};

__device__ __forceinline__ void j1ox_and_prime(double x, double& f, double& fp) {
  if (x < 12.0) {
    const double y = 0.25 * x * x;
    double s = kJ1Ox.c[24], sp = 24.0 * kJ1Ox.c[24];
#pragma unroll
    for (int m = 23; m >= 0; --m) {
      s = fma(s, y, kJ1Ox.c[m]);
      if (m >= 1) sp = fma(sp, y, (double)m * kJ1Ox.c[m]);
    }
    f = s;
    fp = sp * 0.5 * x;  // d/dx sum c_m y^m = sum m c_m y^(m-1) x/2
    return;
  }
  const double inv = 1.0 / x, i2 = inv * inv;
  double P0 = 0.0, Q0 = 0.0, P1 = 0.0, Q1 = 0.0;
#pragma unroll
  for (int k = 5; k >= 0; --k) {
    P0 = fma(P0, i2, kHankel.p0[k]);
    Q0 = fma(Q0, i2, kHankel.q0[k]);
    P1 = fma(P1, i2, kHankel.p1[k]);
    Q1 = fma(Q1, i2, kHankel.q1[k]);
  }
  Q0 *= inv;
  Q1 *= inv;
  double s0, c0, s1, c1;
  sincos(x - 0.7853981633974483, &s0, &c0);   // x - pi/4
  sincos(x - 2.356194490192345, &s1, &c1);    // x - 3 pi/4
  const double amp = sqrt(0.6366197723675814 * inv);
  const double J0 = amp * (P0 * c0 - Q0 * s0);
  const double J1 = amp * (P1 * c1 - Q1 * s1);
  f = J1 * inv;
  fp = (J0 - 2.0 * J1 * inv) * inv;
}

}  // namespace
}  // namespace pii
