// Device helpers for triangle form-factor derivatives, shared by the
// adjoint (adjoint.hip) and the DoF-derivative kernels of the Gauss-Newton
// Hessian (dtheta.hip).  Each translation unit gets its own copies.
//
// For a face with relative phases (0, a, b) (a = k.e1, b = k.e2) and half
// phasors Q_j = e^{i u_j / 2} of the absolute phases:
//   F   = 2 g[u0, u1, u2] = e^{i u0} Phi(a, b),  g(z) = -e^{iz},
//   G_j = dF/du_j = 2 g[u0, u1, u2, u_j]  (node j repeated).
#pragma once

#include "device_common.hpp"

namespace pii {
namespace {

constexpr double kSpanTaylor = 1.0;
// Lanes with a span below kLaneTaylor take the Taylor sums; the main formula
// above it (its FP32 error over the span squared stays near 1e-6 of G there).
constexpr float kLaneTaylor = 0.3f;
constexpr float kSincPoly = 0.5f;

// e^{i 2 pi r} for r in revolutions (|r| <= 0.5) on the hardware
// sine/cosine (revolution input): max abs error 1.25e-7 on MI355X.
__device__ __forceinline__ float2 phasor_rev_a(double r) {
  const float f = (float)r;
  return make_float2(__builtin_amdgcn_cosf(f), __builtin_amdgcn_sinf(f));
}

__device__ __forceinline__ float2 half_phasor_a(double u) {
  const double t = u * 0.07957747154594767;  // 1/(4 pi)
  return phasor_rev_a(t - rint(t));
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Phi(a, b) = sum_{n <= NO} 2 i^n h_n(a, b) / (n+2)!  (FP64), returned as
// (re, im).  |h_n| <= (n+1) H^n for span H, so the first omitted term is
// below 2 (NO+2) H^(NO+1) / (NO+3)!.
struct PhiCoef {
  double c[12];  // 2/(n+2)! with the sign of the real factor of i^n
  constexpr PhiCoef() : c() {
    double f = 1.0;
    for (int n = 0; n < 12; ++n) {
      c[n] = (((n / 2) % 2 == 0) ? 1.0 : -1.0) * f;
      f /= (double)(n + 3);
    }
  }
};
__device__ constexpr PhiCoef kPhiCoef{};

template <int NO>
__device__ __forceinline__ double2 phi_taylor(double a, double b) {
  static_assert(NO >= 1 && NO <= 11, "order range");
  double an = 1.0, hn = 1.0, re = 1.0, im = 0.0;
#pragma unroll
  for (int n = 1; n <= NO; ++n) {
    an *= a;
    hn = fma(b, hn, an);
    if (n & 1) im = fma(hn, kPhiCoef.c[n], im);
    else re = fma(hn, kPhiCoef.c[n], re);
  }
  return make_double2(re, im);
}

// 2 g[0, a, b, d] by its Taylor series (relative to u0) to order NO, for d
// in {0, a, b}: returns the three values for d = 0, a, b.  The order-n term
// has |h3_n| <= (n+1)(n+2)/2 H^n and factor 2/(n+3)!.
template <int NO>
__device__ __forceinline__ void dd3_taylor(double a, double b, double2& T0, double2& T1,
                                           double2& T2) {
  static_assert(NO >= 1 && NO <= 11, "order range");
  // h2_n = h_n(a, b); h3_n(d) = d h3_{n-1}(d) + h2_n.  Coefficients
  // c_n = 2 i^(n+1) / (n+3)!  : i/3, -1/12, -i/60, 1/360, ...
  double an = 1.0, h2 = 1.0, hA = 1.0, hB = 1.0;
  double r0 = 0.0, i0 = 1.0 / 3.0, r1 = 0.0, i1 = 1.0 / 3.0, r2 = 0.0, i2 = 1.0 / 3.0;
  const double c[11] = {-1.0 / 12.0, -1.0 / 60.0, 1.0 / 360.0, 1.0 / 2520.0, -1.0 / 20160.0,
                        -1.0 / 181440.0, 1.0 / 1814400.0, 1.0 / 19958400.0,
                        -1.0 / 239500800.0, -1.0 / 3113510400.0, 1.0 / 43589145600.0};
#pragma unroll
  for (int n = 1; n <= NO; ++n) {
    an *= a;
    h2 = fma(b, h2, an);
    hA = fma(a, hA, h2);
    hB = fma(b, hB, h2);
    const double cf = c[n - 1];
    if (n & 1) {  // i^(n+1) real
      r0 += cf * h2; r1 += cf * hA; r2 += cf * hB;
    } else {
      i0 += cf * h2; i1 += cf * hA; i2 += cf * hB;
    }
  }
  T0 = make_double2(r0, i0);
  T1 = make_double2(r1, i1);
  T2 = make_double2(r2, i2);
}

// F, G_0..2 by the Taylor sums for one lane: NO = 7 keeps the first omitted
// terms below 1e-8 for spans up to kLaneTaylor, NO = 11 up to kSpanTaylor.
template <int NO>
__device__ __forceinline__ void tri_taylor_adj(double al, double be, float2 Q0, float2& F,
                                               float2& G0, float2& G1, float2& G2) {
  double2 T0, T1, T2;
  dd3_taylor<NO>(al, be, T0, T1, T2);
  const double2 tf = phi_taylor<NO>(al, be);
  const float2 P0 = make_float2(Q0.x * Q0.x - Q0.y * Q0.y, 2.0f * Q0.x * Q0.y);  // e^{i u0}
  const float2 fF = make_float2((float)tf.x, (float)tf.y);
  F = make_float2(P0.x * fF.x - P0.y * fF.y, P0.x * fF.y + P0.y * fF.x);
  const float2 f0 = make_float2((float)T0.x, (float)T0.y);
  const float2 f1 = make_float2((float)T1.x, (float)T1.y);
  const float2 f2 = make_float2((float)T2.x, (float)T2.y);
  G0 = make_float2(P0.x * f0.x - P0.y * f0.y, P0.x * f0.y + P0.y * f0.x);
  G1 = make_float2(P0.x * f1.x - P0.y * f1.y, P0.x * f1.y + P0.y * f1.x);
  G2 = make_float2(P0.x * f2.x - P0.y * f2.y, P0.x * f2.y + P0.y * f2.x);
}

// A DPP move (VALU only, no LDS): quad_perm [1,0,3,2] (0xb1) and [2,3,0,1]
// (0x4e) exchange within quads, row_half_mirror (0x141) and row_mirror
// (0x140) within rows of 16 lanes.  Lanes a move does not write receive 0.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, kRowMask, 0xf, false));
}

// Adds the wave totals of 12 per-lane values to acc[0..11] (FP64, LDS; one
// wave owns acc).  Two rounds of lane swaps fold the values pairwise:
// v_permlane32_swap exchanges lanes 32-63 of its first operand with lanes
// 0-31 of its second, so one add leaves value 2i's half sums in lanes 0-31
// and value 2i+1's in lanes 32-63; v_permlane16_swap exchanges the odd rows
// (16 lanes) of its first operand with the even rows of its second, so after
// one more add row r of register j holds 16 partial sums of value
// 4j + {0, 2, 1, 3}[r].  Four DPP steps (quad_perm [1,0,3,2], [2,3,0,1],
// row_half_mirror, row_mirror) give every lane its row's total, and lanes 0,
// 16, 32, 48 add the three registers' totals to their accumulators: 9 swaps,
// 21 adds and 3 LDS atomics in place of 12 separate wave reductions.  Every
// lane of the wave must be active.
__device__ __forceinline__ void wave_totals12_add(const float (&v)[12], double* acc, int lane) {
  float w[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[2 * i]),
                                                    __float_as_uint(v[2 * i + 1]), false, false);
    w[i] = __uint_as_float(s[0]) + __uint_as_float(s[1]);
  }
  float z[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const auto s = __builtin_amdgcn_permlane16_swap(__float_as_uint(w[2 * j]),
                                                    __float_as_uint(w[2 * j + 1]), false, false);
    z[j] = __uint_as_float(s[0]) + __uint_as_float(s[1]);
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) z[j] += dpp_move<0xb1, 0xf>(z[j]);
#pragma unroll
  for (int j = 0; j < 3; ++j) z[j] += dpp_move<0x4e, 0xf>(z[j]);
#pragma unroll
  for (int j = 0; j < 3; ++j) z[j] += dpp_move<0x141, 0xf>(z[j]);
#pragma unroll
  for (int j = 0; j < 3; ++j) z[j] += dpp_move<0x140, 0xf>(z[j]);
  if ((lane & 15) == 0) {
    const int r = lane >> 4;
    const int o = r == 1 ? 2 : (r == 2 ? 1 : r);
#pragma unroll
    for (int j = 0; j < 3; ++j) atomicAdd(acc + 4 * j + o, (double)z[j]);
  }
}

// sinc'(h) = (cos h - sinc h) / h from cos h and sinc h; odd polynomial below
// |h| = kSincPoly (truncation < 1e-9 there).  Both evaluated, then selected.
__device__ __forceinline__ float dsinc_from(float h, float cos_h, float sinc_h) {
  const float h2 = h * h;
  float p = h2 * (1.0f / 45360.0f) - (1.0f / 840.0f);
  p = p * h2 + (1.0f / 30.0f);
  p = p * h2 - (1.0f / 3.0f);
  p = p * h;
  const float r = (cos_h - sinc_h) * __builtin_amdgcn_rcpf(h);
  return fabsf(h) < kSincPoly ? p : r;
}

__device__ __forceinline__ float sinc_sel(float h, float sin_h) {
  const float h2 = h * h;
  float p = h2 * (1.0f / 362880.0f) - (1.0f / 5040.0f);
  p = p * h2 + (1.0f / 120.0f);
  p = p * h2 - (1.0f / 6.0f);
  p = p * h2 + 1.0f;
  const float r = sin_h * __builtin_amdgcn_rcpf(h);
  return fabsf(h) < kSincPoly ? p : r;
}

// F = 2 g[u0, u1, u2] and G_j = dF/du_j for relative phases (0, a, b) and
// half phasors Q0..Q2 of the absolute phases, span >= kSpanTaylor.  With the
// middle node m and the outer p, q:  F = -2i T / D,  T = Q_m N,
// N = Q_q S(h_q) - Q_p S(h_p),  D = u_q - u_p,  h the half gaps, S = sinc,
// dQ/du = (i/2) Q, and
//   G_q = -2i (Q_m dN/du_q / D - T / D^2),  dN/du_q =  Q_q (S'_q + i S_q) / 2,
//   G_p = -2i (Q_m dN/du_p / D + T / D^2),  dN/du_p =  Q_p (S'_p - i S_p) / 2,
//   G_m = i F - G_p - G_q   (translation: sum_j dF/du_j = i F).
__device__ __forceinline__ void tri_main_adj(double al, double be, float2 Q0, float2 Q1,
                                             float2 Q2, float2& F, float2& G0, float2& G1,
                                             float2& G2) {
  const float a = (float)al, b = (float)be;
  const bool m0 = (al * be <= 0.0);
  const bool m1 = (!m0) & (fabsf(a) < fabsf(b));
  const bool m01 = m0 | m1;
  const float um = m0 ? 0.0f : (m1 ? a : b);
  const float up = m0 ? a : 0.0f;
  const float uq = m01 ? b : a;
  const float2 Qm = make_float2(m0 ? Q0.x : (m1 ? Q1.x : Q2.x), m0 ? Q0.y : (m1 ? Q1.y : Q2.y));
  const float2 Qp = make_float2(m0 ? Q1.x : Q0.x, m0 ? Q1.y : Q0.y);
  const float2 Qq = make_float2(m01 ? Q2.x : Q1.x, m01 ? Q2.y : Q1.y);
  const float hp = 0.5f * (um - up), hq = 0.5f * (uq - um);
  const float iD = __builtin_amdgcn_rcpf(uq - up);
  // e^{i h_p} = Q_m conj Q_p, e^{i h_q} = Q_q conj Q_m
  const float sin_p = Qm.y * Qp.x - Qm.x * Qp.y, cos_p = Qm.x * Qp.x + Qm.y * Qp.y;
  const float sin_q = Qq.y * Qm.x - Qq.x * Qm.y, cos_q = Qq.x * Qm.x + Qq.y * Qm.y;
  const float Sp = sinc_sel(hp, sin_p), Sq = sinc_sel(hq, sin_q);
  const float dSp = dsinc_from(hp, cos_p, Sp), dSq = dsinc_from(hq, cos_q, Sq);
  const float2 N = make_float2(Qq.x * Sq - Qp.x * Sp, Qq.y * Sq - Qp.y * Sp);
  const float2 T = cmulf(Qm, N);
  F = make_float2(2.0f * T.y * iD, -2.0f * T.x * iD);
  const float2 dNq = cmulf(Qq, make_float2(0.5f * dSq, 0.5f * Sq));
  const float2 dNp = cmulf(Qp, make_float2(0.5f * dSp, -0.5f * Sp));
  const float2 Mq = cmulf(Qm, dNq), Mp = cmulf(Qm, dNp);
  const float iD2 = iD * iD;
  // -2i (x + i y) = 2y - 2i x
  const float2 cq = make_float2(Mq.x * iD - T.x * iD2, Mq.y * iD - T.y * iD2);
  const float2 cp = make_float2(Mp.x * iD + T.x * iD2, Mp.y * iD + T.y * iD2);
  const float2 Gq = make_float2(2.0f * cq.y, -2.0f * cq.x);
  const float2 Gp = make_float2(2.0f * cp.y, -2.0f * cp.x);
  const float2 Gm = make_float2(-F.y - Gp.x - Gq.x, F.x - Gp.y - Gq.y);
  G0 = make_float2(m0 ? Gm.x : Gp.x, m0 ? Gm.y : Gp.y);
  G1 = make_float2(m0 ? Gp.x : (m1 ? Gm.x : Gq.x), m0 ? Gp.y : (m1 ? Gm.y : Gq.y));
  G2 = make_float2(m01 ? Gq.x : Gm.x, m01 ? Gq.y : Gm.y);
}

}  // namespace
}  // namespace pii
