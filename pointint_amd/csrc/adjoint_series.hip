// Adjoint of the low-|k| moment series (the shells a triangle object takes
// from the series in the forward pass).
//
// Forward (moments.hip):  A(kappa d) = e^{i kappa d.c} sum_n (i kappa)^n c_n H_n(d),
// c_n = 2/(n+2)!,  H_n(d) = sum_f s_f h_n(w_f0, w_f1, w_f2),  s_f = d.S_f,
// w_fj = d.(v_fj - c).  The expansion centre c is a constant of the call (A
// does not depend on it), so for L = sum_ij G_ij V_ij
//
//   dL/dv = (1/8 pi^3) sum_q w_q Re(dA(k_q)/dv conj B(k_q))
//         = sum_d sum_n Gamma_n(d) dH_n(d)/dv,
//   Gamma_n(d) = c_n/(8 pi^3) sum_{s < nser} Re[(i kappa_s)^n e^{i kappa_s d.c} w conj B(kappa_s d)].
//
// Per face and direction, with D_j,m = h_m(w0, w1, w2, w_j) (node j repeated,
// dh_{m+1}/dw_j = D_j,m):
//
//   dH/dv_j = s g_j d + gs dS.d/dv_j,   gs = sum_n Gamma_n h_n,
//   g_j = sum_m Gamma_{m+1} D_j,m,   sum_j D_j,m = (m+3) h_m,
//   d(d.S)/dv0 = (e1 - e2) x d / 2,  d(d.S)/dv1 = e2 x d / 2,  d(d.S)/dv2 = d x e1 / 2,
//
// so g_2 follows from g_0, g_1 and sum_m (m+3) Gamma_{m+1} h_m.  Summed over
// directions, the cross products need only Sd = sum_d gs d.
//
// Kernels: k_series_gamma (one lane per direction, loop over the object's
// series shells) writes Gamma_n and (m+3) Gamma_{m+1} per (object, direction);
// k_series_adjoint_tri (one lane per face, loop over directions, the
// direction's Gamma read with scalar loads) accumulates the face's vertex
// gradients and adds them to the world-vertex gradient.
#include "device_common.hpp"
#include "pii_internal.hpp"
#include "series_common.hpp"

namespace pii {

namespace {

constexpr int P = kSeriesOrder;
constexpr int kLanes = 64;

struct GammaCoef {
  double c[P + 1];
  constexpr GammaCoef() : c() {
    double f = 1.0;  // 2/(n+2)!
    for (int n = 0; n <= P; ++n) {
      c[n] = f;
      f = f / (double)(n + 3);
    }
  }
};
__device__ constexpr GammaCoef kGammaCoef{};

__global__ void __launch_bounds__(kLanes) k_series_gamma(const ObjDesc* __restrict__ objs,
                                                         const double* __restrict__ krep, int Qh,
                                                         int Mh, const float2* __restrict__ Bw,
                                                         int ldB, double* __restrict__ gam) {
  const ObjDesc& d = objs[blockIdx.y];
  const int nser = d.nser;
  if (nser <= 0) return;
  const int j = blockIdx.x * kLanes + threadIdx.x;
  if (j >= Mh) return;
  const double dc = krep[3 * (size_t)Qh + j] * d.ctr[0] + krep[4 * (size_t)Qh + j] * d.ctr[1] +
                    krep[5 * (size_t)Qh + j] * d.ctr[2];
  const float2* Bcol = Bw + (size_t)d.col * ldB;
  double G[P + 1];
#pragma unroll
  for (int n = 0; n <= P; ++n) G[n] = 0.0;
  for (int s = 0; s < nser; ++s) {
    const int q = s * Mh + j;
    const double kap = krep[6 * (size_t)Qh + q];
    if (kap < 1e-10) continue;  // A(0) = 0 contributes nothing
    // w conj(B) = sqrt(w) * conj(sqrt(w) B)
    const double sw = sqrt(krep[7 * (size_t)Qh + q]);
    const float2 b = Bcol[q];
    const double ox = sw * (double)b.x, oy = -sw * (double)b.y;
    double sn, cs;
    sincos(kap * dc, &sn, &cs);
    double zx = cs * ox - sn * oy, zy = cs * oy + sn * ox;
#pragma unroll
    for (int n = 0; n <= P; ++n) {
      G[n] += zx;
      // z *= i kappa
      const double t = zx;
      zx = -kap * zy;
      zy = kap * t;
    }
  }
  double* out = gam + ((size_t)blockIdx.y * Mh + j) * kGammaStride;
#pragma unroll
  for (int n = 0; n <= P; ++n) out[n] = G[n] * kGammaCoef.c[n] * kInv8Pi3;
#pragma unroll
  for (int m = 0; m < P; ++m)
    out[P + 1 + m] = (double)(m + 3) * G[m + 1] * kGammaCoef.c[m + 1] * kInv8Pi3;
}

// Direction loop of one face at series order PP (the object's order: its
// forward series stops there too).  Returns the sums behind the face's
// vertex gradients: Gj = sum_d (S.d) (dH/dw_j-series) d and Sd = sum_d gs d.
template <int PP>
__device__ __forceinline__ void series_adjoint_dirs(const PII_CONST_AS double* __restrict__ kr,
                                                    int Qh, int Mh,
                                                    const PII_CONST_AS double* __restrict__ ga,
                                                    const double (&x)[9], const double (&S)[3],
                                                    double (&G0)[3], double (&G1)[3],
                                                    double (&G2)[3], double (&Sd)[3]) {
  for (int j = 0; j < Mh; ++j) {
    const double dx = kr[3 * (size_t)Qh + j], dy = kr[4 * (size_t)Qh + j],
                 dz = kr[5 * (size_t)Qh + j];
    const auto* g = ga + (size_t)j * kGammaStride;
    const double w0 = dx * x[0] + dy * x[1] + dz * x[2];
    const double w1 = dx * x[3] + dy * x[4] + dz * x[5];
    const double w2 = dx * x[6] + dy * x[7] + dz * x[8];
    const double s = dx * S[0] + dy * S[1] + dz * S[2];
    double b = 1.0, hab = 1.0, habc = 1.0, D0 = 1.0, D1 = 1.0;
    double gs = g[0], g0 = g[1], g1 = g[1], gt = g[P + 1];
#pragma unroll
    for (int n = 1; n <= PP; ++n) {
      b *= w1;
      hab = fma(w0, hab, b);
      habc = fma(w2, habc, hab);
      gs = fma(g[n], habc, gs);
      if (n < PP) {
        D0 = fma(w0, D0, habc);
        D1 = fma(w1, D1, habc);
        g0 = fma(g[n + 1], D0, g0);
        g1 = fma(g[n + 1], D1, g1);
        gt = fma(g[P + 1 + n], habc, gt);
      }
    }
    const double sg0 = s * g0, sg1 = s * g1, sg2 = s * (gt - g0 - g1);
    G0[0] = fma(sg0, dx, G0[0]); G0[1] = fma(sg0, dy, G0[1]); G0[2] = fma(sg0, dz, G0[2]);
    G1[0] = fma(sg1, dx, G1[0]); G1[1] = fma(sg1, dy, G1[1]); G1[2] = fma(sg1, dz, G1[2]);
    G2[0] = fma(sg2, dx, G2[0]); G2[1] = fma(sg2, dy, G2[1]); G2[2] = fma(sg2, dz, G2[2]);
    Sd[0] = fma(gs, dx, Sd[0]); Sd[1] = fma(gs, dy, Sd[1]); Sd[2] = fma(gs, dz, Sd[2]);
  }
}

__global__ void __launch_bounds__(256) k_series_adjoint_tri(
    const ObjDesc* __restrict__ objs, const double* __restrict__ krep, int Qh, int Mh,
    const double* __restrict__ gam, double* __restrict__ gworld, const int* __restrict__ goff) {
  const ObjDesc& d = objs[blockIdx.y];
  if (d.nser <= 0) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.n) return;
  const double* r = d.mrec + (size_t)kRecMoment * e;
  const double x[9] = {r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8]};
  const double S[3] = {r[9], r[10], r[11]};
  double G0[3] = {0, 0, 0}, G1[3] = {0, 0, 0}, G2[3] = {0, 0, 0}, Sd[3] = {0, 0, 0};
  const auto* kr = as_const(krep);
  const auto* ga = as_const(gam) + (size_t)blockIdx.y * Mh * kGammaStride;
#define PII_SADJ_BODY(PP) series_adjoint_dirs<PP>(kr, Qh, Mh, ga, x, S, G0, G1, G2, Sd)
  PII_SERIES_SWITCH(PII_SADJ_BODY)
#undef PII_SADJ_BODY
  const double x0 = x[0], y0 = x[1], z0 = x[2];
  const double x1 = x[3], y1 = x[4], z1 = x[5];
  const double x2 = x[6], y2 = x[7], z2 = x[8];
  // area-vector terms: v0: (e1 - e2) x Sd / 2, v1: e2 x Sd / 2, v2: Sd x e1 / 2
  const double e1x = x1 - x0, e1y = y1 - y0, e1z = z1 - z0;
  const double e2x = x2 - x0, e2y = y2 - y0, e2z = z2 - z0;
  const double ax = e1x - e2x, ay = e1y - e2y, az = e1z - e2z;
  double gv[3][3];
  gv[0][0] = G0[0] + 0.5 * (ay * Sd[2] - az * Sd[1]);
  gv[0][1] = G0[1] + 0.5 * (az * Sd[0] - ax * Sd[2]);
  gv[0][2] = G0[2] + 0.5 * (ax * Sd[1] - ay * Sd[0]);
  gv[1][0] = G1[0] + 0.5 * (e2y * Sd[2] - e2z * Sd[1]);
  gv[1][1] = G1[1] + 0.5 * (e2z * Sd[0] - e2x * Sd[2]);
  gv[1][2] = G1[2] + 0.5 * (e2x * Sd[1] - e2y * Sd[0]);
  gv[2][0] = G2[0] + 0.5 * (Sd[1] * e1z - Sd[2] * e1y);
  gv[2][1] = G2[1] + 0.5 * (Sd[2] * e1x - Sd[0] * e1z);
  gv[2][2] = G2[2] + 0.5 * (Sd[0] * e1y - Sd[1] * e1x);
  const int f = (int)d.patch_faces[e];
  double* out = gworld + goff[blockIdx.y];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int v = d.faces[3 * f + k];
#pragma unroll
    for (int c = 0; c < 3; ++c) atomicAdd(out + 3 * (size_t)v + c, gv[k][c]);
  }
}

}  // namespace

void launch_series_adjoint_tri(hipStream_t s, const ObjDesc* d_objs, int N, int max_faces,
                               const double* d_krep, int Qh, int Mh, const float2* d_Bw, int ldB,
                               double* d_gam, double* d_gworld, const int* d_goff) {
  if (N <= 0 || Mh <= 0 || max_faces <= 0) return;
  hipLaunchKernelGGL(k_series_gamma, dim3((Mh + kLanes - 1) / kLanes, N), dim3(kLanes), 0, s,
                     d_objs, d_krep, Qh, Mh, d_Bw, ldB, d_gam);
  hipLaunchKernelGGL(k_series_adjoint_tri, dim3((max_faces + 255) / 256, N), dim3(256), 0, s,
                     d_objs, d_krep, Qh, Mh, d_gam, d_gworld, d_goff);
}

}  // namespace pii
