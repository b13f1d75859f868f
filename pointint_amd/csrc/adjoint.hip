// Adjoint of the volume matrix: vertex gradients of L = sum_ij G_ij V_ij
// without materialising dA/dvertex.
//
//   dL/dv = (1/8 pi^3) sum_q w_q Re( dA(k_q)/dv * conj(B(k_q)) ),
//   B_i   = sum_j (G_ij + G_ji) A_j,
//
// the vector-Jacobian product of the reference's per-vertex gradient
// (cpp/src/dof/triangle_mesh_dof.cpp:76-143, contracted as in
// cpp/src/compute_intersection_volume_multi_object.cpp:186-233).
//
// Per face, with u_j = k.v_j and A_face = kappa * 2 g[u0, u1, u2]
// (kappa = khat.S, g(z) = -e^{iz}, second divided difference), the
// derivatives are
//
//   dA_face/dv_j = kappa * G_j * k + 2 g[u0,u1,u2] * dkappa/dv_j,
//   G_j = d(2 g[u0,u1,u2])/du_j = 2 g[u0, u1, u2, u_j]  (node j repeated),
//   dkappa/dv0 = khat x (e2 - e1)/2, dkappa/dv1 = -khat x e2 / 2,
//   dkappa/dv2 = khat x e1 / 2.
//
// For spans of at least kSpanTaylor, G_j comes from differentiating the
// forward kernel's middle-node formula (tri_main_adj); below, from the Taylor
// series 2 g[0, a, b, d] = sum_n 2 i^(n+1) h_n(a, b, d) / (n+3)! (times
// e^{i u0}) in FP64.
//
// This kernel covers the nodes from each object's qbeg (the high-|k| shells);
// adjoint_series.hip covers the shells below.
#include "device_common.hpp"
#include "pii_internal.hpp"

namespace pii {

namespace {

constexpr int kLanes = 64;
constexpr int kWaves = 4;
constexpr double kSpanTaylor = 1.0;
constexpr float kSincPoly = 0.5f;

__device__ __forceinline__ float2 phasor_rev_a(double r) {
  const double j = rint(4.0 * r);
  const float x = (float)((r - 0.25 * j) * 6.283185307179586);
  const float x2 = x * x;
  float s = x2 * (-1.0f / 39916800.0f) + (1.0f / 362880.0f);
  s = s * x2 - (1.0f / 5040.0f);
  s = s * x2 + (1.0f / 120.0f);
  s = s * x2 - (1.0f / 6.0f);
  s = (s * x2) * x + x;
  float c = x2 * (-1.0f / 3628800.0f) + (1.0f / 40320.0f);
  c = c * x2 - (1.0f / 720.0f);
  c = c * x2 + (1.0f / 24.0f);
  c = c * x2 - 0.5f;
  c = c * x2 + 1.0f;
  const int q = ((int)j) & 3;
  float2 o;
  o.x = (q == 0) ? c : (q == 1) ? -s : (q == 2) ? -c : s;
  o.y = (q == 0) ? s : (q == 1) ? c : (q == 2) ? -s : -c;
  return o;
}

__device__ __forceinline__ float2 half_phasor_a(double u) {
  const double t = u * 0.07957747154594767;  // 1/(4 pi)
  return phasor_rev_a(t - rint(t));
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Phi(a, b) = sum_n 2 i^n h_n(a, b) / (n+2)!  (FP64), returned as (re, im).
__device__ __forceinline__ double2 phi_taylor(double a, double b) {
  double an = 1.0, hn = 1.0, re = 1.0, im = 0.0;
  an *= a; hn = fma(b, hn, an); im += hn * (1.0 / 3.0);
  an *= a; hn = fma(b, hn, an); re -= hn * (1.0 / 12.0);
  an *= a; hn = fma(b, hn, an); im -= hn * (1.0 / 60.0);
  an *= a; hn = fma(b, hn, an); re += hn * (1.0 / 360.0);
  an *= a; hn = fma(b, hn, an); im += hn * (1.0 / 2520.0);
  an *= a; hn = fma(b, hn, an); re -= hn * (1.0 / 20160.0);
  an *= a; hn = fma(b, hn, an); im -= hn * (1.0 / 181440.0);
  an *= a; hn = fma(b, hn, an); re += hn * (1.0 / 1814400.0);
  an *= a; hn = fma(b, hn, an); im += hn * (1.0 / 19958400.0);
  an *= a; hn = fma(b, hn, an); re -= hn * (1.0 / 239500800.0);
  an *= a; hn = fma(b, hn, an); im -= hn * (1.0 / 3113510400.0);
  return make_double2(re, im);
}

// 2 g[0, a, b, d] by its Taylor series (relative to u0), for d in {0, a, b}:
// returns the three values for d = 0, a, b.
__device__ __forceinline__ void dd3_taylor(double a, double b, double2& T0, double2& T1,
                                           double2& T2) {
  // h2_n = h_n(a, b); h3_n(d) = d h3_{n-1}(d) + h2_n.  Coefficients
  // c_n = 2 i^(n+1) / (n+3)!  : i/3, -1/12, -i/60, 1/360, ...
  double an = 1.0, h2 = 1.0, hA = 1.0, hB = 1.0;
  double r0 = 0.0, i0 = 1.0 / 3.0, r1 = 0.0, i1 = 1.0 / 3.0, r2 = 0.0, i2 = 1.0 / 3.0;
  const double c[11] = {-1.0 / 12.0, -1.0 / 60.0, 1.0 / 360.0, 1.0 / 2520.0, -1.0 / 20160.0,
                        -1.0 / 181440.0, 1.0 / 1814400.0, 1.0 / 19958400.0,
                        -1.0 / 239500800.0, -1.0 / 3113510400.0, 1.0 / 43589145600.0};
#pragma unroll
  for (int n = 1; n <= 11; ++n) {
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

// Sum over the 64 lanes of a wave with DPP moves (VALU only, no LDS):
// quad_perm [1,0,3,2] and [2,3,0,1] give quad sums, row_half_mirror and
// row_mirror row sums, row_bcast:15 (rows 1, 3) and row_bcast:31 (rows 2, 3)
// carry them into lane 63, which holds the total.  Lanes a move does not
// write receive 0.  Every lane of the wave must be active.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, kRowMask, 0xf, false));
}
__device__ __forceinline__ float wave_total(float v) {
  v += dpp_move<0xb1, 0xf>(v);
  v += dpp_move<0x4e, 0xf>(v);
  v += dpp_move<0x141, 0xf>(v);
  v += dpp_move<0x140, 0xf>(v);
  v += dpp_move<0x142, 0xa>(v);
  v += dpp_move<0x143, 0xc>(v);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
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

// One workgroup per (patch, node range): one lane per node in blocks of 64
// nodes; per block the waves fill an LDS table of the patch vertices' half
// phasors, then stream the patch's faces (a quarter each, wave-uniform, scalar
// loads).  Per face and node the lane forms s = Re(F w conj B) and
// r_j = kappa Re(G_j w conj B); the wave sums r_j k and s khat over its 64
// nodes and adds them to the face's LDS accumulator.  At the end the
// accumulators become vertex gradients (cross products with the edges for
// the dkappa/dv terms) added to the world-vertex gradient.
__global__ void __launch_bounds__(256) k_adjoint_tri(const ObjDesc* __restrict__ objs,
                                                     const double* __restrict__ krep, int Qh,
                                                     const float2* __restrict__ Bw, int ldB,
                                                     int kchunk, double* __restrict__ gworld,
                                                     const int* __restrict__ goff) {
  __shared__ float2 tab[kPatchVerts][kLanes];
  __shared__ double acc[kPatchFaces][12];
  const ObjDesc& d = objs[blockIdx.z];
  const int p = blockIdx.x;
  if (p >= d.npatch) return;
  // nodes below qbeg come from the moment series (adjoint_series.hip)
  const int kbeg = d.qbeg + blockIdx.y * kchunk;
  const int kend = min(Qh, kbeg + kchunk);
  if (kbeg >= kend) return;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const auto* hdr = as_const(d.patch_hdr);
  const int vbeg = hdr[4 * p], vcnt = hdr[4 * p + 1];
  const int fbeg = hdr[4 * p + 2], fcnt = hdr[4 * p + 3];
  for (int i = threadIdx.x; i < fcnt * 12; i += blockDim.x) (&acc[0][0])[i] = 0.0;
  const auto* pverts = as_const(d.patch_verts);
  const auto* world = as_const(d.world);
  const auto* rec = as_const(d.rec);
  const float2* Bcol = Bw + (size_t)d.col * ldB;

  for (int q0 = kbeg; q0 < kend; q0 += kLanes) {
    const int q = q0 + lane;
    const int qc = q < kend ? q : kend - 1;
    const double kx = krep[qc], ky = krep[(size_t)Qh + qc], kz = krep[2 * (size_t)Qh + qc];
    const float hx = (float)krep[3 * (size_t)Qh + qc], hy = (float)krep[4 * (size_t)Qh + qc],
                hz = (float)krep[5 * (size_t)Qh + qc];
    const double km = krep[6 * (size_t)Qh + qc];
    // w conj(B) = sqrt(w) conj(sqrt(w) B); zero for lanes past the range and
    // for k = 0 (A(0) = 0)
    const float sw = (q < kend && km >= 1e-10) ? (float)sqrt(krep[7 * (size_t)Qh + qc]) : 0.0f;
    const float2 bq = Bcol[qc];
    const float ox = bq.x * sw, oy = -bq.y * sw;
    __syncthreads();  // the previous block's table reads are done
    for (int sv = wave; sv < vcnt; sv += kWaves) {
      const int gv = pverts[vbeg + sv];
      const double u = kx * world[3 * gv] + ky * world[3 * gv + 1] + kz * world[3 * gv + 2];
      tab[sv][lane] = half_phasor_a(u);
    }
    __syncthreads();
    for (int f = wave; f < fcnt; f += kWaves) {
      const auto* r = rec + (size_t)kRecFastTri * (fbeg + f);
      const double al = kx * r[0] + ky * r[1] + kz * r[2];
      const double be = kx * r[3] + ky * r[4] + kz * r[5];
      const auto* rf = reinterpret_cast<const PII_CONST_AS float*>(r + 6);
      const float kap = 0.5f * (hx * rf[0] + hy * rf[1] + hz * rf[2]);
      const uint32_t loc = reinterpret_cast<const PII_CONST_AS uint32_t*>(rf)[3];
      const float2 Q0 = tab[loc & 0xff][lane], Q1 = tab[(loc >> 8) & 0xff][lane],
                   Q2 = tab[(loc >> 16) & 0xff][lane];
      const double H = fmax(0.0, fmax(al, be)) - fmin(0.0, fmin(al, be));
      float2 F, G0, G1, G2;
      if (H < kSpanTaylor) {
        double2 T0, T1, T2;
        dd3_taylor(al, be, T0, T1, T2);
        const double2 tf = phi_taylor(al, be);
        const float2 P0 = cmulf(Q0, Q0);  // e^{i u0}
        F = cmulf(P0, make_float2((float)tf.x, (float)tf.y));
        G0 = cmulf(P0, make_float2((float)T0.x, (float)T0.y));
        G1 = cmulf(P0, make_float2((float)T1.x, (float)T1.y));
        G2 = cmulf(P0, make_float2((float)T2.x, (float)T2.y));
      } else {
        tri_main_adj(al, be, Q0, Q1, Q2, F, G0, G1, G2);
      }
      const float sf = F.x * ox - F.y * oy;
      const float r0 = kap * (G0.x * ox - G0.y * oy);
      const float r1 = kap * (G1.x * ox - G1.y * oy);
      const float r2 = kap * (G2.x * ox - G2.y * oy);
      // per-block sums over the wave's 64 nodes in FP32 (relative rounding
      // 6e-8 of the block's terms), accumulated over blocks in FP64
      const float fx = (float)kx, fy = (float)ky, fz = (float)kz;
      const float v[12] = {r0 * fx, r0 * fy, r0 * fz, r1 * fx, r1 * fy, r1 * fz,
                           r2 * fx, r2 * fy, r2 * fz,
                           sf * hx, sf * hy, sf * hz};
      float t[12];
#pragma unroll
      for (int i = 0; i < 12; ++i) t[i] = wave_total(v[i]);
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 12; ++i) acc[f][i] += (double)t[i];
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < fcnt; t += blockDim.x) {
    const double* r = d.rec + (size_t)kRecFastTri * (fbeg + t);
    const double e1x = r[0], e1y = r[1], e1z = r[2], e2x = r[3], e2y = r[4], e2z = r[5];
    const double* R = acc[t];
    const double* Sk = acc[t] + 9;
    // dkappa terms: v0: Sk x (e2 - e1)/2, v1: -Sk x e2 / 2, v2: Sk x e1 / 2
    const double dx = e2x - e1x, dy = e2y - e1y, dz = e2z - e1z;
    double g[3][3];
    g[0][0] = R[0] + 0.5 * (Sk[1] * dz - Sk[2] * dy);
    g[0][1] = R[1] + 0.5 * (Sk[2] * dx - Sk[0] * dz);
    g[0][2] = R[2] + 0.5 * (Sk[0] * dy - Sk[1] * dx);
    g[1][0] = R[3] - 0.5 * (Sk[1] * e2z - Sk[2] * e2y);
    g[1][1] = R[4] - 0.5 * (Sk[2] * e2x - Sk[0] * e2z);
    g[1][2] = R[5] - 0.5 * (Sk[0] * e2y - Sk[1] * e2x);
    g[2][0] = R[6] + 0.5 * (Sk[1] * e1z - Sk[2] * e1y);
    g[2][1] = R[7] + 0.5 * (Sk[2] * e1x - Sk[0] * e1z);
    g[2][2] = R[8] + 0.5 * (Sk[0] * e1y - Sk[1] * e1x);
    const int f = (int)d.patch_faces[fbeg + t];
    double* G = gworld + goff[blockIdx.z];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int v = d.faces[3 * f + j];
#pragma unroll
      for (int c = 0; c < 3; ++c) atomicAdd(G + 3 * (size_t)v + c, g[j][c] * kInv8Pi3);
    }
  }
}

// Y = X C for complex X (float2, object-major, ldX) and real symmetric C (N x N,
// FP64): Y[i][q] = sum_j C[j][i] X[j][q].  One thread per (q, 8 outputs).
__global__ void __launch_bounds__(256) k_mix_columns(const float2* __restrict__ X, int N, int ldX,
                                                     int Qh, const double* __restrict__ C,
                                                     float2* __restrict__ Y) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0 = blockIdx.y * 8;
  if (q >= Qh) return;
  float accx[8] = {0, 0, 0, 0, 0, 0, 0, 0}, accy[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < N; ++j) {
    const float2 x = X[(size_t)j * ldX + q];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int i = i0 + m;
      const float c = (i < N) ? (float)C[(size_t)j * N + i] : 0.0f;
      accx[m] += c * x.x;
      accy[m] += c * x.y;
    }
  }
#pragma unroll
  for (int m = 0; m < 8; ++m)
    if (i0 + m < N) Y[(size_t)(i0 + m) * ldX + q] = make_float2(accx[m], accy[m]);
}

}  // namespace

void launch_mix_columns(hipStream_t s, const float2* d_X, int N, int ldX, int Qh,
                        const double* d_C, float2* d_Y) {
  if (N <= 0 || Qh <= 0) return;
  hipLaunchKernelGGL(k_mix_columns, dim3((Qh + 255) / 256, (N + 7) / 8), dim3(256), 0, s, d_X, N,
                     ldX, Qh, d_C, d_Y);
}

void launch_adjoint_tri(hipStream_t s, const ObjDesc* d_objs, int nobj, int max_patches,
                        const double* d_krep, int Qh, int qbeg_min, const float2* d_Bw, int ldB,
                        int ksplit, double* d_gworld, const int* d_goff) {
  const int span = Qh - qbeg_min;
  if (nobj <= 0 || max_patches <= 0 || span <= 0) return;
  int kchunk = (span + ksplit - 1) / ksplit;
  kchunk = (kchunk + kLanes - 1) / kLanes * kLanes;
  hipLaunchKernelGGL(k_adjoint_tri, dim3(max_patches, ksplit, nobj), dim3(kLanes * kWaves), 0, s,
                     d_objs, d_krep, Qh, d_Bw, ldB, kchunk, d_gworld, d_goff);
}

}  // namespace pii
