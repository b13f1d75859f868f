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
// Per face, with u_j = k.v_j, F = gamma... written as A_face = kappa * 2 g[u0, u1, u2]
// (kappa = khat.S, g(z) = -e^{iz}, second divided difference), the
// derivatives are
//
//   dA_face/dv_j = kappa * G_j * k + 2 g[u0,u1,u2] * dkappa/dv_j,
//   G_j = d(2 g[u0,u1,u2])/du_j = 2 g[u0, u1, u2, u_j]  (node j repeated),
//   dkappa/dv0 = khat x (e2 - e1)/2, dkappa/dv1 = -khat x e2 / 2,
//   dkappa/dv2 = khat x e1 / 2.
//
// With the nodes sorted a <= b <= c the third differences follow from five
// second differences, each evaluated by the same cancellation-free
// half-phasor formula as the forward kernel, or by Taylor series when the
// span is below kSpanTaylor:
//   G_a = (2g[a,b,c] - 2g[a,a,b]) / (c - a)
//   G_b = (2g[b,b,c] - 2g[a,b,b]) / (c - a)
//   G_c = (2g[b,c,c] - 2g[a,b,c]) / (c - a)
// and for a total span below kSpanTaylor directly from
//   2 g[0, a, b, d] = sum_n 2 i^(n+1) h_n(a, b, d) / (n+3)!   (times e^{i u0}).
//
// Mapping: one lane owns one face of a patch and loops over k-nodes in
// batches of kKB; per batch the workgroup fills an LDS table of the patch
// vertices' half phasors for those nodes.  Per face the lane accumulates
// sum_k r_j k (j = 0..2) and sum_k s khat in FP64; the cross products with
// the edges and the scatter to vertices happen once at the end.
#include "device_common.hpp"
#include "pii_internal.hpp"

namespace pii {

namespace {

constexpr int kKB = 32;              // k-nodes per LDS batch
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

__device__ __forceinline__ float sinc_a(float h, float sin_h) {
  if (fabsf(h) < kSincPoly) {
    const float h2 = h * h;
    float p = h2 * (1.0f / 362880.0f) - (1.0f / 5040.0f);
    p = p * h2 + (1.0f / 120.0f);
    p = p * h2 - (1.0f / 6.0f);
    return p * h2 + 1.0f;
  }
  return sin_h * __builtin_amdgcn_rcpf(h);
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

// 2 g[x0, x1, x2] for nodes given relative to a reference node r (offsets
// d0 <= d1 <= d2, FP64) with half phasors Q0, Q1, Q2 of the absolute nodes and
// the full phasor Pr of the reference node (for the Taylor branch, which
// expands around the smallest node).
__device__ __forceinline__ float2 dd2(double d0, double d1, double d2, float2 Q0, float2 Q1,
                                      float2 Q2) {
  const double span = d2 - d0;
  if (span < kSpanTaylor) {
    const double2 t = phi_taylor(d1 - d0, d2 - d0);
    const float2 P0 = cmulf(Q0, Q0);  // e^{i x0}
    return cmulf(P0, make_float2((float)t.x, (float)t.y));
  }
  // middle node x1:  2 g = -2i Q1 (Q2 sinc(hq) - Q0 sinc(hp)) / (x2 - x0)
  const double hp = 0.5 * (d1 - d0), hq = 0.5 * (d2 - d1);
  const float sp = sinc_a((float)hp, Q1.y * Q0.x - Q1.x * Q0.y);
  const float sq = sinc_a((float)hq, Q2.y * Q1.x - Q2.x * Q1.y);
  const float2 n = make_float2(Q2.x * sq - Q0.x * sp, Q2.y * sq - Q0.y * sp);
  const float2 t = cmulf(Q1, n);
  const float f = 2.0f * __builtin_amdgcn_rcpf((float)span);
  return make_float2(t.y * f, -t.x * f);
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

__global__ void __launch_bounds__(128) k_adjoint_tri(const ObjDesc* __restrict__ objs,
                                                     const double* __restrict__ krep, int Qh,
                                                     const float2* __restrict__ Bw, int ldB,
                                                     int kchunk, double* __restrict__ gworld,
                                                     const int* __restrict__ goff) {
  __shared__ float2 tab[kKB][kPatchVerts];
  const ObjDesc& d = objs[blockIdx.z];
  const int p = blockIdx.x;
  if (p >= d.npatch) return;
  const int vbeg = d.patch_hdr[4 * p], vcnt = d.patch_hdr[4 * p + 1];
  const int fbeg = d.patch_hdr[4 * p + 2], fcnt = d.patch_hdr[4 * p + 3];
  // nodes below qbeg come from the moment series (adjoint_series.hip)
  const int kbeg = d.qbeg + blockIdx.y * kchunk;
  const int kend = min(Qh, kbeg + kchunk);
  if (kbeg >= kend) return;
  const int t = threadIdx.x;
  const bool active = t < fcnt;
  const int fi = active ? t : 0;
  const double* r = d.rec + (size_t)kRecFastTri * (fbeg + fi);
  const double e1x = r[0], e1y = r[1], e1z = r[2], e2x = r[3], e2y = r[4], e2z = r[5];
  const float* rf = reinterpret_cast<const float*>(r + 6);
  const float S2x = rf[0], S2y = rf[1], S2z = rf[2];
  const uint32_t loc = reinterpret_cast<const uint32_t*>(rf)[3];
  const int l0 = loc & 0xff, l1 = (loc >> 8) & 0xff, l2 = (loc >> 16) & 0xff;
  const float2* Bcol = Bw + (size_t)d.col * ldB;

  double R[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double Sk[3] = {0, 0, 0};

  for (int k0 = kbeg; k0 < kend; k0 += kKB) {
    const int nk = min(kKB, kend - k0);
    __syncthreads();
    for (int e = t; e < nk * vcnt; e += blockDim.x) {
      const int kk = e / vcnt, s = e - kk * vcnt;
      const int q = k0 + kk;
      const int gv = d.patch_verts[vbeg + s];
      const double* W = d.world + 3 * (size_t)gv;
      const double u = krep[q] * W[0] + krep[(size_t)Qh + q] * W[1] + krep[2 * (size_t)Qh + q] * W[2];
      tab[kk][s] = half_phasor_a(u);
    }
    __syncthreads();
    if (!active) continue;
    for (int kk = 0; kk < nk; ++kk) {
      const int q = k0 + kk;
      const double kx = krep[q], ky = krep[(size_t)Qh + q], kz = krep[2 * (size_t)Qh + q];
      const float hx = (float)krep[3 * (size_t)Qh + q], hy = (float)krep[4 * (size_t)Qh + q],
                  hz = (float)krep[5 * (size_t)Qh + q];
      const double km = krep[6 * (size_t)Qh + q];
      if (km < 1e-10) continue;
      // conj(B) scaled by sqrt(w): Bw holds sqrt(w) B; times sqrt(w) again.
      const float sw = (float)sqrt(krep[7 * (size_t)Qh + q]);
      const float2 b = Bcol[q];
      const float bx = b.x * sw, by = -b.y * sw;   // sqrt(w) * w^{1/2} conj(B) = w conj(B)
      const double al = kx * e1x + ky * e1y + kz * e1z;
      const double be = kx * e2x + ky * e2y + kz * e2z;
      const float kap = 0.5f * (hx * S2x + hy * S2y + hz * S2z);
      const float2 Q0 = tab[kk][l0], Q1 = tab[kk][l1], Q2 = tab[kk][l2];
      // sort nodes {0 (v0), al (v1), be (v2)}
      double x[3] = {0.0, al, be};
      float2 Qs[3] = {Q0, Q1, Q2};
      int id[3] = {0, 1, 2};
#define PII_SWAP(i, j)                                               \
  if (x[j] < x[i]) {                                                 \
    double tx = x[i]; x[i] = x[j]; x[j] = tx;                        \
    float2 tq = Qs[i]; Qs[i] = Qs[j]; Qs[j] = tq;                    \
    int ti = id[i]; id[i] = id[j]; id[j] = ti;                       \
  }
      PII_SWAP(0, 1)
      PII_SWAP(1, 2)
      PII_SWAP(0, 1)
#undef PII_SWAP
      const double H = x[2] - x[0];
      float2 F, Gs[3];  // Gs in sorted order (a, b, c)
      if (H < kSpanTaylor) {
        double2 T0, T1, T2;
        dd3_taylor(al, be, T0, T1, T2);
        const double2 tf = phi_taylor(al, be);
        const float2 P0 = cmulf(Q0, Q0);
        F = cmulf(P0, make_float2((float)tf.x, (float)tf.y));
        const float2 G0 = cmulf(P0, make_float2((float)T0.x, (float)T0.y));
        const float2 G1 = cmulf(P0, make_float2((float)T1.x, (float)T1.y));
        const float2 G2 = cmulf(P0, make_float2((float)T2.x, (float)T2.y));
        // back to sorted order
        for (int s = 0; s < 3; ++s) Gs[s] = (id[s] == 0) ? G0 : (id[s] == 1) ? G1 : G2;
      } else {
        F = dd2(x[0], x[1], x[2], Qs[0], Qs[1], Qs[2]);
        const float2 Faab = dd2(x[0], x[0], x[1], Qs[0], Qs[0], Qs[1]);
        const float2 Fabb = dd2(x[0], x[1], x[1], Qs[0], Qs[1], Qs[1]);
        const float2 Fbbc = dd2(x[1], x[1], x[2], Qs[1], Qs[1], Qs[2]);
        const float2 Fbcc = dd2(x[1], x[2], x[2], Qs[1], Qs[2], Qs[2]);
        const float ih = __builtin_amdgcn_rcpf((float)H);
        Gs[0] = make_float2((F.x - Faab.x) * ih, (F.y - Faab.y) * ih);
        Gs[1] = make_float2((Fbbc.x - Fabb.x) * ih, (Fbbc.y - Fabb.y) * ih);
        Gs[2] = make_float2((Fbcc.x - F.x) * ih, (Fbcc.y - F.y) * ih);
      }
      // s = Re(F w conj B); r_j = kappa Re(G_j w conj B)
      const float s = F.x * bx - F.y * by;
      Sk[0] += (double)(s * hx); Sk[1] += (double)(s * hy); Sk[2] += (double)(s * hz);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const float rj = kap * (Gs[m].x * bx - Gs[m].y * by);
        const int j = id[m];
        const double rd = (double)rj;
        R[j][0] = fma(rd, kx, R[j][0]);
        R[j][1] = fma(rd, ky, R[j][1]);
        R[j][2] = fma(rd, kz, R[j][2]);
      }
    }
  }
  if (!active) return;
  // dkappa terms: v0: Sk x (e2 - e1)/2, v1: -Sk x e2 / 2, v2: Sk x e1 / 2
  const double dx = e2x - e1x, dy = e2y - e1y, dz = e2z - e1z;
  double g[3][3];
  g[0][0] = R[0][0] + 0.5 * (Sk[1] * dz - Sk[2] * dy);
  g[0][1] = R[0][1] + 0.5 * (Sk[2] * dx - Sk[0] * dz);
  g[0][2] = R[0][2] + 0.5 * (Sk[0] * dy - Sk[1] * dx);
  g[1][0] = R[1][0] - 0.5 * (Sk[1] * e2z - Sk[2] * e2y);
  g[1][1] = R[1][1] - 0.5 * (Sk[2] * e2x - Sk[0] * e2z);
  g[1][2] = R[1][2] - 0.5 * (Sk[0] * e2y - Sk[1] * e2x);
  g[2][0] = R[2][0] + 0.5 * (Sk[1] * e1z - Sk[2] * e1y);
  g[2][1] = R[2][1] + 0.5 * (Sk[2] * e1x - Sk[0] * e1z);
  g[2][2] = R[2][2] + 0.5 * (Sk[0] * e1y - Sk[1] * e1x);
  const int f = (int)d.patch_faces[fbeg + t];
  double* G = gworld + goff[blockIdx.z];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int v = d.faces[3 * f + j];
#pragma unroll
    for (int c = 0; c < 3; ++c) atomicAdd(G + 3 * (size_t)v + c, g[j][c] * kInv8Pi3);
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
  kchunk = (kchunk + kKB - 1) / kKB * kKB;
  hipLaunchKernelGGL(k_adjoint_tri, dim3(max_patches, ksplit, nobj), dim3(128), 0, s, d_objs,
                     d_krep, Qh, d_Bw, ldB, kchunk, d_gworld, d_goff);
}

}  // namespace pii
