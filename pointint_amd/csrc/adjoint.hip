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
// For spans of at least kLaneTaylor, G_j comes from differentiating the
// forward kernel's middle-node formula (tri_main_adj); below, from the Taylor
// series 2 g[0, a, b, d] = sum_n 2 i^(n+1) h_n(a, b, d) / (n+3)! (times
// e^{i u0}) in FP64, to order 7.
//
// This kernel covers the nodes from each object's qbeg (the high-|k| shells);
// adjoint_series.hip covers the shells below.
#include "device_common.hpp"
#include "pii_internal.hpp"
#include "tri_derivs.hpp"

namespace pii {

namespace {

constexpr int kLanes = 64;
constexpr int kWaves = 4;
// One workgroup per (patch, node range): one lane per node in blocks of 64
// nodes; per block the waves fill an LDS table of the patch vertices' half
// phasors, then stream the patch's faces (a quarter each, wave-uniform, scalar
// loads).  Per face and node the lane forms s = Re(F w conj B) and
// r_j = kappa Re(G_j w conj B); the wave sums r_j k and s khat over its 64
// nodes (lane swaps, then DPP) and adds them to the face's LDS accumulator.  At the end the
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
      // span max(|a|, |b|, |a - b|) in FP32
      const float fa = (float)al, fb = (float)be;
      const float H = fmaxf(fmaxf(fabsf(fa), fabsf(fb)), fabsf(fa - fb));
      float2 F, G0, G1, G2;
      if (H < kLaneTaylor) {
        tri_taylor_adj<7>(al, be, Q0, F, G0, G1, G2);
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
      wave_totals12_add(v, acc[f], lane);
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

namespace {

__global__ void __launch_bounds__(256) k_affine_dof_grads(const AffineGradJob* __restrict__ jobs) {
  __shared__ double red[12][256];
  const AffineGradJob jb = jobs[blockIdx.x];
  const int tid = threadIdx.x;
  double acc[12];
#pragma unroll
  for (int t = 0; t < 12; ++t) acc[t] = 0.0;
  for (int v = tid; v < jb.nv; v += blockDim.x) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const double g = jb.gw[3 * (size_t)v + r];
      acc[r] += g;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        acc[3 + 3 * r + c] = fma(g, jb.verts[3 * (size_t)v + c], acc[3 + 3 * r + c]);
    }
  }
#pragma unroll
  for (int t = 0; t < 12; ++t) red[t][tid] = acc[t];
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (tid < h)
#pragma unroll
      for (int t = 0; t < 12; ++t) red[t][tid] += red[t][tid + h];
    __syncthreads();
  }
  if (tid < 12) jb.out[tid] = red[tid][0];
}

}  // namespace

void launch_affine_dof_grads(hipStream_t s, const AffineGradJob* d_jobs, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_affine_dof_grads, dim3(n), dim3(256), 0, s, d_jobs);
}

}  // namespace pii
