// Affine-DoF derivatives of the form factor at every representative node,
// for the Gauss-Newton Hessian blocks of the volume matrix
// (cpp/src/compute_intersection_volume_multi_object.cpp:236-292):
//
//   H_ij = (1/8 pi^3) sum_q w_q Re( dA_i/dtheta_i (dA_j/dtheta_j)^H ),
//
// a weighted Gram over derivative columns.  Object i owns columns 12 i + t,
// t = [t_x, t_y, t_z, A_00, A_01, ..., A_22] (the AffineDoF layout), holding
// X_t(q) = sqrt(w_q) dA/dtheta_t(k_q) in FP32 and the column max |X|.  On a
// centrally symmetric grid dA(-k)/dtheta = -conj(dA(k)/dtheta), so the
// representative nodes with summed weights give the full sum, as for V.
//
// Triangles, world vertices v_j = A vref_j + t (the derivative of
// cpp/src/dof/affine_dof.cpp:171-232, taken through the vertices):
//   dA_face/dv_j = kappa G_j k + F khat x m_j,
//   m_0 = (e2 - e1)/2, m_1 = -e2/2, m_2 = e1/2,  F = 2 g[u], G_j = dF/du_j,
//   dA/dt = i k A,
//   dA/dA_rc = sum_faces [ kappa k_r sum_j G_j vref_j,c + F ([khat]_x N)_rc ],
//   N_bc = sum_j m_j,b vref_j,c  (per face, from dtheta records).
// Disks / Gaussians (affine_dof.cpp:233-316): with k' = A^T k, c' = A c + t,
// reference centre c and normal n,
//   dA/dt = i k A_e,
//   dA/dA_rc = e^{i k.c'} [ i A_par k_r c_c + S khat_r n_c
//                           + (k'.n / |k|) dS k_r (k'_c - (k'.n) n_c) ],
//   disk: S = 2 area J1(x)/x (area for x < 1e-10), x = rho r,
//         dS = 2 rho area (J1(x)/x)'/r (-rho^2 area / 4 for x < 1e-10);
//   Gaussian: S = w exp(-sigma^2 r^2 / 2), dS = -sigma^2 S.
#include "device_common.hpp"
#include "pii_internal.hpp"
#include "tri_derivs.hpp"
#include "dtheta_columns.hpp"
#include "bessel_f64.hpp"

namespace pii {

namespace {

constexpr int kLanes = 64;
constexpr int kWaves = 4;
constexpr int kRecDtheta = 18;  // floats per face: vref_0..2 (9), N (9, row b, column c)

// Per face (patch order): reference corners and N_bc = sum_j m_j,b vref_j,c.
__global__ void __launch_bounds__(256) k_dtheta_records(const ObjDesc* __restrict__ objs,
                                                        const int* __restrict__ eprefix, int N,
                                                        int total, float* __restrict__ out,
                                                        const long long* __restrict__ ooff) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int o = find_owner(eprefix, N, i);
  const ObjDesc& d = objs[o];
  if (d.type != PII_GEOM_TRIANGLE) return;
  const int e = i - eprefix[o];
  const int f = (int)d.patch_faces[e];
  const int i0 = d.faces[3 * f], i1 = d.faces[3 * f + 1], i2 = d.faces[3 * f + 2];
  const double* W = d.world;
  double e1[3], e2[3], m[3][3], vr[3][3];
  for (int c = 0; c < 3; ++c) {
    e1[c] = W[3 * i1 + c] - W[3 * i0 + c];
    e2[c] = W[3 * i2 + c] - W[3 * i0 + c];
    vr[0][c] = d.verts[3 * i0 + c];
    vr[1][c] = d.verts[3 * i1 + c];
    vr[2][c] = d.verts[3 * i2 + c];
  }
  for (int c = 0; c < 3; ++c) {
    m[0][c] = 0.5 * (e2[c] - e1[c]);
    m[1][c] = -0.5 * e2[c];
    m[2][c] = 0.5 * e1[c];
  }
  float* r = out + ooff[o] + (size_t)kRecDtheta * e;
  for (int j = 0; j < 3; ++j)
    for (int c = 0; c < 3; ++c) r[3 * j + c] = (float)vr[j][c];
  for (int b = 0; b < 3; ++b)
    for (int c = 0; c < 3; ++c)
      r[9 + 3 * b + c] = (float)(m[0][b] * vr[0][c] + m[1][b] * vr[1][c] + m[2][b] * vr[2][c]);
}


// First node of the per-element derivative kernels: the derivative series
// covers the nodes below qbeg of its objects.
__device__ __forceinline__ int dtheta_first_node(const ObjDesc& d) {
  return (d.type == PII_GEOM_TRIANGLE && d.dser) ? d.qbeg : 0;
}

// One workgroup per (object, 64 nodes from dtheta_first_node[, element
// split]).  Split launches store the 20 sums per node to part[((z * nobj +
// obj) * nblk * 64 + node) * 20 + i] (node counted from the first node) for
// k_dtheta_finish.
__global__ void __launch_bounds__(kLanes * kWaves) k_dtheta_affine(
    const ObjDesc* __restrict__ objs, const double* __restrict__ krep, int Qh,
    const float* __restrict__ drec, const long long* __restrict__ doff, float2* __restrict__ Xt,
    int ldX, unsigned* __restrict__ colmax, double* __restrict__ part) {
  __shared__ float2 tab[kPatchVerts][kLanes];
  const BlockObj bo = xcd_block_map();
  const ObjDesc& d = objs[bo.obj];
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // objects with a derivative series start after its shells (uniform exit)
  const int q0 = dtheta_first_node(d) + bo.blk * kLanes;
  if (q0 >= ldX) return;
  const int q = q0 + lane;
  const int qc = q < Qh ? q : Qh - 1;
  const double kx = krep[qc], ky = krep[(size_t)Qh + qc], kz = krep[2 * (size_t)Qh + qc];
  const float hx = (float)krep[3 * (size_t)Qh + qc], hy = (float)krep[4 * (size_t)Qh + qc],
              hz = (float)krep[5 * (size_t)Qh + qc];
  const double km = krep[6 * (size_t)Qh + qc];
  const float fk[3] = {(float)kx, (float)ky, (float)kz};
  const float fh[3] = {hx, hy, hz};
  // accumulators: A (for dA/dt = i k A) and dA/dA_rc, FP32 per patch / element
  // block, FP64 across them in LDS (acc[wave][i][lane], i = 0..19: A re, im,
  // then dA/dA_rc re, im): FP64 registers for them would push the kernel past
  // 256 VGPRs, i.e. to one wave per SIMD
  __shared__ double acc[kWaves][20][kLanes];
#pragma unroll
  for (int i = 0; i < 20; ++i) acc[wave][i][lane] = 0.0;
  auto flush = [&](const float (&pA)[2], const float (&pg)[9][2]) {
    acc[wave][0][lane] += (double)pA[0];
    acc[wave][1][lane] += (double)pA[1];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      acc[wave][2 + 2 * i][lane] += (double)pg[i][0];
      acc[wave][3 + 2 * i][lane] += (double)pg[i][1];
    }
  };

  if (d.type == PII_GEOM_TRIANGLE) {
    const auto* hdr = as_const(d.patch_hdr);
    const auto* pverts = as_const(d.patch_verts);
    const auto* world = as_const(d.world);
    const auto* rec = as_const(d.rec);
    const auto* dr = as_const(drec) + doff[bo.obj];
    const int pchunk = (d.npatch + (int)gridDim.z - 1) / (int)gridDim.z;
    const int pb = (int)blockIdx.z * pchunk, pe = min(d.npatch, pb + pchunk);
    for (int p = pb; p < pe; ++p) {
      const int vbeg = hdr[4 * p], vcnt = hdr[4 * p + 1];
      const int fbeg = hdr[4 * p + 2], fcnt = hdr[4 * p + 3];
      for (int s = wave; s < vcnt; s += kWaves) {
        const int gv = pverts[vbeg + s];
        const double u = kx * world[3 * gv] + ky * world[3 * gv + 1] + kz * world[3 * gv + 2];
        tab[s][lane] = half_phasor_a(u);
      }
      __syncthreads();
      // per-patch sums: kappa F, S_c = sum_f kappa_f sum_j G_j vref_j,c and
      // M = sum_f F_f N_f; the node's k and [khat]_x, the same for every
      // face, are applied once per patch (dA/dA_rc = k_r S_c + ([khat]_x M)_rc)
      float pA[2] = {0.f, 0.f}, pS[3][2] = {}, pM[9][2] = {};
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
        const auto* t = dr + (size_t)kRecDtheta * (fbeg + f);  // vref (9), N (9)
        pA[0] += kap * F.x;
        pA[1] += kap * F.y;
        const float2 K0 = make_float2(kap * G0.x, kap * G0.y);
        const float2 K1 = make_float2(kap * G1.x, kap * G1.y);
        const float2 K2 = make_float2(kap * G2.x, kap * G2.y);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          // kappa sum_j G_j vref_j,c
          pS[c][0] += K0.x * t[c] + K1.x * t[3 + c] + K2.x * t[6 + c];
          pS[c][1] += K0.y * t[c] + K1.y * t[3 + c] + K2.y * t[6 + c];
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
          pM[i][0] += F.x * t[9 + i];
          pM[i][1] += F.y * t[9 + i];
        }
      }
      float pg[9][2];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          // ([khat]_x M)_rc from column c of M
          const float m0 = pM[c][e], m1 = pM[3 + c][e], m2 = pM[6 + c][e];
          pg[c][e] = fk[0] * pS[c][e] + (fh[1] * m2 - fh[2] * m1);
          pg[3 + c][e] = fk[1] * pS[c][e] + (fh[2] * m0 - fh[0] * m2);
          pg[6 + c][e] = fk[2] * pS[c][e] + (fh[0] * m1 - fh[1] * m0);
        }
      flush(pA, pg);
      __syncthreads();
    }
  } else {
    const auto* p = as_const(d.dofs);
    const auto* rec = as_const(d.rec);
    const auto* el = as_const(d.el);
    const bool disk = d.type == PII_GEOM_DISK;
    // k' = A^T k
    const double kp[3] = {p[3] * kx + p[6] * ky + p[9] * kz, p[4] * kx + p[7] * ky + p[10] * kz,
                          p[5] * kx + p[8] * ky + p[11] * kz};
    const double kp2 = kp[0] * kp[0] + kp[1] * kp[1] + kp[2] * kp[2];
    const double inv_km = km > 0.0 ? 1.0 / km : 0.0;
    float pA[2] = {0.f, 0.f}, pg[9][2] = {};
    int cnt = 0;
    const int echunk = (d.n + (int)gridDim.z - 1) / (int)gridDim.z;
    const int eb = (int)blockIdx.z * echunk, ee = min(d.n, eb + echunk);
    for (int e = eb + wave; e < ee; e += kWaves) {
      const auto* r = rec + (size_t)kRecSurfel * e;   // c' (world), A n, rho|sigma, area|w
      const auto* s = el + (size_t)kRecSurfel * e;    // c, n (reference), rho|sigma, area|w
      const double kdn = kp[0] * s[3] + kp[1] * s[4] + kp[2] * s[5];
      double r2 = kp2 - kdn * kdn;
      double S, dS;
      if (disk) {
        r2 = fmax(0.0, r2);
        const double rr = sqrt(r2);
        const double x = s[6] * rr;
        if (x < 1e-10) {
          S = s[7];
          dS = -0.25 * s[6] * s[6] * s[7];
        } else {
          double f, fp;
          j1ox_and_prime(x, f, fp);
          S = 2.0 * s[7] * f;
          dS = 2.0 * fp / rr * s[6] * s[7];
        }
      } else {
        S = s[7] * exp(-0.5 * s[6] * s[6] * r2);
        dS = -s[6] * s[6] * S;
      }
      const double Apar = S * kdn * inv_km;
      const double tau = kdn * inv_km * dS;
      const float2 ph = half_phasor_a(2.0 * (kx * r[0] + ky * r[1] + kz * r[2]));
      // e^{i u} from the half-phasor routine on 2u
      const float fA = (float)Apar;
      pA[0] += fA * ph.x;
      pA[1] += fA * ph.y;
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float im = (float)(Apar * (rr == 0 ? kx : rr == 1 ? ky : kz) * s[c]);
          const float re = (float)(S * (rr == 0 ? hx : rr == 1 ? hy : hz) * s[3 + c] +
                                   tau * (rr == 0 ? kx : rr == 1 ? ky : kz) *
                                       (kp[c] - kdn * s[3 + c]));
          // (re + i im) e^{i u}
          pg[3 * rr + c][0] += re * ph.x - im * ph.y;
          pg[3 * rr + c][1] += re * ph.y + im * ph.x;
        }
      }
      if (++cnt == 64) {
        flush(pA, pg);
#pragma unroll
        for (int i = 0; i < 9; ++i) pg[i][0] = pg[i][1] = 0.f;
        pA[0] = pA[1] = 0.f;
        cnt = 0;
      }
    }
    flush(pA, pg);
  }

  // Sum the four waves' partial sums of the same 64 nodes in a fixed order,
  // ((w0 + w1) + w2) + w3 (each wave reads only what it wrote itself before
  // the barrier).
  __syncthreads();
  if (wave != 0) return;
  double v[20];
#pragma unroll
  for (int i = 0; i < 20; ++i)
    v[i] = ((acc[0][i][lane] + acc[1][i][lane]) + acc[2][i][lane]) + acc[3][i][lane];
  if (km < 1e-10) {  // dA/dtheta = 0 at k = 0 (affine_dof.cpp:181)
#pragma unroll
    for (int i = 0; i < 20; ++i) v[i] = 0.0;
  }
  if (part) {
    double* o = part + (((size_t)blockIdx.z * gridDim.y + bo.obj) * gridDim.x * kLanes +
                        (size_t)bo.blk * kLanes + lane) * 20;
#pragma unroll
    for (int i = 0; i < 20; ++i) o[i] = v[i];
    return;
  }
  write_columns(v, q, Qh, kx, ky, kz, krep, d.col, Xt, ldX, colmax, lane);
}

// Finishes a split launch: each of the four waves sums every fourth split of
// the nodes' partials (splits in increasing order), the four sums are added
// in the fixed order ((w0 + w1) + w2) + w3 through LDS, and wave 0 writes the
// columns.  One workgroup per (object, 64 nodes).
__global__ void __launch_bounds__(kLanes * kWaves) k_dtheta_finish(
    const ObjDesc* __restrict__ objs, const double* __restrict__ krep, int Qh,
    float2* __restrict__ Xt, int ldX, unsigned* __restrict__ colmax,
    const double* __restrict__ part, int nsplit) {
  __shared__ double red[kWaves][20][kLanes];
  const int obj = blockIdx.y, blk = blockIdx.x;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const ObjDesc& d = objs[obj];
  const int q0 = dtheta_first_node(d) + blk * kLanes;
  if (q0 >= ldX) return;  // uniform over the workgroup
  const int q = q0 + lane;
  double v[20];
#pragma unroll
  for (int i = 0; i < 20; ++i) v[i] = 0.0;
  const size_t stride = (size_t)gridDim.y * gridDim.x * kLanes * 20;
  const double* src = part + ((size_t)obj * gridDim.x * kLanes + (size_t)blk * kLanes + lane) * 20;
  for (int z = wave; z < nsplit; z += kWaves) {
#pragma unroll
    for (int i = 0; i < 20; ++i) v[i] += src[(size_t)z * stride + i];
  }
#pragma unroll
  for (int i = 0; i < 20; ++i) red[wave][i][lane] = v[i];
  __syncthreads();
  if (wave != 0) return;
  const int qc = q < Qh ? q : Qh - 1;
  const double kx = krep[qc], ky = krep[(size_t)Qh + qc], kz = krep[2 * (size_t)Qh + qc];
#pragma unroll
  for (int i = 0; i < 20; ++i) v[i] = ((red[0][i][lane] + red[1][i][lane]) + red[2][i][lane]) +
                                      red[3][i][lane];
  write_columns(v, q, Qh, kx, ky, kz, krep, d.col, Xt, ldX, colmax, lane);
}

}  // namespace

void launch_dtheta_records(hipStream_t s, const ObjDesc* d_objs, const int* d_eprefix, int N,
                           int total_e, float* d_out, const long long* d_off) {
  if (total_e <= 0) return;
  hipLaunchKernelGGL(k_dtheta_records, dim3((total_e + 255) / 256), dim3(256), 0, s, d_objs,
                     d_eprefix, N, total_e, d_out, d_off);
}

int dtheta_slots() {
  static std::atomic<int> cache[16];
  return resident_workgroups(k_dtheta_affine, kLanes * kWaves, cache);
}

size_t dtheta_split_doubles(int N, int ldX, int nsplit) {
  return nsplit > 1 ? (size_t)nsplit * N * ((ldX + kLanes - 1) / kLanes) * kLanes * 20 : 0;
}

void launch_dtheta_affine(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_krep,
                          int Qh, const float* d_drec, const long long* d_doff, float2* d_Xt,
                          int ldX, unsigned* d_colmax, int nsplit, double* d_part) {
  if (N <= 0 || Qh <= 0) return;
  const int nblk = (ldX + kLanes - 1) / kLanes;
  const bool split = nsplit > 1 && d_part;
  double* part = split ? d_part : nullptr;
  hipLaunchKernelGGL(k_dtheta_affine, dim3(nblk, N, split ? nsplit : 1), dim3(kLanes * kWaves), 0,
                     s, d_objs, d_krep, Qh, d_drec, d_doff, d_Xt, ldX, d_colmax, part);
  if (split)
    hipLaunchKernelGGL(k_dtheta_finish, dim3(nblk, N), dim3(kLanes * kWaves), 0, s, d_objs,
                       d_krep, Qh, d_Xt, ldX, d_colmax, part, nsplit);
}

}  // namespace pii
