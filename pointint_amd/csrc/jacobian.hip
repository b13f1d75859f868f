// dA(k_q)/dtheta at every node of a k-grid, in FP64 with the reference's
// element derivatives: the Jacobian that the reference's gradient plug-ins
// write (CudaComputeAkGradientFunc, cpp/include/dof/cuda/dof_cuda_interface.hpp:34-40:
// compute_Ak_gradient_affine_{triangle,disk,gaussian}_cuda_wrapper,
// cpp/src/dof/cuda/affine_dof_cuda.cu:932-1329, and
// compute_Ak_gradient_triangle_mesh_cuda_wrapper,
// cpp/src/dof/cuda/triangle_mesh_dof_cuda.cu:387-466), with the maths of
// cpp/src/dof/affine_dof.cpp:172-322 and cpp/src/dof/triangle_mesh_dof.cpp:76-143.
// Output: row-major Q x D complex, out[q * D + t], or (for the Gram of
// derivative columns) column-major, out[t * ld + q], on the representative
// nodes, then scaled by sqrt(w) into FP32 columns.
//
// AffineDoF: one lane per node streams the object's elements (wave-uniform,
// so through the scalar cache) and keeps the 12 complex sums in registers.
// TriangleMeshDoF: one lane per (node, face) adds the face's three vertex
// derivatives to out[q][3 dv + c] with FP64 atomics (the output cleared
// first).  The volume path never builds this matrix (it contracts with the
// adjoint instead); this is the reference's testing helper and plug-in
// output, sized Q x D by the caller's choice.
#include "bessel_f64.hpp"
#include "device_common.hpp"
#include "pii_internal.hpp"

namespace pii {

namespace {

// dPhi/da, dPhi/db in FP64 (cpp/src/element_functions.cpp:73-93): -i E''((a+b)/2)
// for both when |b - a| < 1e-3, else -2i (t - E'(a)/d) and 2i (t - E'(b)/d)
// with t = (E(b) - E(a))/d^2, d = b - a.
__device__ __forceinline__ void ref_Phi_grad(double a, double b, cplx& dA, cplx& dB) {
  const double d = b - a;
  if (fabs(d) < 1e-3) {
    const cplx e = ref_Epp(0.5 * (a + b));
    dA = {e.y, -e.x};
    dB = dA;
    return;
  }
  const cplx eb = ref_E(b), ea = ref_E(a);
  const cplx pa = ref_Ep(a), pb = ref_Ep(b);
  const double inv = 1.0 / d, inv2 = inv * inv;
  const double tx = (eb.x - ea.x) * inv2, ty = (eb.y - ea.y) * inv2;
  const double ux = tx - pa.x * inv, uy = ty - pa.y * inv;  // -2i u
  const double vx = tx - pb.x * inv, vy = ty - pb.y * inv;  //  2i v
  dA = {2.0 * uy, -2.0 * ux};
  dB = {-2.0 * vy, 2.0 * vx};
}

__device__ __forceinline__ cplx cexpi(double p) {
  double s, c;
  sincos(p, &s, &c);
  return {c, s};
}

// AffineDoF (dofs [t, A row-major]): one lane per node q.
template <bool kColMajor>
__global__ void __launch_bounds__(256) k_jacobian_affine(const ObjDesc* __restrict__ objs,
                                                         const double* __restrict__ kfull, int Q,
                                                         int ld, double2* __restrict__ out) {
  const ObjDesc& d = objs[0];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  const double kv[3] = {kfull[q], kfull[(size_t)Q + q], kfull[2 * (size_t)Q + q]};
  const double km = kfull[6 * (size_t)Q + q];
  cplx g[12];
  for (int i = 0; i < 12; ++i) g[i] = {0.0, 0.0};
  if (km >= 1e-10) {  // dA/dtheta = 0 at k = 0 (affine_dof.cpp:181)
    const double* p = d.dofs;
    double A[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) A[r][c] = p[3 + 3 * r + c];
    const double t[3] = {p[0], p[1], p[2]};
    const double kh[3] = {kv[0] / km, kv[1] / km, kv[2] / km};
    if (d.type == PII_GEOM_TRIANGLE) {
      // KA = [khat]_x A, per lane; the third term per face is row r of KA
      // crossed with S (= KA [S]_x)
      double KA[3][3];
      for (int c = 0; c < 3; ++c) {
        KA[0][c] = -kh[2] * A[1][c] + kh[1] * A[2][c];
        KA[1][c] = kh[2] * A[0][c] - kh[0] * A[2][c];
        KA[2][c] = -kh[1] * A[0][c] + kh[0] * A[1][c];
      }
      for (int f = 0; f < d.n; ++f) {
        const double* V = d.verts;
        const int i0 = d.faces[3 * f], i1 = d.faces[3 * f + 1], i2 = d.faces[3 * f + 2];
        double a[3], e1[3], e2[3];
        for (int c = 0; c < 3; ++c) {
          a[c] = V[3 * i0 + c];
          e1[c] = V[3 * i1 + c] - a[c];
          e2[c] = V[3 * i2 + c] - a[c];
        }
        const double S[3] = {0.5 * (e1[1] * e2[2] - e1[2] * e2[1]),
                             0.5 * (e1[2] * e2[0] - e1[0] * e2[2]),
                             0.5 * (e1[0] * e2[1] - e1[1] * e2[0])};
        double ap[3], e1p[3], e2p[3];
        for (int r = 0; r < 3; ++r) {
          ap[r] = A[r][0] * a[0] + A[r][1] * a[1] + A[r][2] * a[2] + t[r];
          e1p[r] = A[r][0] * e1[0] + A[r][1] * e1[1] + A[r][2] * e1[2];
          e2p[r] = A[r][0] * e2[0] + A[r][1] * e2[1] + A[r][2] * e2[2];
        }
        const double Sp[3] = {0.5 * (e1p[1] * e2p[2] - e1p[2] * e2p[1]),
                              0.5 * (e1p[2] * e2p[0] - e1p[0] * e2p[2]),
                              0.5 * (e1p[0] * e2p[1] - e1p[1] * e2p[0])};
        const double al = kv[0] * e1p[0] + kv[1] * e1p[1] + kv[2] * e1p[2];
        const double be = kv[0] * e2p[0] + kv[1] * e2p[1] + kv[2] * e2p[2];
        const double ga = kh[0] * Sp[0] + kh[1] * Sp[1] + kh[2] * Sp[2];
        const cplx phi = ref_Phi(al, be);
        const cplx ph = cexpi(kv[0] * ap[0] + kv[1] * ap[1] + kv[2] * ap[2]);
        const cplx pp = cmul(ph, phi);                  // e^{i k.a'} Phi
        const cplx At = {pp.x * ga, pp.y * ga};         // A_tri
        cplx dPa, dPb;
        ref_Phi_grad(al, be, dPa, dPb);
        const cplx gpa = {ga * (ph.x * dPa.x - ph.y * dPa.y), ga * (ph.x * dPa.y + ph.y * dPa.x)};
        const cplx gpb = {ga * (ph.x * dPb.x - ph.y * dPb.y), ga * (ph.x * dPb.y + ph.y * dPb.x)};
        // dA/dt_r = i k_r A_tri
        for (int r = 0; r < 3; ++r) {
          g[r].x -= kv[r] * At.y;
          g[r].y += kv[r] * At.x;
        }
        for (int r = 0; r < 3; ++r) {
          // (KA row r) x S
          const double X[3] = {KA[r][1] * S[2] - KA[r][2] * S[1], KA[r][2] * S[0] - KA[r][0] * S[2],
                               KA[r][0] * S[1] - KA[r][1] * S[0]};
          for (int c = 0; c < 3; ++c) {
            // i k_r a_c A_tri + ga e^{i k.a'} k_r (dPa e1_c + dPb e2_c) - X_c e^{i k.a'} Phi
            const double ka = kv[r] * a[c];
            g[3 + 3 * r + c].x += -ka * At.y + kv[r] * (gpa.x * e1[c] + gpb.x * e2[c]) - X[c] * pp.x;
            g[3 + 3 * r + c].y += ka * At.x + kv[r] * (gpa.y * e1[c] + gpb.y * e2[c]) - X[c] * pp.y;
          }
        }
      }
    } else {
      // k' = A^T k
      const double kp[3] = {A[0][0] * kv[0] + A[1][0] * kv[1] + A[2][0] * kv[2],
                            A[0][1] * kv[0] + A[1][1] * kv[1] + A[2][1] * kv[2],
                            A[0][2] * kv[0] + A[1][2] * kv[1] + A[2][2] * kv[2]};
      const double kp2 = kp[0] * kp[0] + kp[1] * kp[1] + kp[2] * kp[2];
      const bool disk = d.type == PII_GEOM_DISK;
      for (int e = 0; e < d.n; ++e) {
        const double* s = d.el + (size_t)kRecSurfel * e;  // c, n (reference), rho|sigma, area|w
        const double kdn = kp[0] * s[3] + kp[1] * s[4] + kp[2] * s[5];
        double S, dS;
        if (disk) {  // affine_dof.cpp:240-269
          const double r2 = fmax(0.0, kp2 - kdn * kdn);
          const double r = sqrt(r2), x = s[6] * r;
          if (x < 1e-10) {
            S = s[7];
            dS = -0.25 * s[6] * s[6] * s[7];
          } else {
            double fx, fpx;
            j1ox_and_prime(x, fx, fpx);
            S = 2.0 * s[7] * fx;
            dS = 2.0 * fpx / r * s[6] * s[7];
          }
        } else {     // affine_dof.cpp:284-311 (no clamp of r^2)
          const double r2 = kp2 - kdn * kdn;
          S = s[7] * exp(-0.5 * s[6] * s[6] * r2);
          dS = -s[6] * s[6] * S;
        }
        double cp[3];
        for (int r = 0; r < 3; ++r) cp[r] = A[r][0] * s[0] + A[r][1] * s[1] + A[r][2] * s[2] + t[r];
        const cplx ph = cexpi(kv[0] * cp[0] + kv[1] * cp[1] + kv[2] * cp[2]);
        const double Ap = S * kdn / km;
        const cplx Ae = {ph.x * Ap, ph.y * Ap};
        const double tau = kdn / km * dS;
        const double kperp[3] = {kp[0] - kdn * s[3], kp[1] - kdn * s[4], kp[2] - kdn * s[5]};
        for (int r = 0; r < 3; ++r) {
          g[r].x -= kv[r] * Ae.y;
          g[r].y += kv[r] * Ae.x;
        }
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            // i k_r c_c A_e + khat_r n_c e^{i k.c'} S + e^{i k.c'} (k'.n/|k|) dS k_r kperp_c
            const double ka = kv[r] * s[c];
            const double re = kh[r] * s[3 + c] * S + tau * kv[r] * kperp[c];
            g[3 + 3 * r + c].x += -ka * Ae.y + re * ph.x;
            g[3 + 3 * r + c].y += ka * Ae.x + re * ph.y;
          }
      }
    }
  }
  for (int i = 0; i < 12; ++i)
    out[kColMajor ? (size_t)i * ld + q : (size_t)q * 12 + i] = make_double2(g[i].x, g[i].y);
}

// TriangleMeshDoF (dofs: the world positions of the DoF vertices): one lane
// per (node, face).  Row-major output: the lanes of a wave take consecutive
// faces of one node; column-major: consecutive nodes of one face, so that
// their atomics go to consecutive addresses of one column.
template <bool kColMajor>
__global__ void __launch_bounds__(256) k_jacobian_mesh(const ObjDesc* __restrict__ objs,
                                                       const double* __restrict__ kfull, int Q,
                                                       int D, int ld, double* __restrict__ out) {
  const ObjDesc& d = objs[0];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)Q * d.n) return;
  const int q = kColMajor ? (int)(i % Q) : (int)(i / d.n);
  const int f = kColMajor ? (int)(i / Q) : (int)(i % d.n);
  const double km = kfull[6 * (size_t)Q + q];
  if (km < 1e-10) return;
  int dv[3];
  for (int j = 0; j < 3; ++j) dv[j] = d.dofvid[d.faces[3 * f + j]];
  if (dv[0] < 0 || dv[1] < 0 || dv[2] < 0) return;  // triangle_mesh_dof.cpp:88
  const double kv[3] = {kfull[q], kfull[(size_t)Q + q], kfull[2 * (size_t)Q + q]};
  const double kh[3] = {kv[0] / km, kv[1] / km, kv[2] / km};
  const double* X = d.dofs;
  double v0[3], e1[3], e2[3];
  for (int c = 0; c < 3; ++c) {
    v0[c] = X[3 * dv[0] + c];
    e1[c] = X[3 * dv[1] + c] - v0[c];
    e2[c] = X[3 * dv[2] + c] - v0[c];
  }
  const double sn[3] = {0.5 * (e1[1] * e2[2] - e1[2] * e2[1]), 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]),
                        0.5 * (e1[0] * e2[1] - e1[1] * e2[0])};
  const double kap = kh[0] * sn[0] + kh[1] * sn[1] + kh[2] * sn[2];
  // dkappa/dv_j: khat x (e2 - e1)/2, -khat x e2/2, khat x e1/2
  const double m[3][3] = {{0.5 * (e2[0] - e1[0]), 0.5 * (e2[1] - e1[1]), 0.5 * (e2[2] - e1[2])},
                          {-0.5 * e2[0], -0.5 * e2[1], -0.5 * e2[2]},
                          {0.5 * e1[0], 0.5 * e1[1], 0.5 * e1[2]}};
  const cplx T = cexpi(kv[0] * v0[0] + kv[1] * v0[1] + kv[2] * v0[2]);
  const double ke1 = kv[0] * e1[0] + kv[1] * e1[1] + kv[2] * e1[2];
  const double ke2 = kv[0] * e2[0] + kv[1] * e2[1] + kv[2] * e2[2];
  const cplx psi = ref_Phi(ke1, ke2);
  cplx dPa, dPb;
  ref_Phi_grad(ke1, ke2, dPa, dPb);
  // coefficients of kappa k: i psi - dPa - dPb, dPa, dPb
  const cplx coef[3] = {{-psi.y - dPa.x - dPb.x, psi.x - dPa.y - dPb.y}, dPa, dPb};
  for (int j = 0; j < 3; ++j) {
    const double dk[3] = {kh[1] * m[j][2] - kh[2] * m[j][1], kh[2] * m[j][0] - kh[0] * m[j][2],
                          kh[0] * m[j][1] - kh[1] * m[j][0]};
    for (int c = 0; c < 3; ++c) {
      // T (psi dkappa_c + kappa k_c coef_j)
      const double ux = psi.x * dk[c] + kap * kv[c] * coef[j].x;
      const double uy = psi.y * dk[c] + kap * kv[c] * coef[j].y;
      const size_t t = 3 * (size_t)dv[j] + c;
      double* o = out + 2 * (kColMajor ? t * ld + q : (size_t)q * D + t);
      atomicAdd(o, T.x * ux - T.y * uy);
      atomicAdd(o + 1, T.x * uy + T.y * ux);
    }
  }
}

// X column t = sqrt(w_q) J[:, t] in FP32 for the representative nodes q < Qh,
// zero up to ldX, and colmax[t] = the largest |component| of the column (as
// the form-factor columns).  Grid (ldX / 256 rounded up, D), whole waves.
__global__ void __launch_bounds__(256) k_jacobian_to_x(const double2* __restrict__ J, int ldJ,
                                                       const double* __restrict__ krep, int Qh,
                                                       float2* __restrict__ X, int ldX,
                                                       unsigned* __restrict__ colmax) {
  const int t = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  float m = 0.0f;
  if (q < ldX) {
    float2 x = make_float2(0.0f, 0.0f);
    if (q < Qh) {
      const double sw = sqrt(krep[7 * (size_t)Qh + q]);
      const double2 v = J[(size_t)t * ldJ + q];
      x = make_float2((float)(v.x * sw), (float)(v.y * sw));
      m = fmaxf(fabsf(x.x), fabsf(x.y));
    }
    X[(size_t)t * ldX + q] = x;
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(colmax + t, __float_as_uint(m));
}

}  // namespace

void launch_jacobian_columns(hipStream_t s, const ObjDesc* d_obj, int dof_kind, int n_elems,
                             const double* d_krep, int Qh, int D, double2* d_J, int ldJ,
                             float2* d_X, int ldX, unsigned* d_colmax) {
  if (Qh <= 0 || D <= 0) return;
  if (dof_kind == PII_DOF_AFFINE) {
    hipLaunchKernelGGL(k_jacobian_affine<true>, dim3((Qh + 255) / 256), dim3(256), 0, s, d_obj,
                       d_krep, Qh, ldJ, d_J);
  } else {
    PII_HIP(hipMemsetAsync(d_J, 0, sizeof(double2) * (size_t)ldJ * D, s));
    const long long total = (long long)Qh * n_elems;
    if (total > 0)
      hipLaunchKernelGGL(k_jacobian_mesh<true>, dim3((unsigned)((total + 255) / 256)), dim3(256),
                         0, s, d_obj, d_krep, Qh, D, ldJ, reinterpret_cast<double*>(d_J));
  }
  hipLaunchKernelGGL(k_jacobian_to_x, dim3((ldX + 255) / 256, D), dim3(256), 0, s, d_J, ldJ,
                     d_krep, Qh, d_X, ldX, d_colmax);
}

void launch_jacobian(hipStream_t s, const ObjDesc* d_obj, int type, int dof_kind, int n_elems,
                     const double* d_kfull, int Q, int D, double2* d_out) {
  if (Q <= 0 || D <= 0) return;
  if (dof_kind == PII_DOF_AFFINE) {
    hipLaunchKernelGGL(k_jacobian_affine<false>, dim3((Q + 255) / 256), dim3(256), 0, s, d_obj,
                       d_kfull, Q, 0, d_out);
  } else {
    PII_HIP(hipMemsetAsync(d_out, 0, sizeof(double2) * (size_t)Q * D, s));
    const long long total = (long long)Q * n_elems;
    if (total > 0)
      hipLaunchKernelGGL(k_jacobian_mesh<false>, dim3((unsigned)((total + 255) / 256)), dim3(256),
                         0, s, d_obj, d_kfull, Q, D, 0, reinterpret_cast<double*>(d_out));
  }
  (void)type;
}

}  // namespace pii
