// Spherical-harmonic coefficients of the triangle moment series, computed
// once per geometry (see sh_common.hpp for the expansion and the layout).
//
// For each geometry, in its reference frame about ref_c:
//   1. moment records {v0 - c, v1 - c, v2 - c, S} of every face;
//   2. the raw sums H_n, n <= kSeriesOrder, at the kNx x kNphi quadrature
//      directions (Gauss-Legendre nodes x_i in cos(theta), phi_j = 2 pi j /
//      kNphi), by the direct series kernel of moments.hip;
//   3. per order m: the Fourier sums F(i, n) = sum_j H_n(x_i, phi_j)
//      {cos, sin}(m phi_j), then a_nlm / b_nlm = c_m (2 pi / kNphi)
//      sum_i w_i Pt_l^m(x_i) F(i, n), c_0 = 1, c_m = 2.  H_n on a ring is a
//      trigonometric polynomial of degree <= kL in phi, so the sum over the
//      kNphi = 2 kL + 1 points is exact; in x the integrand is a polynomial
//      of degree <= 2 kL, which kNx = kL + 1 Gauss-Legendre nodes integrate
//      exactly.
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.hpp"
#include "pii_capi_internal.hpp"
#include "pii_host.hpp"
#include "pii_internal.hpp"
#include "sh_common.hpp"

namespace pii {

namespace {

using sh::kL;
using sh::kNphi;
using sh::kNq;
using sh::kNx;
constexpr int P = kSeriesOrder;

__device__ constexpr sh::Recurrence kRec{};
__device__ constexpr sh::Layout kLay{};

// Step 1.  Cache descriptors: verts / faces of the geometry, rctr = ref_c,
// mrec the records, n faces.
__global__ void __launch_bounds__(256) k_sh_records(const ObjDesc* __restrict__ objs) {
  const ObjDesc& d = objs[blockIdx.y];
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < d.n; f += gridDim.x * blockDim.x) {
    const int i0 = d.faces[3 * f], i1 = d.faces[3 * f + 1], i2 = d.faces[3 * f + 2];
    const double* V = d.verts;
    const double ax = V[3 * i0], ay = V[3 * i0 + 1], az = V[3 * i0 + 2];
    const double bx = V[3 * i1], by = V[3 * i1 + 1], bz = V[3 * i1 + 2];
    const double cx = V[3 * i2], cy = V[3 * i2 + 1], cz = V[3 * i2 + 2];
    const double e1x = bx - ax, e1y = by - ay, e1z = bz - az;
    const double e2x = cx - ax, e2y = cy - ay, e2z = cz - az;
    double* m = d.mrec + (size_t)kRecMoment * f;
    const double c0 = d.rctr[0], c1 = d.rctr[1], c2 = d.rctr[2];
    m[0] = ax - c0; m[1] = ay - c1; m[2] = az - c2;
    m[3] = bx - c0; m[4] = by - c1; m[5] = bz - c2;
    m[6] = cx - c0; m[7] = cy - c1; m[8] = cz - c2;
    // S = (e1 x e2) / 2, as the pre-pass
    m[9] = 0.5 * (e1y * e2z - e1z * e2y);
    m[10] = 0.5 * (e1z * e2x - e1x * e2z);
    m[11] = 0.5 * (e1x * e2y - e1y * e2x);
  }
}

// Step 3: one workgroup per (order m, object).  H: the raw sums of step 2
// (nblk blocks of 64 directions per object); xw: x_i then w_i; the
// coefficients go to the cache descriptor's rec pointer.
__global__ void __launch_bounds__(256) k_sh_project(const ObjDesc* __restrict__ objs,
                                                    const double* __restrict__ H,
                                                    const double* __restrict__ xw, int nblk) {
  __shared__ double tw[2][kNphi];         // cos(m phi_j), sin(m phi_j)
  __shared__ double F[kNx][P + 1];        // Fourier sums of one component
  __shared__ double Pt[kNx][kL + 1];      // c_m (2 pi / kNphi) w_i Pt_l^m(x_i)
  const int m = blockIdx.x;
  const ObjDesc& d = objs[blockIdx.y];
  double* out = d.rec;
  const int tid = threadIdx.x;
  for (int j = tid; j < kNphi; j += blockDim.x) {
    // m phi_j reduced exactly: 2 pi ((m j) mod kNphi) / kNphi
    double s, c;
    sincospi(2.0 * (double)((m * j) % kNphi) / kNphi, &s, &c);
    tw[0][j] = c;
    tw[1][j] = s;
  }
  const double cm = (m == 0 ? 1.0 : 2.0) * 2.0 * kPi / kNphi;
  for (int i = tid; i < kNx; i += blockDim.x) {
    const double x = xw[i], wt = xw[kNx + i];
    const double st = sqrt(fmax(0.0, (1.0 - x) * (1.0 + x)));
    double sm = 1.0;  // sin^m(theta)
    for (int t = 0; t < m; ++t) sm *= st;
    double qa = 0.0, qb = 0.0;
    for (int l = 0; l <= kL; ++l) {
      double q = 0.0;
      if (l >= m) {
        q = l == m ? kRec.seed[m] : kRec.a[l][m] * fma(x, qa, -kRec.b[l][m] * qb);
        qb = qa;
        qa = q;
      }
      Pt[i][l] = q * sm * wt * cm;
    }
  }
  __syncthreads();
  const double* Ho = H + (size_t)blockIdx.y * nblk * (P + 1) * 64;
  const int base = kLay.C[m];
  const int npair = kLay.C[kL + 1] - base;
  double* om = out + 2 * (size_t)kLay.mstart[m];
  for (int comp = 0; comp < 2; ++comp) {
    if (comp == 1 && m == 0) {
      for (int p = tid; p < npair; p += blockDim.x) om[2 * p + 1] = 0.0;
      break;
    }
    for (int t = tid; t < kNx * (P + 1); t += blockDim.x) {
      const int i = t / (P + 1), n = t - i * (P + 1);
      double acc = 0.0;
      for (int j = 0; j < kNphi; ++j) {
        const int q = i * kNphi + j;
        acc = fma(Ho[((size_t)(q >> 6) * (P + 1) + n) * 64 + (q & 63)], tw[comp][j], acc);
      }
      F[i][n] = acc;
    }
    __syncthreads();
    for (int p = tid; p < npair; p += blockDim.x) {
      int l = m;
      while (kLay.C[l + 1] - base <= p) ++l;
      const int n = sh::n0(l) + 2 * (p - (kLay.C[l] - base));
      double acc = 0.0;
      for (int i = 0; i < kNx; ++i) acc = fma(Pt[i][l], F[i][n], acc);
      om[2 * p + comp] = acc;
    }
    __syncthreads();
  }
}

}  // namespace

void cache_moments(Context& c, int n, GeometryDev* const* gs) {
  constexpr int kChunk = 32;
  hipStream_t s = c.stream;
  constexpr sh::Layout lay{};
  const size_t coef_bytes = sizeof(double) * 2 * (size_t)lay.total;
  // quadrature directions (rows 3..5 of a k-grid-like SoA, row stride Qd) and
  // the Gauss-Legendre nodes / weights (rows 0, 1)
  const int Qd = (kNq + 63) / 64 * 64;
  const int nblk = Qd / 64;
  std::vector<double> x(kNx), w(kNx);
  gauss_legendre(kNx, -1.0, 1.0, x.data(), w.data());
  std::vector<double> h_dirs(8 * (size_t)Qd, 0.0);
  for (int i = 0; i < kNx; ++i) {
    h_dirs[i] = x[i];
    h_dirs[kNx + i] = w[i];
    const double st = std::sqrt(std::max(0.0, (1.0 - x[i]) * (1.0 + x[i])));
    for (int j = 0; j < kNphi; ++j) {
      const int q = i * kNphi + j;
      const double phi = 2.0 * kPi * j / kNphi;
      h_dirs[3 * (size_t)Qd + q] = st * std::cos(phi);
      h_dirs[4 * (size_t)Qd + q] = st * std::sin(phi);
      h_dirs[5 * (size_t)Qd + q] = x[i];
    }
  }
  std::vector<GeometryDev*> todo;
  for (int i = 0; i < n; ++i)
    if (gs[i]->type == PII_GEOM_TRIANGLE && gs[i]->n >= kShMinFaces) todo.push_back(gs[i]);
  if (todo.empty()) return;
  DevBuf d_dirs, d_objs, d_mrec, d_H;
  d_dirs.ensure(sizeof(double) * h_dirs.size());
  PII_HIP(hipMemcpyAsync(d_dirs.p, h_dirs.data(), sizeof(double) * h_dirs.size(),
                         hipMemcpyHostToDevice, s));
  d_objs.ensure(sizeof(ObjDesc) * kChunk);
  d_H.ensure(sizeof(double) * (size_t)kChunk * nblk * (P + 1) * 64);
  std::vector<ObjDesc> h(kChunk);
  for (size_t c0 = 0; c0 < todo.size(); c0 += kChunk) {
    const int m = (int)std::min<size_t>(kChunk, todo.size() - c0);
    size_t rec = 0;
    int maxf = 0;
    for (int k = 0; k < m; ++k) {
      rec += align_up((size_t)todo[c0 + k]->n * kRecMoment, 32);
      maxf = std::max(maxf, todo[c0 + k]->n);
    }
    d_mrec.ensure(sizeof(double) * rec);
    rec = 0;
    for (int k = 0; k < m; ++k) {
      GeometryDev& g = *todo[c0 + k];
      g.shc.ensure(coef_bytes);
      ObjDesc& d = h[k];
      std::memset(&d, 0, sizeof(d));
      d.type = PII_GEOM_TRIANGLE;
      d.n = g.n;
      d.nv = g.nv;
      d.nser = 1;
      d.nord = kSeriesOrder;
      d.verts = g.verts.as<double>();
      d.faces = g.faces.as<int32_t>();
      for (int t = 0; t < 3; ++t) d.rctr[t] = d.ctr[t] = g.ref_c[t];
      d.mrec = d_mrec.as<double>() + rec;
      d.rec = g.shc.as<double>();   // coefficient output of k_sh_project
      rec += align_up((size_t)g.n * kRecMoment, 32);
    }
    PII_HIP(hipMemcpyAsync(d_objs.p, h.data(), sizeof(ObjDesc) * m, hipMemcpyHostToDevice, s));
    const ObjDesc* objs = d_objs.as<ObjDesc>();
    hipLaunchKernelGGL(k_sh_records, dim3(std::min((maxf + 255) / 256, 64), m), dim3(256), 0, s,
                       objs);
    launch_series_tri_raw(s, objs, m, d_dirs.as<double>(), Qd, kNq, d_H.as<double>());
    hipLaunchKernelGGL(k_sh_project, dim3(kL + 1, m), dim3(256), 0, s, objs,
                       d_H.as<double>(), d_dirs.as<double>(), nblk);
    PII_HIP(hipGetLastError());
    // the descriptors and records are rewritten by the next chunk
    PII_HIP(hipStreamSynchronize(s));
  }
}

}  // namespace pii
