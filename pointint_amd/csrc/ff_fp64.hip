// Plain FP64 form factor: the reference formula evaluated in FP64 on the GPU.
//
// This is the "plain FP64 GPU path" the fast path is measured against and the
// parity anchor of the device code: same element maths as
// cpp/src/dof/affine_dof.cpp:55-170 and cpp/src/dof/triangle_mesh_dof.cpp:18-74
// (E/Phi with the 1e-4 / 1e-3 Taylor windows, FP64 sincos), over the full
// k-grid.  One lane owns one k-node; the wave streams the object's element
// records (wave-uniform, so they come through the scalar cache) and
// accumulates in FP64.  A launch with few workgroups (a pair of objects on a
// small grid) splits each object's elements over gridDim.z workgroups per
// node block; k_ff_fp64_finish adds their FP64 partial sums in split order.
#include "device_common.hpp"
#include "pii_internal.hpp"

namespace pii {

namespace {

__global__ void __launch_bounds__(256) k_ff_fp64(const ObjDesc* __restrict__ objs,
                                                 const double* __restrict__ kfull, int Q,
                                                 double2* __restrict__ J, int ldJ,
                                                 double2* __restrict__ part) {
  const ObjDesc& d = objs[blockIdx.y];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  const double kx = kfull[q], ky = kfull[(size_t)Q + q], kz = kfull[2 * (size_t)Q + q];
  const double hx = kfull[3 * (size_t)Q + q], hy = kfull[4 * (size_t)Q + q],
               hz = kfull[5 * (size_t)Q + q];
  const double km = kfull[6 * (size_t)Q + q];
  double ar = 0.0, ai = 0.0;
  // elements [eb, ee) of this split
  const int chunk = (d.n + (int)gridDim.z - 1) / (int)gridDim.z;
  const int eb = (int)blockIdx.z * chunk, ee = min(d.n, eb + chunk);
  if (km >= 1e-10) {
    if (d.type == PII_GEOM_TRIANGLE) {
      const double* rec = d.rec;
      for (int f = eb; f < ee; ++f) {
        const double* r = rec + (size_t)kRecPlainTri * f;
        const double al = kx * r[3] + ky * r[4] + kz * r[5];
        const double be = kx * r[6] + ky * r[7] + kz * r[8];
        const double ga = hx * r[9] + hy * r[10] + hz * r[11];
        const double ph = kx * r[0] + ky * r[1] + kz * r[2];
        const cplx phi = ref_Phi(al, be);
        double s, c;
        sincos(ph, &s, &c);
        // A_tri = gamma * e^{i ph} * Phi
        ar += ga * (c * phi.x - s * phi.y);
        ai += ga * (c * phi.y + s * phi.x);
      }
    } else {
      // |A^T k|^2 once per lane (cpp/src/dof/affine_dof.cpp:116)
      const double* p = d.dofs;
      const double kpx = p[3] * kx + p[6] * ky + p[9] * kz;
      const double kpy = p[4] * kx + p[7] * ky + p[10] * kz;
      const double kpz = p[5] * kx + p[8] * ky + p[11] * kz;
      const double kp2 = kpx * kpx + kpy * kpy + kpz * kpz;
      const double inv_km = 1.0 / km;
      const double* rec = d.rec;
      const bool disk = (d.type == PII_GEOM_DISK);
      for (int e = eb; e < ee; ++e) {
        const double* r = rec + (size_t)kRecSurfel * e;
        const double kdn = kx * r[3] + ky * r[4] + kz * r[5];
        double S;
        if (disk) {  // affine_dof.cpp:118-123
          const double r2 = fmax(0.0, kp2 - kdn * kdn);
          const double rr = r[6] * sqrt(r2);
          S = (rr < 1e-10) ? r[7] : 2.0 * r[7] * j1_over_x_f64(rr);
        } else {     // affine_dof.cpp:151-155 (no clamp, as on the CPU path)
          const double r2 = kp2 - kdn * kdn;
          S = r[7] * exp(-0.5 * r[6] * r[6] * r2);
        }
        const double Ap = S * kdn * inv_km;
        double s, c;
        sincos(kx * r[0] + ky * r[1] + kz * r[2], &s, &c);
        ar += Ap * c;
        ai += Ap * s;
      }
    }
  }
  if (part)
    part[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * Q + q] = make_double2(ar, ai);
  else
    J[(size_t)d.col * ldJ + q] = make_double2(ar, ai);
}

// J of a split launch: the nsplit partial sums of each (object, node) in
// split order.
__global__ void __launch_bounds__(256) k_ff_fp64_finish(const ObjDesc* __restrict__ objs, int Q,
                                                        int nsplit,
                                                        const double2* __restrict__ part,
                                                        double2* __restrict__ J, int ldJ) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  double ar = 0.0, ai = 0.0;
  for (int z = 0; z < nsplit; ++z) {
    const double2 v = part[((size_t)z * gridDim.y + blockIdx.y) * Q + q];
    ar += v.x;
    ai += v.y;
  }
  J[(size_t)objs[blockIdx.y].col * ldJ + q] = make_double2(ar, ai);
}

constexpr size_t kFp64SplitBytes = 8u << 20;

int ff_fp64_slots() {
  static std::atomic<int> cache[16];
  return resident_workgroups(k_ff_fp64, 256, cache);
}

}  // namespace

int ff_fp64_split(int N, int Q, int max_elems) {
  if (N <= 0 || Q <= 0) return 1;
  const int wgs = (Q + 255) / 256 * N;
  const size_t by_mem = kFp64SplitBytes / ((size_t)Q * N * sizeof(double2));
  const int cap = (int)std::max<size_t>(1, std::min<size_t>(by_mem, 64));
  return std::min(split_rounds(wgs, ff_fp64_slots(), max_split(max_elems, 64)), cap);
}

size_t ff_fp64_split_values(int N, int Q, int nsplit) {
  return nsplit > 1 ? (size_t)nsplit * N * Q : 0;
}

void launch_ff_fp64(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_kfull, int Q,
                    double2* d_J, int ldJ, int nsplit, double2* d_part) {
  if (N <= 0 || Q <= 0) return;
  if (nsplit > 1 && d_part) {
    hipLaunchKernelGGL(k_ff_fp64, dim3((Q + 255) / 256, N, nsplit), dim3(256), 0, s, d_objs,
                       d_kfull, Q, d_J, ldJ, d_part);
    hipLaunchKernelGGL(k_ff_fp64_finish, dim3((Q + 255) / 256, N), dim3(256), 0, s, d_objs, Q,
                       nsplit, d_part, d_J, ldJ);
    return;
  }
  hipLaunchKernelGGL(k_ff_fp64, dim3((Q + 255) / 256, N), dim3(256), 0, s, d_objs, d_kfull, Q,
                     d_J, ldJ, nullptr);
}

}  // namespace pii
