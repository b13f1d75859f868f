// Per-call pre-pass: apply each object's DoFs once, producing world-space
// vertices and the element records the form-factor kernels stream.  One
// batched launch covers every object of a call (flat index over the prefix
// sums of vertex / element counts).
#include "device_common.hpp"
#include "pii_internal.hpp"

namespace pii {

namespace {

__device__ __forceinline__ void world_vertex(const ObjDesc* __restrict__ objs,
                                             const int* __restrict__ vprefix, int N, int i) {
  const int o = find_owner(vprefix, N, i);
  const ObjDesc& d = objs[o];
  const int v = i - vprefix[o];
  double w0, w1, w2;
  if (d.dof_kind == PII_DOF_AFFINE) {
    // cpp/src/dof/affine_dof.cpp:83 -- x' = A x + t, A row-major in dofs[3:12]
    const double* p = d.dofs;
    const double x = d.verts[3 * v], y = d.verts[3 * v + 1], z = d.verts[3 * v + 2];
    w0 = p[3] * x + p[4] * y + p[5] * z + p[0];
    w1 = p[6] * x + p[7] * y + p[8] * z + p[1];
    w2 = p[9] * x + p[10] * y + p[11] * z + p[2];
  } else {
    // cpp/src/dof/triangle_mesh_dof.cpp:46-47 -- vertices are the DoFs
    const int dv = d.dofvid[v];
    if (dv >= 0) {
      w0 = d.dofs[3 * dv]; w1 = d.dofs[3 * dv + 1]; w2 = d.dofs[3 * dv + 2];
    } else {
      w0 = w1 = w2 = 0.0;  // degenerate: the face contributes exactly zero
    }
  }
  d.world[3 * v] = w0;
  d.world[3 * v + 1] = w1;
  d.world[3 * v + 2] = w2;
}

// Both kernels cover the flat index range [0, total); total < 0 takes it from
// the device prefix sums (prefix[N]), so that a launch of a fixed grid can be
// replayed for batches of other sizes (grid-stride loops).
__global__ void __launch_bounds__(256) k_world_vertices(const ObjDesc* __restrict__ objs,
                                                        const int* __restrict__ vprefix, int N,
                                                        int total_arg) {
  const int total = total_arg >= 0 ? total_arg : vprefix[N];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x)
    world_vertex(objs, vprefix, N, i);
}

// Element records.  fast != 0 writes the patch-ordered triangle layout and,
// for objects with a moment series, the moment records.
__device__ __forceinline__ void element_record(const ObjDesc* __restrict__ objs,
                                               const int* __restrict__ eprefix, int N, int i,
                                               int fast) {
  const int o = find_owner(eprefix, N, i);
  const ObjDesc& d = objs[o];
  const int e = i - eprefix[o];
  if (d.type == PII_GEOM_TRIANGLE) {
    const int f = fast ? (int)d.patch_faces[e] : e;
    const int i0 = d.faces[3 * f], i1 = d.faces[3 * f + 1], i2 = d.faces[3 * f + 2];
    const double* W = d.world;
    const double ax = W[3 * i0], ay = W[3 * i0 + 1], az = W[3 * i0 + 2];
    const double e1x = W[3 * i1] - ax, e1y = W[3 * i1 + 1] - ay, e1z = W[3 * i1 + 2] - az;
    const double e2x = W[3 * i2] - ax, e2y = W[3 * i2 + 1] - ay, e2z = W[3 * i2 + 2] - az;
    // S = (e1 x e2)/2: cpp/src/dof/affine_dof.cpp:86, triangle_mesh_dof.cpp:54
    const double sx = 0.5 * (e1y * e2z - e1z * e2y);
    const double sy = 0.5 * (e1z * e2x - e1x * e2z);
    const double sz = 0.5 * (e1x * e2y - e1y * e2x);
    if (!fast) {
      double* r = d.rec + (size_t)kRecPlainTri * e;
      r[0] = ax; r[1] = ay; r[2] = az;
      r[3] = e1x; r[4] = e1y; r[5] = e1z;
      r[6] = e2x; r[7] = e2y; r[8] = e2z;
      r[9] = sx; r[10] = sy; r[11] = sz;
    } else {
      double* r = d.rec + (size_t)kRecFastTri * e;
      r[0] = e1x; r[1] = e1y; r[2] = e1z;
      r[3] = e2x; r[4] = e2y; r[5] = e2z;
      float* rf = reinterpret_cast<float*>(r + 6);
      rf[0] = (float)(2.0 * sx);
      rf[1] = (float)(2.0 * sy);
      rf[2] = (float)(2.0 * sz);
      // the corners' slots in the patch's phasor table, 8 bits each, and
      // their offsets in a per-wave table of k_ff_fast
      const uint32_t loc = d.patch_floc[e] & 0xffffffu;
      uint32_t* ru = reinterpret_cast<uint32_t*>(rf);
      ru[3] = loc;
      ru[4] = (loc & 0xffu) * kWaveNodes;
      ru[5] = ((loc >> 8) & 0xffu) * kWaveNodes;
      ru[6] = ((loc >> 16) & 0xffu) * kWaveNodes;
      ru[7] = 0u;
      if (d.nser > 0 && d.mrec) {
        // corners relative to the series centre, and the area vector
        // (objects with cached harmonics have none)
        double* m = d.mrec + (size_t)kRecMoment * e;
        const double c0 = d.ctr[0], c1 = d.ctr[1], c2 = d.ctr[2];
        m[0] = ax - c0; m[1] = ay - c1; m[2] = az - c2;
        m[3] = W[3 * i1] - c0; m[4] = W[3 * i1 + 1] - c1; m[5] = W[3 * i1 + 2] - c2;
        m[6] = W[3 * i2] - c0; m[7] = W[3 * i2 + 1] - c1; m[8] = W[3 * i2 + 2] - c2;
        m[9] = sx; m[10] = sy; m[11] = sz;
      }
    }
  } else {
    // cpp/src/dof/affine_dof.cpp:115-118: c' = A c + t; k'.n = (A^T k).n = k.(A n)
    const double* p = d.dofs;
    const double* s = d.el + (size_t)kRecSurfel * e;
    const double cx = s[0], cy = s[1], cz = s[2], nx = s[3], ny = s[4], nz = s[5];
    double* r = d.rec + (size_t)kRecSurfel * e;
    r[0] = p[3] * cx + p[4] * cy + p[5] * cz + p[0];
    r[1] = p[6] * cx + p[7] * cy + p[8] * cz + p[1];
    r[2] = p[9] * cx + p[10] * cy + p[11] * cz + p[2];
    r[3] = p[3] * nx + p[4] * ny + p[5] * nz;
    r[4] = p[6] * nx + p[7] * ny + p[8] * nz;
    r[5] = p[9] * nx + p[10] * ny + p[11] * nz;
    r[6] = s[6];
    r[7] = s[7];
  }
}

__global__ void __launch_bounds__(256) k_records(const ObjDesc* __restrict__ objs,
                                                 const int* __restrict__ eprefix, int N,
                                                 int total_arg, int fast) {
  const int total = total_arg >= 0 ? total_arg : eprefix[N];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x)
    element_record(objs, eprefix, N, i, fast);
}

__global__ void __launch_bounds__(256) k_clear_u32(unsigned* __restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

__global__ void k_stamp(unsigned long long* __restrict__ out, int i) {
  if (threadIdx.x == 0) out[i] = __builtin_amdgcn_s_memrealtime();
}

// Centre and squared radius of a vertex-DoF object's DoF vertices (those a
// geometry vertex refers to), as the host plan computes them from host DoFs:
// out[5 j ..] = {c (bounding-box centre), max |x - c|^2, valid}.  One
// workgroup per object; min / max / max are exact, so the result does not
// depend on the reduction order.
__global__ void __launch_bounds__(256) k_dof_bounds(const DofBoundsJob* __restrict__ jobs,
                                                    double* __restrict__ out) {
  __shared__ double red[6][256];
  const DofBoundsJob jb = jobs[blockIdx.x];
  const int tid = threadIdx.x;
  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int v = tid; v < jb.nv; v += blockDim.x) {
    const int dv = jb.dofvid[v];
    if (dv < 0) continue;
    for (int k = 0; k < 3; ++k) {
      const double x = jb.dofs[3 * (size_t)dv + k];
      lo[k] = fmin(lo[k], x);
      hi[k] = fmax(hi[k], x);
    }
  }
  for (int k = 0; k < 3; ++k) {
    red[k][tid] = lo[k];
    red[3 + k][tid] = hi[k];
  }
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (tid < h)
      for (int k = 0; k < 3; ++k) {
        red[k][tid] = fmin(red[k][tid], red[k][tid + h]);
        red[3 + k][tid] = fmax(red[3 + k][tid], red[3 + k][tid + h]);
      }
    __syncthreads();
  }
  const bool valid = red[0][0] <= red[3][0];
  double c[3];
  for (int k = 0; k < 3; ++k) c[k] = valid ? 0.5 * (red[k][0] + red[3 + k][0]) : 0.0;
  __syncthreads();
  double r2 = 0.0;
  for (int v = tid; v < jb.nv; v += blockDim.x) {
    const int dv = jb.dofvid[v];
    if (dv < 0) continue;
    const double x = jb.dofs[3 * (size_t)dv] - c[0], y = jb.dofs[3 * (size_t)dv + 1] - c[1],
                 z = jb.dofs[3 * (size_t)dv + 2] - c[2];
    r2 = fmax(r2, fma(z, z, fma(y, y, x * x)));
  }
  red[0][tid] = r2;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (tid < h) red[0][tid] = fmax(red[0][tid], red[0][tid + h]);
    __syncthreads();
  }
  if (tid == 0) {
    double* o = out + 5 * (size_t)blockIdx.x;
    o[0] = c[0];
    o[1] = c[1];
    o[2] = c[2];
    o[3] = valid ? red[0][0] : 0.0;
    o[4] = valid ? 1.0 : 0.0;
  }
}

}  // namespace

void launch_dof_bounds(hipStream_t s, const DofBoundsJob* d_jobs, int n, double* d_out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_dof_bounds, dim3(n), dim3(256), 0, s, d_jobs, d_out);
}

void launch_prepass_fixed(hipStream_t s, const ObjDesc* d_objs, const int* d_vprefix,
                          const int* d_eprefix, int N, int fast, int grid) {
  hipLaunchKernelGGL(k_world_vertices, dim3(grid), dim3(256), 0, s, d_objs, d_vprefix, N, -1);
  hipLaunchKernelGGL(k_records, dim3(grid), dim3(256), 0, s, d_objs, d_eprefix, N, -1, fast);
}

void launch_clear_u32(hipStream_t s, unsigned* d_p, int n) {
  if (n <= 0) return;
  const unsigned grid = (unsigned)std::min(1024, (n + 255) / 256);
  hipLaunchKernelGGL(k_clear_u32, dim3(grid), dim3(256), 0, s, d_p, n);
}

void launch_stamp(hipStream_t s, unsigned long long* d_out, int i) {
  hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, d_out, i);
}

void launch_prepass(hipStream_t s, const ObjDesc* d_objs, const int* d_vprefix, int total_v,
                    const int* d_eprefix, int total_e, int N, int fast) {
  if (total_v > 0) {
    hipLaunchKernelGGL(k_world_vertices, dim3((total_v + 255) / 256), dim3(256), 0, s, d_objs,
                       d_vprefix, N, total_v);
  }
  if (total_e > 0) {
    hipLaunchKernelGGL(k_records, dim3((total_e + 255) / 256), dim3(256), 0, s, d_objs,
                       d_eprefix, N, total_e, fast);
  }
}

}  // namespace pii
