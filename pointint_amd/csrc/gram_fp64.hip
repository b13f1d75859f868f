// Plain FP64 Gram: V_ij = (1/8 pi^3) sum_q w_q Re(J_qi conj J_qj)
// (cpp/src/compute_intersection_volume_multi_object.cu:74-116 and
//  cpp/src/compute_intersection_volume_multi_object.cpp:98-149).
//
// J is object-major (column i is contiguous in q).  A workgroup computes one
// 32 x 32 tile of the upper triangle of V over one slice of q (split-K); the
// slices are summed in a second, deterministic pass that also mirrors V.
#include "device_common.hpp"
#include <algorithm>

#include "pii_internal.hpp"

namespace pii {

namespace {

constexpr int kT = 32;   // tile edge (objects)
constexpr int kQc = 32;  // q-chunk staged in LDS

// T = double2 (form factors, weights applied here) or float2 (fast-path X,
// sqrt(w) already folded in, w == nullptr).
template <class T>
__global__ void __launch_bounds__(256) k_gram_fp64(const T* __restrict__ J, int Q, int N,
                                                   int ldJ, const double* __restrict__ w,
                                                   double* __restrict__ part, int qslice) {
  // Upper-triangular tile pair from the linear index.
  const int nt = (N + kT - 1) / kT;
  int t = blockIdx.x, ti = 0;
  while (t >= nt - ti) { t -= nt - ti; ++ti; }
  const int tj = ti + t;
  const int split = blockIdx.y;
  const int q_beg = split * qslice;
  const int q_end = min(Q, q_beg + qslice);

  __shared__ double2 sa[kT][kQc + 1];
  __shared__ double2 sb[kT][kQc + 1];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};

  for (int q0 = q_beg; q0 < q_end; q0 += kQc) {
    // 32 x 32 complex per operand: 4 elements per thread each.
    for (int e = tid; e < kT * kQc; e += 256) {
      const int r = e / kQc, c = e % kQc;
      const int q = q0 + c;
      const int oi = ti * kT + r, oj = tj * kT + r;
      double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
      if (q < q_end) {
        const double wq = w ? w[q] : 1.0;
        if (oi < N) {
          const T v = J[(size_t)oi * ldJ + q];
          a = make_double2((double)v.x, (double)v.y);
        }
        if (oj < N) {
          const T v = J[(size_t)oj * ldJ + q];
          b = make_double2((double)v.x * wq, (double)v.y * wq);
        }
      }
      sa[r][c] = a;
      sb[r][c] = b;
    }
    __syncthreads();
#pragma unroll 4
    for (int c = 0; c < kQc; ++c) {
      const double2 a0 = sa[2 * ty][c], a1 = sa[2 * ty + 1][c];
      const double2 b0 = sb[2 * tx][c], b1 = sb[2 * tx + 1][c];
      acc[0][0] += a0.x * b0.x + a0.y * b0.y;
      acc[0][1] += a0.x * b1.x + a0.y * b1.y;
      acc[1][0] += a1.x * b0.x + a1.y * b0.y;
      acc[1][1] += a1.x * b1.x + a1.y * b1.y;
    }
    __syncthreads();
  }
  double* P = part + (size_t)split * N * N;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const int i = ti * kT + 2 * ty + a, j = tj * kT + 2 * tx + b;
      if (i < N && j < N) P[(size_t)i * N + j] = acc[a][b];
    }
}

__global__ void __launch_bounds__(256) k_gram_reduce(const double* __restrict__ part, int nsplit,
                                                     int N, double scale,
                                                     double* __restrict__ V) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)N * N) return;
  const int i = (int)(idx / N), j = (int)(idx % N);
  if (i > j) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += part[(size_t)k * N * N + idx];
  s *= scale;
  V[(size_t)i * N + j] = s;
  V[(size_t)j * N + i] = s;
}

}  // namespace

int gram_fp64_splits(int Q, int N) {
  const int nt = (N + kT - 1) / kT;
  const int tiles = nt * (nt + 1) / 2;
  int s = (2048 + tiles - 1) / tiles;        // aim for >= 2048 workgroups
  s = std::max(1, std::min(s, (Q + 1023) / 1024));  // but >= 1024 nodes per slice
  return s;
}

template <class T>
static void gram_fp64_impl(hipStream_t s, const T* d_J, int Q, int N, int ldJ, const double* d_w,
                           double* d_part, int nsplit, double* d_V) {
  if (N <= 0) return;
  const int nt = (N + kT - 1) / kT;
  const int tiles = nt * (nt + 1) / 2;
  int qslice = (Q + nsplit - 1) / nsplit;
  qslice = (qslice + kQc - 1) / kQc * kQc;
  hipLaunchKernelGGL(k_gram_fp64<T>, dim3(tiles, nsplit), dim3(256), 0, s, d_J, Q, N, ldJ, d_w,
                     d_part, qslice);
  const size_t total = (size_t)N * N;
  hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     d_part, nsplit, N, kInv8Pi3, d_V);
}

void launch_gram_fp64(hipStream_t s, const double2* d_J, int Q, int N, int ldJ,
                      const double* d_w, double* d_part, int nsplit, double* d_V) {
  gram_fp64_impl(s, d_J, Q, N, ldJ, d_w, d_part, nsplit, d_V);
}

void launch_gram_fp64_x(hipStream_t s, const float2* d_X, int Qh, int N, int ldX, double* d_part,
                        int nsplit, double* d_V) {
  gram_fp64_impl(s, d_X, Qh, N, ldX, (const double*)nullptr, d_part, nsplit, d_V);
}

}  // namespace pii
