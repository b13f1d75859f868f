// Split-precision Gram on the gfx950 matrix cores.
//
// V = X^T X / (8 pi^3) with X (K x N, column-major, K = 2 Qh) the FP32
// weighted form factors sqrt(w) [Re A; Im A] written by the fast form-factor
// kernel.  Each FP32 entry is split into a high and a low 16-bit piece,
// x = hi + lo (+ O(2^-16 x) for BF16, O(2^-22 x) for FP16), and the product is
// assembled from three MFMA products hi.hi + hi.lo + lo.hi with FP32
// accumulation.  The FP32 accumulators are flushed into FP64 every 256 K
// values (compensates the long FP32 sums), split-K slices are summed in FP64
// by a deterministic second pass.
//
// FP16 mode scales every column by a power of two 2^e_i (chosen from the
// column's max |x| so that the max lands near 2^14) to keep the low pieces out
// of the subnormal range; the reduction divides by 2^(e_i + e_j) exactly.
//
// Tile: 64 x 64 outputs per workgroup (4 waves, 32 x 32 each as 2 x 2 MFMA
// 16x16x32 tiles), upper-triangular tile pairs only.
#include "device_common.hpp"
#include "pii_internal.hpp"

namespace pii {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;
constexpr int kKStep = 32;       // K per MFMA 16x16x32
constexpr int kFlush = 8;        // K steps between FP32 -> FP64 flushes (256 K values)

__device__ __forceinline__ float col_scale(const unsigned* colmax, int i, int N, int fp16) {
  if (!fp16 || i >= N) return 1.0f;
  const float m = __uint_as_float(colmax[i]);
  if (!(m > 0.0f)) return 1.0f;
  int e;
  frexpf(m, &e);                 // m = f * 2^e, f in [0.5, 1)
  return ldexpf(1.0f, 14 - e);   // max |x| * scale in [2^13, 2^14)
}

template <int FP16>
struct Split;

template <>
struct Split<1> {
  typedef half8 T;
  static __device__ __forceinline__ void run(const float (&x)[8], float s, T& hi, T& lo) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const float v = x[i] * s;
      const _Float16 h = (_Float16)v;
      hi[i] = h;
      lo[i] = (_Float16)(v - (float)h);
    }
  }
  static __device__ __forceinline__ f32x4 mma(T a, T b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

template <>
struct Split<0> {
  typedef bf16x8 T;
  static __device__ __forceinline__ void run(const float (&x)[8], float, T& hi, T& lo) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const __bf16 h = (__bf16)x[i];
      hi[i] = h;
      lo[i] = (__bf16)(x[i] - (float)h);
    }
  }
  static __device__ __forceinline__ f32x4 mma(T a, T b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

template <int FP16>
__global__ void __launch_bounds__(256) k_gram_split(const float* __restrict__ X, int N, int ldK,
                                                    const unsigned* __restrict__ colmax,
                                                    double* __restrict__ part, int kslice) {
  typedef Split<FP16> S;
  typedef typename S::T T;
  const int nt = (N + kTile - 1) / kTile;
  int t = blockIdx.x, ti = 0;
  while (t >= nt - ti) { t -= nt - ti; ++ti; }
  const int tj = ti + t;
  const int split = blockIdx.y;
  const int kbeg = split * kslice;
  const int kend = min(ldK, kbeg + kslice);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int row0 = ti * kTile + (wave >> 1) * 32;
  const int col0 = tj * kTile + (wave & 1) * 32;

  int ri[2], cj[2];
  float rs[2], cs[2];
  const float* pa[2];
  const float* pb[2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    ri[a] = row0 + a * 16 + l16;
    cj[a] = col0 + a * 16 + l16;
    rs[a] = col_scale(colmax, ri[a], N, FP16);
    cs[a] = col_scale(colmax, cj[a], N, FP16);
    pa[a] = X + (size_t)min(ri[a], N - 1) * ldK + kq * 8;
    pb[a] = X + (size_t)min(cj[a], N - 1) * ldK + kq * 8;
  }
  const bool rv[2] = {ri[0] < N, ri[1] < N};
  const bool cv[2] = {cj[0] < N, cj[1] < N};

  f32x4 acc[2][2];
  double acc64[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[a][b][r] = 0.0;
    }

  int step = 0;
  for (int k0 = kbeg; k0 < kend; k0 += kKStep) {
    T ahi[2], alo[2], bhi[2], blo[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      float xa[8], xb[8];
      const float4 a0 = *reinterpret_cast<const float4*>(pa[a] + k0);
      const float4 a1 = *reinterpret_cast<const float4*>(pa[a] + k0 + 4);
      const float4 b0 = *reinterpret_cast<const float4*>(pb[a] + k0);
      const float4 b1 = *reinterpret_cast<const float4*>(pb[a] + k0 + 4);
      xa[0] = rv[a] ? a0.x : 0.f; xa[1] = rv[a] ? a0.y : 0.f; xa[2] = rv[a] ? a0.z : 0.f;
      xa[3] = rv[a] ? a0.w : 0.f; xa[4] = rv[a] ? a1.x : 0.f; xa[5] = rv[a] ? a1.y : 0.f;
      xa[6] = rv[a] ? a1.z : 0.f; xa[7] = rv[a] ? a1.w : 0.f;
      xb[0] = cv[a] ? b0.x : 0.f; xb[1] = cv[a] ? b0.y : 0.f; xb[2] = cv[a] ? b0.z : 0.f;
      xb[3] = cv[a] ? b0.w : 0.f; xb[4] = cv[a] ? b1.x : 0.f; xb[5] = cv[a] ? b1.y : 0.f;
      xb[6] = cv[a] ? b1.z : 0.f; xb[7] = cv[a] ? b1.w : 0.f;
      S::run(xa, rs[a], ahi[a], alo[a]);
      S::run(xb, cs[a], bhi[a], blo[a]);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc[a][b] = S::mma(ahi[a], bhi[b], acc[a][b]);
        acc[a][b] = S::mma(ahi[a], blo[b], acc[a][b]);
        acc[a][b] = S::mma(alo[a], bhi[b], acc[a][b]);
      }
    if (++step == kFlush) {
      step = 0;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc64[a][b][r] += (double)acc[a][b][r];
          acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
    }
  }
  double* P = part + (size_t)split * N * N;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // MFMA 16x16 output layout: lane holds column l16, rows 4*kq + r.
        const int i = row0 + a * 16 + 4 * kq + r;
        const int j = col0 + b * 16 + l16;
        if (i < N && j < N) P[(size_t)i * N + j] = acc64[a][b][r] + (double)acc[a][b][r];
      }
}

__global__ void __launch_bounds__(256) k_gram_split_reduce(const double* __restrict__ part,
                                                           int nsplit, int N,
                                                           const unsigned* __restrict__ colmax,
                                                           int fp16, double scale,
                                                           double* __restrict__ V) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)N * N) return;
  const int i = (int)(idx / N), j = (int)(idx % N);
  if (i > j) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += part[(size_t)k * N * N + idx];
  const double si = (double)col_scale(colmax, i, N, fp16);
  const double sj = (double)col_scale(colmax, j, N, fp16);
  s *= scale / (si * sj);
  V[(size_t)i * N + j] = s;
  V[(size_t)j * N + i] = s;
}

}  // namespace

int gram_split_splits(int K, int N) {
  const int nt = (N + kTile - 1) / kTile;
  const int tiles = nt * (nt + 1) / 2;
  int s = (1024 + tiles - 1) / tiles;               // >= ~1024 workgroups
  s = std::max(1, std::min(s, K / (kKStep * kFlush * 2)));  // >= 512 K per slice
  return s;
}

void launch_gram_split(hipStream_t s, int fp16, const float* d_X, int N, int ldK,
                       const unsigned* d_colmax, double* d_part, int nsplit, double* d_V) {
  if (N <= 0) return;
  const int nt = (N + kTile - 1) / kTile;
  const int tiles = nt * (nt + 1) / 2;
  const int unit = kKStep * kFlush;
  int kslice = (ldK + nsplit - 1) / nsplit;
  kslice = (kslice + unit - 1) / unit * unit;
  if (fp16)
    hipLaunchKernelGGL(k_gram_split<1>, dim3(tiles, nsplit), dim3(256), 0, s, d_X, N, ldK,
                       d_colmax, d_part, kslice);
  else
    hipLaunchKernelGGL(k_gram_split<0>, dim3(tiles, nsplit), dim3(256), 0, s, d_X, N, ldK,
                       d_colmax, d_part, kslice);
  const size_t total = (size_t)N * N;
  hipLaunchKernelGGL(k_gram_split_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     d_part, nsplit, N, d_colmax, fp16, kInv8Pi3, d_V);
}

}  // namespace pii
