// Split-precision Gram on the gfx950 matrix cores.
//
// V = X^T X / (8 pi^3) with X (K x ncols, column-major, K = 2 ldX FP32
// values per column: Re, Im of sqrt(w) A at each representative node) written
// by the fast form-factor kernels.  Two passes:
//
// 1. k_split_planes: every FP32 value is split once into NP 16-bit pieces,
//    x = p0 + p1 (+ p2), written as NP planes in the operand order of
//    v_mfma_f32_16x16x32_{bf16,f16}: for each group of 16 columns and each
//    step of 32 K values, 64 lanes x 8 values (lane l: column l % 16, K values
//    8 (l / 16) .. + 7), 1 KiB per plane, so a wave loads one operand as one
//    contiguous 1 KiB block.
//      BF16 (PII_GRAM_BF16X3): three pieces (8 + 8 + 8 significant bits, the
//        full FP32 significand; BF16 has the FP32 exponent range) and six
//        products p0 p0' + p0 p1' + p1 p0' + p0 p2' + p2 p0' + p1 p1'; the
//        dropped terms are below 2^-24 of |x x'|.
//      FP16 (PII_GRAM_FP16X3): two pieces (11 + 11 bits) of the column scaled
//        by 2^e (max |x| 2^e in [2^13, 2^14), keeping the low pieces out of
//        the FP16 subnormals) and three products p0 p0' + p0 p1' + p1 p0'.
// 2. k_gram_planes: 128 x 128 output tiles of the upper triangle per
//    workgroup (8 waves of 32 x 64), split over K; the planes of a K step are
//    staged in LDS (double-buffered) once per workgroup and read by every wave
//    that needs them.  p0 p0' accumulates in one FP32 MFMA accumulator, the
//    correction products in a second; the first is added into FP64 registers
//    every 2 K steps (64 values), the second every 8, so that FP32 rounding
//    of long sums stays below the rounding of X itself.  The FP64 partials of
//    the K slices are summed in slice order by k_gram_planes_reduce
//    (deterministic).
//
// One launch serves a list of Grams ("queries"): query q owns columns
// [c0, c0 + n) of the planes (c0 a multiple of 16) and its own slices, so
// its result does not depend on the other queries of the launch.  The work
// items (tile x slice of every query) and V entries are read from a device
// plan; both kernels loop over them from a fixed grid, so a captured launch
// can be replayed for another plan of the same capacity.
#include "device_common.hpp"
#include "pii_internal.hpp"

namespace pii {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTileCols = 128;             // output tile side
constexpr int kTileGroups = kTileCols / 16;
constexpr int kWaves = 8;                  // 4 x 2 waves of 32 x 64 outputs
constexpr int kFlushMain = 2;              // K steps between FP64 flushes of p0 p0'
constexpr int kFlushCorr = 8;              // ... of the correction products

__device__ __forceinline__ float col_scale(const unsigned* colmax, int c, int fp16) {
  if (!fp16) return 1.0f;
  const float m = __uint_as_float(colmax[c]);
  if (!(m > 0.0f) || !(m < 3.0e38f)) return 1.0f;
  int e;
  frexpf(m, &e);                 // m = f * 2^e, f in [0.5, 1)
  return ldexpf(1.0f, 14 - e);   // max |x| * scale in [2^13, 2^14)
}

template <int FP16>
struct Piece;

template <>
struct Piece<0> {
  typedef bf16x8 T;
  static __device__ __forceinline__ void split(float x, __bf16* out, int np) {
    float r = x;
    for (int p = 0; p < np; ++p) {
      const __bf16 h = (__bf16)r;
      out[p] = h;
      r = r - (float)h;      // exact: the residual fits the FP32 significand
    }
  }
  static __device__ __forceinline__ f32x4 mma(T a, T b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

template <>
struct Piece<1> {
  typedef half8 T;
  static __device__ __forceinline__ void split(float x, _Float16* out, int np) {
    float r = x;
    for (int p = 0; p < np; ++p) {
      const _Float16 h = (_Float16)r;
      out[p] = h;
      r = r - (float)h;
    }
  }
  static __device__ __forceinline__ f32x4 mma(T a, T b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

template <int FP16> constexpr int pieces() { return FP16 ? 2 : 3; }

// Planes: np x G groups x S steps x 64 lanes x 16 bytes.  One thread per
// (group, step, lane) chunk in output order (coalesced 16-byte stores).
// Columns at or beyond the plan's ncols (when a plan is given) are not
// converted: the planes are laid out for G groups, the launch covers the
// columns the plan's queries use.
template <int FP16>
__global__ void __launch_bounds__(256) k_split_planes(const float* __restrict__ X, int ldK,
                                                      int ncols, int G, int S,
                                                      const unsigned* __restrict__ colmax,
                                                      const GramPlanHeader* __restrict__ plan,
                                                      uint4* __restrict__ planes) {
  constexpr int NP = pieces<FP16>();
  typedef typename std::conditional<FP16 != 0, _Float16, __bf16>::type E;
  const size_t nchunk = (size_t)G * S * 64;
  if (plan) ncols = min(ncols, plan->ncols);
  const size_t nconv = (size_t)((ncols + 15) / 16) * S * 64;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nconv;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int lane = (int)(idx & 63);
    const size_t gs = idx >> 6;
    const int s = (int)(gs % S);
    const int g = (int)(gs / S);
    const int c = g * 16 + (lane & 15);
    const int k = s * 32 + 8 * (lane >> 4);
    float x[8];
    if (c < ncols) {
      const float4 a = *reinterpret_cast<const float4*>(X + (size_t)c * ldK + k);
      const float4 b = *reinterpret_cast<const float4*>(X + (size_t)c * ldK + k + 4);
      const float sc = col_scale(colmax, c, FP16);
      x[0] = a.x * sc; x[1] = a.y * sc; x[2] = a.z * sc; x[3] = a.w * sc;
      x[4] = b.x * sc; x[5] = b.y * sc; x[6] = b.z * sc; x[7] = b.w * sc;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) x[e] = 0.0f;
    }
    E pc[NP][8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      E t[NP];
      Piece<FP16>::split(x[e], t, NP);
#pragma unroll
      for (int p = 0; p < NP; ++p) pc[p][e] = t[p];
    }
#pragma unroll
    for (int p = 0; p < NP; ++p)
      planes[(size_t)p * nchunk + idx] = *reinterpret_cast<const uint4*>(pc[p]);
  }
}

__device__ __forceinline__ void decode_tile(int t, int nt, int& ti, int& tj) {
  ti = 0;
  while (t >= nt - ti) {
    t -= nt - ti;
    ++ti;
  }
  tj = ti + t;
}

__device__ __forceinline__ int find_query(const GramQuery* q, int nq, long long key,
                                          long long GramQuery::*field) {
  int lo = 0, hi = nq - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (q[mid].*field <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int FP16>
__global__ void __launch_bounds__(kWaves * 64) k_gram_planes(const uint4* __restrict__ planes,
                                                             int G, int S,
                                                             const GramPlanHeader* __restrict__ plan,
                                                             double* __restrict__ part) {
  constexpr int NP = pieces<FP16>();
  typedef Piece<FP16> Pc;
  typedef typename Pc::T T;
  // [buffer][side][group][piece][lane]
  __shared__ uint4 lds[2][2][kTileGroups][NP][64];
  const GramQuery* qs = reinterpret_cast<const GramQuery*>(plan + 1);
  const int nq = plan->nq;
  const long long nitems = plan->nitems;
  const size_t nchunk = (size_t)G * S * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;   // rows 32 wr.., cols 64 wc..

  for (long long item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int qi = find_query(qs, nq, item, &GramQuery::item0);
    const GramQuery q = qs[qi];
    // item = 8 (octet * ntiles + tile) + slice % 8 (see gram_plan_layout);
    // slices past the query's last are empty items
    const int local = (int)(item - q.item0);
    const int ot = local >> 3;
    const int t = ot % q.ntiles, slice = 8 * (ot / q.ntiles) + (local & 7);
    if (slice >= q.nslice) continue;   // uniform over the workgroup
    const int nt = (q.n + kTileCols - 1) / kTileCols;
    int ti, tj;
    decode_tile(t, nt, ti, tj);
    const bool diag = ti == tj;
    const int s0 = slice * q.slice_len;
    const int s1 = min(S, s0 + q.slice_len);
    const int r0 = ti * kTileCols, cc0 = tj * kTileCols;   // query-local
    const int gA = (q.c0 + r0) >> 4, gB = (q.c0 + cc0) >> 4;
    // this wave's block intersects the query (and, on a diagonal tile, the
    // upper triangle)
    const bool active = (r0 + 32 * wr < q.n) && (cc0 + 64 * wc < q.n) &&
                        !(diag && 32 * wr >= 64 * wc + 64);

    // chunks of one K step this thread stages: (side, group, piece, lane)
    constexpr int kChunks = 2 * kTileGroups * NP * 64;
    constexpr int kPerThread = kChunks / (kWaves * 64);
    uint4 reg[kPerThread];
    auto fetch = [&](int s) {
#pragma unroll
      for (int r = 0; r < kPerThread; ++r) {
        const int ch = tid + r * kWaves * 64;
        const int ln = ch & 63;
        const int p = (ch >> 6) % NP;
        const int g = (ch >> 6) / NP % kTileGroups;
        const int side = (ch >> 6) / NP / kTileGroups;
        const int col = (side ? cc0 : r0) + 16 * g + (ln & 15);   // query-local
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (col < q.n && !(diag && side)) {
          const int gg = (side ? gB : gA) + g;
          v = planes[(size_t)p * nchunk + ((size_t)gg * S + s) * 64 + ln];
        }
        reg[r] = v;
      }
    };
    auto stage = [&](int buf) {
#pragma unroll
      for (int r = 0; r < kPerThread; ++r) {
        const int ch = tid + r * kWaves * 64;
        (&lds[buf][0][0][0][0])[ch] = reg[r];
      }
    };

    f32x4 accm[2][4], accc[2][4];
    double acc64[2][4][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        accm[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        accc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[a][b][r] = 0.0;
      }
    const int sideB = diag ? 0 : 1;

    if (s0 < s1) {
      fetch(s0);
      stage(0);
    }
    __syncthreads();
    int nm = 0, nc = 0;
    for (int s = s0; s < s1; ++s) {
      const int buf = (s - s0) & 1;
      const bool more = s + 1 < s1;
      if (more) fetch(s + 1);
      if (active) {
        T A[2][NP];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            const uint4 v = lds[buf][0][2 * wr + a][p][lane];
            A[a][p] = *reinterpret_cast<const T*>(&v);
          }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          T B[NP];
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            const uint4 v = lds[buf][sideB][4 * wc + b][p][lane];
            B[p] = *reinterpret_cast<const T*>(&v);
          }
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            accm[a][b] = Pc::mma(A[a][0], B[0], accm[a][b]);
            accc[a][b] = Pc::mma(A[a][0], B[1], accc[a][b]);
            accc[a][b] = Pc::mma(A[a][1], B[0], accc[a][b]);
            if (NP == 3) {
              accc[a][b] = Pc::mma(A[a][0], B[NP - 1], accc[a][b]);
              accc[a][b] = Pc::mma(A[a][NP - 1], B[0], accc[a][b]);
              accc[a][b] = Pc::mma(A[a][1], B[1], accc[a][b]);
            }
          }
        }
        if (++nm == kFlushMain || !more) {
          nm = 0;
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
#pragma unroll
              for (int r = 0; r < 4; ++r) acc64[a][b][r] += (double)accm[a][b][r];
              accm[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            }
        }
        if (++nc == kFlushCorr || !more) {
          nc = 0;
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
#pragma unroll
              for (int r = 0; r < 4; ++r) acc64[a][b][r] += (double)accc[a][b][r];
              accc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            }
        }
      }
      if (more) stage(buf ^ 1);
      __syncthreads();
    }
    if (active) {
      double* P = part + q.part + (size_t)slice * q.n * q.n;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            // MFMA 16x16 output layout: lane holds column lane % 16, rows 4 (lane / 16) + r
            const int i = r0 + 32 * wr + 16 * a + 4 * (lane >> 4) + r;
            const int j = cc0 + 64 * wc + 16 * b + (lane & 15);
            if (i < q.n && j < q.n && i <= j) P[(size_t)i * q.n + j] = acc64[a][b][r];
          }
    }
  }
}

// V of every query: the upper triangle summed over the slices in order,
// scaled and mirrored.
__global__ void __launch_bounds__(256) k_gram_planes_reduce(const double* __restrict__ part,
                                                            const GramPlanHeader* __restrict__ plan,
                                                            const unsigned* __restrict__ colmax,
                                                            int fp16, double scale,
                                                            double* __restrict__ V) {
  const GramQuery* qs = reinterpret_cast<const GramQuery*>(plan + 1);
  const int nq = plan->nq;
  const long long total = plan->nentries;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int qi = find_query(qs, nq, e, &GramQuery::entry0);
    const GramQuery q = qs[qi];
    const long long loc = e - q.entry0;
    const int i = (int)(loc / q.n), j = (int)(loc % q.n);
    if (i > j) continue;
    const double* P = part + q.part + (size_t)i * q.n + j;
    const size_t stride = (size_t)q.n * q.n;
    double s = 0.0;
    for (int k = 0; k < q.nslice; ++k) s += P[k * stride];
    const double si = (double)col_scale(colmax, q.c0 + i, fp16);
    const double sj = (double)col_scale(colmax, q.c0 + j, fp16);
    s *= scale / (si * sj);
    V[q.v + (size_t)i * q.n + j] = s;
    V[q.v + (size_t)j * q.n + i] = s;
  }
}

}  // namespace

int gram_planes_pieces(int fp16) { return fp16 ? 2 : 3; }

size_t gram_plane_bytes(int fp16, int K, int ncols) {
  const size_t G = (size_t)(ncols + 15) / 16, S = (size_t)K / 32;
  return (size_t)gram_planes_pieces(fp16) * G * S * 64 * 16;
}

GramQuery gram_query(int c0, int n, int K) {
  GramQuery q{};
  q.c0 = c0;
  q.n = n;
  const int S = K / 32;
  const int nt = (n + kTileCols - 1) / kTileCols;
  q.ntiles = nt * (nt + 1) / 2;
  // enough slices for about 512 workgroups, at least 8 K steps each, and the
  // partials' traffic (8 n^2 bytes per slice) at most about half of the
  // planes' (2 NP K n bytes)
  long long ns = (512 + q.ntiles - 1) / q.ntiles;
  ns = std::min<long long>(ns, std::max(1, S / 8));
  ns = std::min<long long>(ns, std::max<long long>(1, (3LL * K) / (8LL * n)));
  ns = std::max<long long>(ns, 1);
  if (ns > 8) ns = (ns + 7) / 8 * 8;   // whole groups of eight slices (gram_plan_layout)
  q.slice_len = (int)((S + ns - 1) / ns);
  q.nslice = (S + q.slice_len - 1) / q.slice_len;
  return q;
}

void gram_plan_layout(std::vector<GramQuery>& qs, GramPlanHeader& h) {
  long long item = 0, part = 0, entry = 0, v = 0;
  int ncols = 0;
  for (GramQuery& q : qs) {
    ncols = std::max(ncols, q.c0 + q.n);
    q.item0 = item;
    q.part = part;
    q.entry0 = entry;
    q.v = v;
    // Items in groups of eight slices: item0 + 8 (octet * ntiles + tile) +
    // slice % 8.  Workgroups go to the eight XCDs round-robin by linear id
    // and the grids are multiples of 8, so every item of XCD x has a slice
    // = x (mod 8): each XCD reads only the K ranges of its slices, for all
    // tiles at once, and shares them through its L2 (eight times fewer HBM
    // reads of the planes than with the tiles' slices spread over all XCDs).
    item += 8LL * ((q.nslice + 7) / 8) * q.ntiles;
    part += (long long)q.nslice * q.n * q.n;
    entry += (long long)q.n * q.n;
    v += (long long)q.n * q.n;
  }
  h.nq = (int)qs.size();
  h.ncols = ncols;
  h.nitems = item;
  h.nentries = entry;
  h.part_doubles = part;
}

int gram_planes_grid(long long nitems, int cap) {
  // a multiple of 8 (nitems is; cap is rounded down), so that a workgroup's
  // items all fall on its XCD's slices
  const long long c = std::max(8, cap - cap % 8);
  return (int)std::max<long long>(8, std::min<long long>(nitems, c));
}

void launch_split_planes(hipStream_t s, int fp16, const float* d_X, int ldK, int ncols,
                         const unsigned* d_colmax, void* d_planes, const GramPlanHeader* d_plan) {
  const int G = (ncols + 15) / 16, S = ldK / 32;
  const size_t nchunk = (size_t)G * S * 64;
  const unsigned grid = (unsigned)std::min<size_t>((nchunk + 255) / 256, 16384);
  if (fp16)
    hipLaunchKernelGGL(k_split_planes<1>, dim3(grid), dim3(256), 0, s, d_X, ldK, ncols, G, S,
                       d_colmax, d_plan, static_cast<uint4*>(d_planes));
  else
    hipLaunchKernelGGL(k_split_planes<0>, dim3(grid), dim3(256), 0, s, d_X, ldK, ncols, G, S,
                       d_colmax, d_plan, static_cast<uint4*>(d_planes));
}

void launch_gram_planes(hipStream_t s, int fp16, const void* d_planes, int ldK, int ncols,
                        const GramPlanHeader* d_plan, int grid_items, long long nentries_cap,
                        double* d_part, const unsigned* d_colmax, double* d_V) {
  const int G = (ncols + 15) / 16, S = ldK / 32;
  if (fp16)
    hipLaunchKernelGGL(k_gram_planes<1>, dim3(grid_items), dim3(kWaves * 64), 0, s,
                       static_cast<const uint4*>(d_planes), G, S, d_plan, d_part);
  else
    hipLaunchKernelGGL(k_gram_planes<0>, dim3(grid_items), dim3(kWaves * 64), 0, s,
                       static_cast<const uint4*>(d_planes), G, S, d_plan, d_part);
  const unsigned rgrid =
      (unsigned)std::max<long long>(1, std::min<long long>((nentries_cap + 255) / 256, 8192));
  hipLaunchKernelGGL(k_gram_planes_reduce, dim3(rgrid), dim3(256), 0, s, d_part, d_plan,
                     d_colmax, fp16, kInv8Pi3, d_V);
}

}  // namespace pii
