// Shared pieces of the moment-series kernels (forward series in moments.hip,
// affine derivative series in dseries.hip): the triangle series coefficients,
// the cross-wave reduction of the per-order sums and the switch over the
// per-object series order.
#pragma once

#include "device_common.hpp"

namespace pii {
namespace ser {

constexpr int P = kSeriesOrder;
constexpr int kLanes = 64;
constexpr int kWaves = 4;

// c_n = (-1)^floor(n/2) 2/(n+2)!: the real factor of (i)^n 2/(n+2)! (the i
// of odd n is applied after the Horner sums).
struct SeriesCoef {
  double c[P + 1];
  constexpr SeriesCoef() : c() {
    double f = 1.0;  // 2/(n+2)!, from n = 0: 2/2! = 1
    for (int n = 0; n <= P; ++n) {
      const double sign = ((n / 2) % 2 == 0) ? 1.0 : -1.0;
      c[n] = sign * f;
      f = f / (double)(n + 3);
    }
  }
};
__device__ constexpr SeriesCoef kSeries{};

// Sums the four waves' H_n into wave 0's registers through one LDS buffer,
// in the fixed order ((H0 + H1) + H2) + H3.
// dst(n) = f(n) for n = 0..PP in chunks of 8 orders: the scheduling barriers
// keep the compiler from issuing a whole array of loads or stores at once
// into fresh registers.
template <int PP, class F>
__device__ __forceinline__ void for_chunked(F f) {
#pragma unroll
  for (int n0 = 0; n0 <= PP; n0 += 8) {
#pragma unroll
    for (int n = n0; n < n0 + 8 && n <= PP; ++n) f(n);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// red = (H1 + H2) + H3 in LDS, in that order, the waves' H registers only
// read (a conditional update of H would make the compiler keep two copies of
// the array).  Wave 0 then folds in its own H.
template <int PP>
__device__ __forceinline__ void sum_waves_123(const double (&H)[PP + 1],
                                              double (&red)[PP + 1][kLanes], int wave, int lane) {
  if (wave == 1) for_chunked<PP>([&](int n) { red[n][lane] = H[n]; });
  __syncthreads();
  if (wave == 2) for_chunked<PP>([&](int n) { red[n][lane] += H[n]; });
  __syncthreads();
  if (wave == 3) for_chunked<PP>([&](int n) { red[n][lane] += H[n]; });
  __syncthreads();
}

// o[n * 64] = H[n] for n = 0..PP (the raw sums of a split workgroup).
template <int PP>
__device__ __forceinline__ void store_split(const double (&H)[PP + 1],
                                            double (&red)[PP + 1][kLanes], int wave, int lane,
                                            double* o) {
  sum_waves_123<PP>(H, red, wave, lane);
  if (wave == 0) for_chunked<PP>([&](int n) { o[n * kLanes] = red[n][lane] + H[n]; });
}

// CALL(PP) with PP the series order of triangle object d (the same for the
// whole workgroup: a uniform branch).
#define PII_SERIES_SWITCH(CALL)                          \
  switch (d.nord) {                                      \
    case 32: CALL(32); break;                            \
    case 40: CALL(40); break;                            \
    case 44: CALL(44); break;                            \
    case 48: CALL(48); break;                            \
    case 52: CALL(52); break;                            \
    default: CALL(56); break;                            \
  }
static_assert(kNumSeriesOrders == 6 && kSeriesOrders[0] == 32 && kSeriesOrders[1] == 40 &&
                  kSeriesOrders[2] == 44 && kSeriesOrders[3] == 48 && kSeriesOrders[4] == 52 &&
                  kSeriesOrders[5] == 56,
              "the order switch lists kSeriesOrders");

}  // namespace ser
}  // namespace pii
