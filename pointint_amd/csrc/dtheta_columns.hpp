// Output stage shared by the affine derivative kernels (per element in
// dtheta.hip, moment series in dseries.hip).
#pragma once

#include "device_common.hpp"

namespace pii {

// The 12 columns X_t = sqrt(w) dA/dtheta_t of node q from the sums v = {A,
// dA/dA_rc (row-major)} (re, im interleaved): dA/dt_t = i k_t A.  Called by
// whole waves (the column maxima are wave reductions); lanes with active
// false store nothing and count as zero.
__device__ __forceinline__ void write_columns(const double (&v)[20], int q, int Qh, double kx,
                                              double ky, double kz, const double* krep, int col,
                                              float2* __restrict__ Xt, int ldX,
                                              unsigned* __restrict__ colmax, int lane,
                                              bool active = true) {
  const int qc = q < Qh ? q : Qh - 1;
  const double sw = (q < Qh) ? sqrt(krep[7 * (size_t)Qh + qc]) : 0.0;
  const double kv[3] = {kx, ky, kz};
  float2* X = Xt + (size_t)(12 * col) * ldX;
  for (int t = 0; t < 12; ++t) {
    double re, im;
    if (t < 3) {  // i k_t A
      re = -kv[t] * v[1];
      im = kv[t] * v[0];
    } else {
      re = v[2 + 2 * (t - 3)];
      im = v[3 + 2 * (t - 3)];
    }
    const float2 x = make_float2((float)(re * sw), (float)(im * sw));
    if (active && q < ldX) X[(size_t)t * ldX + q] = (q < Qh) ? x : make_float2(0.f, 0.f);
    float mval = (active && q < Qh) ? fmaxf(fabsf(x.x), fabsf(x.y)) : 0.0f;
    for (int off = 32; off > 0; off >>= 1) mval = fmaxf(mval, __shfl_xor(mval, off));
    if (lane == 0) atomicMax(colmax + 12 * col + t, __float_as_uint(mval));
  }
}

}  // namespace pii
