// Issue-rate probes for the roofline figures of the measurement harness:
// sustained FP64 and FP32 FMA throughput and the FP32 sine rate of the
// device, measured with the same launch shape as the form-factor kernels
// (4 waves per workgroup, many workgroups per CU).  No reference counterpart.
#include "device_common.hpp"
#include "pii_internal.hpp"

namespace pii {

namespace {

constexpr int kChains = 8;  // independent accumulators per lane

// kChains independent FMA chains per lane, iters rounds: 2 * kChains * iters
// flops per lane.  The result is written only when it is never true at run
// time (guard), so the chains are not optimised away.  The multiplier and
// addend are moved to vector registers first: a three-operand FMA that
// reads a scalar register issues at about 60 % of the rate on gfx950
// (measured), which is not the form the kernels use.
template <class T>
__global__ void __launch_bounds__(256) k_fma_probe(int iters, T a, T b, T* __restrict__ sink,
                                                   int guard) {
  T x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = (T)(threadIdx.x + c);
  T av = a, bv = b;
  asm volatile("" : "+v"(av), "+v"(bv));  // no instruction: places them in VGPRs
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = fma(x[c], av, bv);
  }
  T s = 0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  if (guard == (int)threadIdx.x + 1000000) sink[blockIdx.x] = s;
}

// kChains independent hardware-sine chains per lane: kChains * iters sines.
__global__ void __launch_bounds__(256) k_sin_probe(int iters, float* __restrict__ sink,
                                                   int guard) {
  float x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = 0.001f * (float)(threadIdx.x + c);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = __builtin_amdgcn_sinf(x[c]);
  }
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  if (guard == (int)threadIdx.x + 1000000) sink[blockIdx.x] = s;
}

}  // namespace

double probe_issue_rate(hipStream_t s, int kind, int blocks, int iters, void* d_sink) {
  hipEvent_t e0, e1;
  PII_HIP(hipEventCreate(&e0));
  PII_HIP(hipEventCreate(&e1));
  const dim3 grid(blocks), block(256);
  auto launch = [&] {
    if (kind == 0)
      hipLaunchKernelGGL(k_fma_probe<double>, grid, block, 0, s, iters, 1.0000001, 1e-9,
                         static_cast<double*>(d_sink), 0);
    else if (kind == 1)
      hipLaunchKernelGGL(k_fma_probe<float>, grid, block, 0, s, iters, 1.0000001f, 1e-9f,
                         static_cast<float*>(d_sink), 0);
    else
      hipLaunchKernelGGL(k_sin_probe, grid, block, 0, s, iters, static_cast<float*>(d_sink), 0);
  };
  launch();  // warm-up
  PII_HIP(hipEventRecord(e0, s));
  launch();
  PII_HIP(hipEventRecord(e1, s));
  PII_HIP(hipEventSynchronize(e1));
  PII_HIP(hipGetLastError());
  float ms = 0.f;
  PII_HIP(hipEventElapsedTime(&ms, e0, e1));
  PII_HIP(hipEventDestroy(e0));
  PII_HIP(hipEventDestroy(e1));
  const double lanes = (double)blocks * 256.0;
  const double ops = lanes * (double)iters * kChains * (kind == 2 ? 1.0 : 2.0);
  return ops / (ms * 1e-3);
}

}  // namespace pii
