// Device-side types and math shared by the form-factor, Gram and adjoint
// kernels.  Everything here is written for gfx950 (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pii {

// Per-object descriptor of one batched launch.  Built on the host for every
// call; device pointers refer to resident geometry and per-call workspaces.
struct ObjDesc {
  int type;            // PII_GEOM_*
  int n;               // faces / disks / Gaussians
  int nv;              // geometry vertices (triangles)
  int npatch;          // triangle patches
  int dof_kind;        // PII_DOF_*
  int col;             // output column
  // Low-|k| moment series (triangles on a shell-structured k-grid): the first
  // nser shells come from the series (of order nord for triangles),
  // representative nodes q >= qbeg = nser*Mh from the per-face kernel.
  // dser: the affine derivative columns of those shells also come from a
  // series (affine triangle objects with a well-conditioned A).
  int nser, qbeg, nord, dser;
  double ctr[3];       // expansion centre of the series (world frame)
  double rctr[3];      // the same centre in the reference frame (ctr = A rctr + t)
  double kser;         // largest |k| of the object's series shells
  const double* verts;        // nv*3 reference-frame vertices
  const int32_t* faces;       // n*3
  const int32_t* dofvid;      // nv
  const int32_t* patch_hdr;   // npatch*4 {vbeg, vcnt, fbeg, fcnt}
  const int32_t* patch_verts; // geometry vertex id per patch slot
  const uint32_t* patch_faces;// face id per patch-ordered face
  const uint32_t* patch_floc; // 3 x 8-bit local corner slots per patch-ordered face
  const double* el;           // n*8 disks / Gaussians
  const double* dofs;         // device DoFs (12 affine or 3V mesh)
  double* world;              // nv*3 world vertices (written by the pre-pass)
  double* rec;                // element records (written by the pre-pass)
  double* mrec;               // n*12 moment records {v0-c, v1-c, v2-c, S} (nser > 0)
  // Spherical-harmonic form of the triangle series (sh_common.hpp): the
  // geometry's cached coefficients (null: the direct series over mrec) and
  // the scale s of the object's similarity transform A = s R.
  const double* shc;
  double shs;
};


// Element record layouts written by the pre-pass.
//   plain triangle : 12 doubles {a, e1, e2, S}            (world frame)
//   fast triangle  : 10 doubles {e1, e2, (2S.xyz f32, loc u32), (o0, o1, o2, 0) u32}
//                    in patch order; loc holds the corners' slots in the patch's
//                    vertex table (8 bits each), o_j = slot_j * kWaveNodes the
//                    offsets (in table entries) of those slots in a per-wave
//                    table of kWaveNodes nodes (k_ff_fast)
//   disk/Gaussian  :  8 doubles {c' = A c + t, A n, rho|sigma, area|w}
constexpr int kRecPlainTri = 12;
constexpr int kRecFastTri = 10;
constexpr int kWaveNodes = 16;  // nodes per wave in k_ff_fast
constexpr int kRecSurfel = 8;
constexpr int kRecMoment = 12;

// Moment series of the triangle form factor (see moments.hip): order P and the
// largest |k| R it is used for, R the object's radius about the centre.  With
// P = 56 and |k| R <= 15 the tail is below 3.6e-10 of the total area.
constexpr int kSeriesOrder = 56;
constexpr double kSeriesReach = 15.0;
// Triangle objects take the smallest of these orders whose tail bound at the
// object's own largest |k| R is below 3.6e-10 of the area (kernels exist for
// each).
constexpr int kSeriesOrders[] = {32, 40, 44, 48, 52, 56};
constexpr int kNumSeriesOrders = 6;
// Per (object, direction) record of the series adjoint: Gamma_0..P, then
// (m+3) Gamma_{m+1} for m < P; padded to a multiple of 8 doubles.
constexpr int kGammaStride = (2 * kSeriesOrder + 1 + 7) / 8 * 8;
static_assert(kGammaStride >= 2 * kSeriesOrder + 1, "Gamma record too short");

// Read-only data that a kernel indexes with wave-uniform indices (element
// records, world vertices, patch tables) is read through the constant address
// space so that the loads become scalar (SMEM) loads.  Only valid for memory
// the kernel itself does not write.
#define PII_CONST_AS __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const PII_CONST_AS T* as_const(const T* p) {
  return (const PII_CONST_AS T*)(p);
}
// The same data indexed per lane: global (vector memory) loads.
#define PII_GLOBAL_AS __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const PII_GLOBAL_AS T* as_global(const T* p) {
  return (const PII_GLOBAL_AS T*)(p);
}

// Workgroup -> (block, object) for a grid of gridDim.x blocks per object and
// gridDim.y objects that keeps all blocks of one object on one XCD, so that
// the object's element records stay in that XCD's L2.  The command processor
// hands workgroups to the 8 XCDs round-robin by linear id (x fastest), so
// workgroup L runs on XCD L % 8; its (L / 8)-th workgroup there gets object
// 8 (r / nb) + xcd and block r % nb.  Objects past the last full group of 8
// keep the plain mapping.  A bijection either way; only locality depends on
// the dispatch order.
struct BlockObj {
  int blk, obj;
};
__device__ __forceinline__ BlockObj xcd_block_map() {
  const int nb = gridDim.x, N = gridDim.y;
  const int L = blockIdx.y * nb + blockIdx.x;
  const int N8 = N & ~7;
  if (L < N8 * nb) {
    const int r = L >> 3;
    return {r % nb, (r / nb) * 8 + (L & 7)};
  }
  return {(int)blockIdx.x, (int)blockIdx.y};
}

struct cplx {
  double x, y;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// ---------------------------------------------------------------------------
// Reference element functions in FP64 (cpp/src/element_functions.cpp:17-93).
// ---------------------------------------------------------------------------
__device__ __forceinline__ cplx ref_E(double z) {
  if (fabs(z) < 1e-4) {
    const double z2 = z * z, z4 = z2 * z2;
    return {1.0 - z2 / 6.0 + z4 / 120.0, z * 0.5 - z * z2 / 24.0 + z4 * z / 720.0};
  }
  double s, c;
  sincos(z, &s, &c);
  return {s / z, (1.0 - c) / z};
}

__device__ __forceinline__ cplx ref_Ep(double z) {
  if (fabs(z) < 1e-4) {
    const double z2 = z * z, z3 = z2 * z, z4 = z2 * z2;
    return {-z / 3.0 + z3 / 30.0, 0.5 - z2 / 8.0 + z4 / 144.0};
  }
  double s, c;
  sincos(z, &s, &c);
  const double z2 = z * z;
  return {(z * c - s) / z2, (z * s - (1.0 - c)) / z2};
}

__device__ __forceinline__ cplx ref_Epp(double z) {
  const double z2 = z * z;
  if (fabs(z) < 1e-4) return {-1.0 / 3.0 + z2 / 10.0, -z / 4.0 + z2 * z / 36.0};
  double s, c;
  sincos(z, &s, &c);
  const double z3 = z2 * z;
  return {(2.0 * s - 2.0 * z * c - z2 * s) / z3, (2.0 * (1.0 - c) - 2.0 * z * s + z2 * c) / z3};
}

// Phi(a, b) = -2i (E(b) - E(a)) / (b - a); -2i E'((a+b)/2) when |b-a| < 1e-3.
__device__ __forceinline__ cplx ref_Phi(double a, double b) {
  const double d = b - a;
  if (fabs(d) < 1e-3) {
    const cplx e = ref_Ep(0.5 * (a + b));
    return {2.0 * e.y, -2.0 * e.x};  // -2i * e
  }
  const cplx eb = ref_E(b), ea = ref_E(a);
  const double inv = 1.0 / d;
  const double dx = (eb.x - ea.x) * inv, dy = (eb.y - ea.y) * inv;
  return {2.0 * dy, -2.0 * dx};
}

// ---------------------------------------------------------------------------
// J1(x)/x in FP64.  Power series sum_m (-x^2/4)^m / (2 m! (m+1)!) for x <= 12
// (terms summed until below 1e-17 of the running sum), Hankel asymptotic
// expansion J1(x) = sqrt(2/(pi x)) (P cos chi - Q sin chi), chi = x - 3pi/4,
// with P, Q summed to the smallest term beyond.  Accuracy ~1e-10 relative near
// x = 12, better elsewhere.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double j1_over_x_f64(double x) {
  const double ax = fabs(x);
  if (ax < 1e-5) {
    const double x2 = x * x;
    return 0.5 - x2 / 16.0 + x2 * x2 / 384.0;
  }
  if (ax <= 12.0) {
    const double q = 0.25 * x * x;
    double term = 0.5, sum = 0.5;
    for (int m = 1; m < 60; ++m) {
      term *= -q / ((double)m * (double)(m + 1));
      sum += term;
      if (fabs(term) < 1e-17 * fabs(sum)) break;
    }
    return sum;
  }
  // a_k(nu=1) = prod_{j=1..k} (4 - (2j-1)^2) / (k! 8^k)
  const double inv = 1.0 / ax;
  double P = 1.0, Qs = 0.0, ak = 1.0, pw = 1.0, last = 1e300;
  for (int k = 1; k < 40; ++k) {
    const double f = (4.0 - (2.0 * k - 1.0) * (2.0 * k - 1.0)) / (8.0 * k);
    ak *= f;
    pw *= inv;
    const double t = ak * pw;
    if (fabs(t) > last) break;  // asymptotic series: stop at the smallest term
    last = fabs(t);
    // k odd -> Q term with sign (-1)^((k-1)/2); k even -> P term with sign (-1)^(k/2)
    if (k & 1) Qs += ((((k - 1) >> 1) & 1) ? -t : t);
    else P += (((k >> 1) & 1) ? -t : t);
    if (last < 1e-17) break;
  }
  double s, c;
  sincos(ax - 0.75 * 3.14159265358979323846, &s, &c);
  const double j1 = sqrt(2.0 / (3.14159265358979323846 * ax)) * (P * c - Qs * s);
  return j1 / ax;
}

// Binary search of the object owning flat index i in prefix[0..N].
__device__ __forceinline__ int find_owner(const int* prefix, int N, int i) {
  int lo = 0, hi = N;  // prefix[lo] <= i < prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

}  // namespace pii
