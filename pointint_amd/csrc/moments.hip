// Low-|k| form factors from moment series (triangles; disks and Gaussians
// further below).
//
// The reference sums, per face f with area vector S_f, corners v_f0..v_f2 and
// k = kappa d (cpp/src/dof/triangle_mesh_dof.cpp:36-71):
//
//   A(k) = sum_f (d.S_f) e^{i k.v_f0} Phi(k.e1, k.e2)
//        = sum_f (d.S_f) / area_f * int_f e^{i k.x} dS.
//
// Expanding e^{i k.x} about a centre c and using, for a linear function z on
// a triangle, int_f z^n dS = area_f * 2 n! / (n+2)! * h_n(z0, z1, z2) (h_n the
// complete homogeneous symmetric polynomial of the corner values):
//
//   A(kappa d) = e^{i kappa d.c} sum_{n>=0} (i kappa)^n 2/(n+2)! H_n(d),
//   H_n(d)     = sum_f (d.S_f) h_n(w_f0, w_f1, w_f2),   w_fj = d.(v_fj - c).
//
// H_n depends on the direction only, so one pass over the faces per direction
// serves every shell of the grid with kappa R <= kSeriesReach (R bounds
// |v - c|).  Each term is bounded by Area (kappa R)^n / n!, so the tail after
// P terms is below Area X^(P+1)/(P+1)! / (1 - X/(P+2)) at X = kappa R:
// 3.6e-10 Area for P = 56, X = 15.  Each triangle object takes the smallest
// order of kSeriesOrders that meets that bound at its own largest series
// kappa R (nord; one launch, the order a uniform switch per workgroup).
// The largest term, about Area e^X / sqrt(2 pi X), keeps FP64 rounding below
// 4e-11 Area.
//
// h_n by the recurrences (all FP64, 4 operations per order and face)
//   b_n = w1^n,  h_n(w0, w1) = w0 h_{n-1}(w0, w1) + b_n,
//   h_n(w0, w1, w2) = w2 h_{n-1}(w0, w1, w2) + h_n(w0, w1).
#include "device_common.hpp"
#include "pii_internal.hpp"
#include "series_common.hpp"
#include "sh_common.hpp"

namespace pii {

namespace {

using ser::P;
using ser::kLanes;
using ser::kWaves;
using ser::kSeries;
using ser::for_chunked;
using ser::sum_waves_123;
using ser::store_split;

// Scale factors g_n and recurrence factors K_n of the disk and Gaussian
// series (see k_series_surfel).
struct SurfelRecurrence {
  double gd[P + 1], gg[P + 1];  // g_n: disk, Gaussian
  double kd[P], kg[P];          // K_n: disk, Gaussian
  constexpr SurfelRecurrence() : gd(), gg(), kd(), kg() {
    double dm1 = 1.0, d0 = 1.0, gm1 = 1.0, g0 = 1.0;  // g_{n-1}, g_n
    for (int n = 0; n <= P; ++n) {
      gd[n] = d0;
      gg[n] = g0;
      if (n == P) break;
      kd[n] = (2.0 * n + 3.0) * d0 / dm1;
      kg[n] = g0 / gm1;
      const double d1 = dm1 / ((n + 1.0) * (n + 3.0)), g1 = gm1 / (n + 1.0);
      dm1 = d0;
      d0 = d1;
      gm1 = g0;
      g0 = g1;
    }
  }
};
__device__ constexpr SurfelRecurrence kSurfel{};

// Coefficient of the series sums: triangles 2/(n+2)!, disks and Gaussians
// their g_n; each times the real factor (-1)^floor(n/2) of i^n.
enum SeriesKind : int { kKindTri = 0, kKindDisk = 1, kKindGauss = 2 };
template <int kKind>
__device__ __forceinline__ double series_coef(int n) {
  if (kKind == kKindTri) return kSeries.c[n];
  const double g = kKind == kKindDisk ? kSurfel.gd[n] : kSurfel.gg[n];
  return ((n / 2) % 2 == 0) ? g : -g;
}

// Face-split launches (few objects): each workgroup of split z writes its
// raw H_n to part[((z * nobj + obj) * nblk + blk) * (P + 1) * 64 + n * 64 +
// lane] (stride P + 1 whatever the order of the launch); a finishing kernel
// sums the splits in order and evaluates.
__device__ __forceinline__ double* split_slot(double* part, int obj, int blk) {
  return part + (((size_t)blockIdx.z * gridDim.y + obj) * gridDim.x + blk) * (size_t)(P + 1) *
                    kLanes;
}

// Sums the four waves' H_n ((H1 + H2) + H3) + H0, applies series_coef<kKind>
// and leaves the coefficients in LDS; then each
// wave evaluates the series at every fourth of the object's series shells for
// its 64 directions, reading the coefficients from LDS.  out_mode 0: A (FP64)
// to Aout[col*ldA + q]; 1: X = sqrt(w) A (FP32) to Xout[col*ldX + q] and the
// column max |X| to colmax[col].
template <int PP, int kKind>
__device__ __forceinline__ void reduce_and_evaluate(
    double (&H)[PP + 1], double (&red)[PP + 1][kLanes], int wave, int lane, int j, int jc,
    double dx, double dy, double dz, const ObjDesc& d, const double* __restrict__ krep, int Qh,
    int Mh, int out_mode, double2* __restrict__ Aout, int ldA, float2* __restrict__ Xout,
    int ldX, unsigned* __restrict__ colmax) {
  sum_waves_123<PP>(H, red, wave, lane);
  if (wave == 0)
    for_chunked<PP>([&](int n) {
      red[n][lane] = (red[n][lane] + H[n]) * series_coef<kKind>(n);
    });
  __syncthreads();
  const int nser = d.nser;
  const double dc = dx * d.ctr[0] + dy * d.ctr[1] + dz * d.ctr[2];
  float m = 0.0f;
  for (int s = wave; s < nser; s += kWaves) {
    const int q = s * Mh + jc;
    const double kap = krep[6 * (size_t)Qh + q];
    const double k2 = kap * kap;
    // even part E = sum a_{2t} kappa^{2t}, odd part O = sum a_{2t+1} kappa^{2t}
    constexpr int ne = PP / 2, no = (PP - 1) / 2;
    double E = red[2 * ne][lane], O = red[2 * no + 1][lane];
#pragma unroll
    for (int t = ne - 1; t >= 0; --t) E = fma(E, k2, red[2 * t][lane]);
#pragma unroll
    for (int t = no - 1; t >= 0; --t) O = fma(O, k2, red[2 * t + 1][lane]);
    O *= kap;
    double sn, cs;
    sincos(kap * dc, &sn, &cs);
    double ax = cs * E - sn * O, ay = sn * E + cs * O;
    if (kap < 1e-10) { ax = 0.0; ay = 0.0; }  // A(0) = 0 (triangle_mesh_dof.cpp:28)
    if (out_mode == 0) {
      if (j < Mh) Aout[(size_t)d.col * ldA + q] = make_double2(ax, ay);
    } else {
      const double sw = sqrt(krep[7 * (size_t)Qh + q]);
      const float2 x = make_float2((float)(ax * sw), (float)(ay * sw));
      if (j < Mh) {
        Xout[(size_t)d.col * ldX + q] = x;
        m = fmaxf(m, fmaxf(fabsf(x.x), fabsf(x.y)));
      }
    }
  }
  if (out_mode == 1) {
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if (lane == 0) atomicMax(colmax + d.col, __float_as_uint(m));
  }
}

// One workgroup per (object, 64 directions[, face split]) for the objects of
// series order PP (others exit): the four waves take every fourth face, their
// H_n are summed through LDS, then the waves evaluate the series at every
// fourth shell (or, split, store the sums).  out_mode 0: A (FP64) to
// Aout[col*ldA + q]; 1: X = sqrt(w) A (FP32) to Xout[col*ldX + q] and the
// column max |X| to colmax[col].
template <int PP>
__device__ __forceinline__ void series_tri_body(
    const ObjDesc& d, BlockObj bo, double* red_lds, const double* __restrict__ krep, int Qh,
    int Mh, int out_mode, double2* __restrict__ Aout, int ldA, float2* __restrict__ Xout,
    int ldX, unsigned* __restrict__ colmax, double* __restrict__ part) {
  double (&red)[PP + 1][kLanes] = *reinterpret_cast<double (*)[PP + 1][kLanes]>(red_lds);
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int j = bo.blk * kLanes + lane;
  const int jc = j < Mh ? j : Mh - 1;
  const double dx = krep[3 * (size_t)Qh + jc], dy = krep[4 * (size_t)Qh + jc],
               dz = krep[5 * (size_t)Qh + jc];
  double H[PP + 1];
#pragma unroll
  for (int n = 0; n <= PP; ++n) H[n] = 0.0;
  const auto* rec = as_const(d.mrec);
  // faces [fb, fe) of this workgroup's split (all faces when not split)
  const int chunk = (d.n + (int)gridDim.z - 1) / (int)gridDim.z;
  const int fb = (int)blockIdx.z * chunk;
  const int fe = min(d.n, fb + chunk);
  // Projections of a face record: corners relative to the centre and the
  // area vector onto this lane's direction.
  auto project = [&](const auto* r, double& w0, double& w1, double& w2, double& s) {
    w0 = dx * r[0] + dy * r[1] + dz * r[2];
    w1 = dx * r[3] + dy * r[4] + dz * r[5];
    w2 = dx * r[6] + dy * r[7] + dz * r[8];
    s = dx * r[9] + dy * r[10] + dz * r[11];
  };
  double w0 = 0.0, w1 = 0.0, w2 = 0.0, s = 0.0;
  if (fb < fe) project(rec + (size_t)kRecMoment * (fb + wave < fe ? fb + wave : fe - 1), w0, w1, w2, s);
  for (int f = fb + wave; f < fe; f += kWaves) {
    const double c0 = w0, c1 = w1, c2 = w2, cs = s;
    // the next face's record loads are issued before the sums of this face
    // and consumed after them (index clamped to the range)
    const int fn = f + kWaves < fe ? f + kWaves : fe - 1;
    const auto* rn = rec + (size_t)kRecMoment * fn;
    double t[kRecMoment];
#pragma unroll
    for (int i = 0; i < kRecMoment; ++i) t[i] = rn[i];
    // keep the loads issued here: the scheduler would otherwise sink them to
    // their use, exposing their latency
    __builtin_amdgcn_sched_barrier(0);
    double b = 1.0, hab = 1.0, habc = 1.0;
    H[0] += cs;
#pragma unroll
    for (int n = 1; n <= PP; ++n) {
      b *= c1;
      hab = fma(c0, hab, b);
      habc = fma(c2, habc, hab);
      H[n] = fma(cs, habc, H[n]);
    }
    project(t, w0, w1, w2, s);
  }
  if (part) {
    store_split<PP>(H, red, wave, lane, split_slot(part, bo.obj, bo.blk) + lane);
    return;
  }
  reduce_and_evaluate<PP, kKindTri>(H, red, wave, lane, j, jc, dx, dy, dz, d, krep, Qh, Mh,
                                    out_mode, Aout, ldA, Xout, ldX, colmax);
}

__global__ void __launch_bounds__(kLanes * kWaves) k_series_tri(
    const ObjDesc* __restrict__ objs, const double* __restrict__ krep, int Qh, int Mh,
    int out_mode, double2* __restrict__ Aout, int ldA, float2* __restrict__ Xout, int ldX,
    unsigned* __restrict__ colmax, double* __restrict__ part) {
  __shared__ double red[P + 1][kLanes];
  const BlockObj bo = xcd_block_map();
  const ObjDesc& d = objs[bo.obj];
  // uniform over the workgroup; objects with cached harmonics take k_series_sh
  if (d.nser <= 0 || d.type != PII_GEOM_TRIANGLE || d.shc) return;
#define PII_TRI_BODY(PP) \
  series_tri_body<PP>(d, bo, &red[0][0], krep, Qh, Mh, out_mode, Aout, ldA, Xout, ldX, colmax, part)
  PII_SERIES_SWITCH(PII_TRI_BODY)
#undef PII_TRI_BODY
}

// Disks and Gaussians (affine DoFs; cpp/src/dof/affine_dof.cpp:103-167).  For
// k = kappa d, with mu = d.(A n), sg = |A^T d|^2 - mu^2 and w = d.(c' - c):
//   disk:      A_e = area mu J1(x)/x * 2 e^{i kappa d.c'},  x = rho kappa sqrt(sg)
//   Gaussian:  A_e = wt mu exp(-sigma^2 kappa^2 sg / 2) e^{i kappa d.c'}
// With J1(x)/x = sum_m (-1)^m (x/2)^(2m) / (2 m! (m+1)!) and (i kappa)^(2m) =
// (-1)^m kappa^(2m), A_e = a_e e^{i kappa d.c} f_e(i kappa) with
//   disk:      a_e = area mu, f(z) = e^{w z} sum_m u^m z^(2m) / (m! (m+1)!),
//              u = rho^2 max(sg, 0) / 4,
//   Gaussian:  a_e = wt mu,   f(z) = e^{w z + v z^2},  v = sigma^2 sg / 2 (no
//              clamp, as the reference CPU path),
// so H_n = sum_e a_e c_n(e) with c_n the Taylor coefficients of f.  They obey
// three-term recurrences (from the differential equations of f:
// z f'' + (3 - 2 w z) f' + ((w^2 - 4u) z - 3 w) f = 0 for disks,
// f' = (w + 2 v z) f for Gaussians):
//   disk:      (n+1)(n+3) c_{n+1} = (2n+3) w c_n - (w^2 - 4u) c_{n-1}
//   Gaussian:  (n+1) c_{n+1}      = w c_n + 2 v c_{n-1}
// With c_n = g_n e_n and g_{n+1} = g_{n-1} / ((n+1)(n+3)) (disk) or
// g_{n-1} / (n+1) (Gaussian), g_{-1} = g_0 = 1, the coefficient of e_{n-1}
// becomes the element constant -(w^2 - 4u) or 2v:
//   e_{n+1} = K_n w e_n + G e_{n-1},  K_n = (2n+3) g_n / g_{n-1} or g_n / g_{n-1},
// four FP64 operations per order and element (two products, one FMA, the sum
// into H), and g_n is applied once to the sums.  Every term of c_n has the
// sign of w^n (u >= 0; v >= 0 but for rounding), so the forward recurrence
// has no cancellation.
template <bool kDisk>
__global__ void __launch_bounds__(kLanes * kWaves) k_series_surfel(
    const ObjDesc* __restrict__ objs, const double* __restrict__ krep, int Qh, int Mh,
    int out_mode, double2* __restrict__ Aout, int ldA, float2* __restrict__ Xout, int ldX,
    unsigned* __restrict__ colmax, double* __restrict__ part) {
  __shared__ double red[P + 1][kLanes];
  const BlockObj bo = xcd_block_map();
  const ObjDesc& d = objs[bo.obj];
  if (d.nser <= 0 || d.type != (kDisk ? PII_GEOM_DISK : PII_GEOM_GAUSSIAN)) return;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int j = bo.blk * kLanes + lane;
  const int jc = j < Mh ? j : Mh - 1;
  const double dx = krep[3 * (size_t)Qh + jc], dy = krep[4 * (size_t)Qh + jc],
               dz = krep[5 * (size_t)Qh + jc];
  const auto* p = as_const(d.dofs);
  // |A^T d|^2 (A row-major in dofs[3:12])
  const double atx = p[3] * dx + p[6] * dy + p[9] * dz;
  const double aty = p[4] * dx + p[7] * dy + p[10] * dz;
  const double atz = p[5] * dx + p[8] * dy + p[11] * dz;
  const double al = atx * atx + aty * aty + atz * atz;
  const double dc = dx * d.ctr[0] + dy * d.ctr[1] + dz * d.ctr[2];
  double H[P + 1];
#pragma unroll
  for (int n = 0; n <= P; ++n) H[n] = 0.0;
  const auto* rec = as_const(d.rec);
  const int chunk = (d.n + (int)gridDim.z - 1) / (int)gridDim.z;
  const int eb = (int)blockIdx.z * chunk;
  const int ee = min(d.n, eb + chunk);
  // element record: centre c' (0..2), A n (3..5), rho or sigma (6), area or
  // weight (7); the next element's record is loaded before this one's sums
  double t[kRecSurfel] = {};
  if (eb < ee) {
    const auto* r = rec + (size_t)kRecSurfel * min(eb + wave, ee - 1);
#pragma unroll
    for (int i = 0; i < kRecSurfel; ++i) t[i] = r[i];
  }
  for (int e = eb + wave; e < ee; e += kWaves) {
    const double mu = dx * t[3] + dy * t[4] + dz * t[5];
    const double w = dx * t[0] + dy * t[1] + dz * t[2] - dc;
    const double sg = al - mu * mu;
    const double s2 = t[6] * t[6];
    const double a = t[7] * mu;
    // G = 4u - w^2 (disk), 2v (Gaussian)
    const double G = kDisk ? fma(s2, fmax(sg, 0.0), -w * w) : s2 * sg;
    const auto* rn = rec + (size_t)kRecSurfel * min(e + kWaves, ee - 1);
#pragma unroll
    for (int i = 0; i < kRecSurfel; ++i) t[i] = rn[i];
    __builtin_amdgcn_sched_barrier(0);
    double em1 = 0.0, e0 = a;
    H[0] += e0;
#pragma unroll
    for (int n = 0; n < P; ++n) {
      const double e1 = fma(e0 * w, kDisk ? kSurfel.kd[n] : kSurfel.kg[n], G * em1);
      H[n + 1] += e1;
      em1 = e0;
      e0 = e1;
    }
  }
  if (part) {
    store_split<P>(H, red, wave, lane, split_slot(part, bo.obj, bo.blk) + lane);
    return;
  }
  reduce_and_evaluate<P, kDisk ? kKindDisk : kKindGauss>(H, red, wave, lane, j, jc, dx, dy, dz,
                                                         d, krep, Qh, Mh, out_mode, Aout, ldA,
                                                         Xout, ldX, colmax);
}

// Finishes a face-split series launch: the four waves each sum every fourth
// split's H_n for their 64 directions (splits in increasing order), then the
// usual reduction and evaluation.  type (and, for triangles, the order PP)
// selects the objects of the launch.
template <int PP, int kKind>
__device__ __forceinline__ void series_finish_body(
    const ObjDesc& d, BlockObj bo, double* red_lds, const double* __restrict__ krep, int Qh,
    int Mh, int out_mode, double2* __restrict__ Aout, int ldA, float2* __restrict__ Xout,
    int ldX, unsigned* __restrict__ colmax, const double* __restrict__ part, int nsplit) {
  double (&red)[PP + 1][kLanes] = *reinterpret_cast<double (*)[PP + 1][kLanes]>(red_lds);
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int j = bo.blk * kLanes + lane;
  const int jc = j < Mh ? j : Mh - 1;
  const double dx = krep[3 * (size_t)Qh + jc], dy = krep[4 * (size_t)Qh + jc],
               dz = krep[5 * (size_t)Qh + jc];
  double H[PP + 1];
#pragma unroll
  for (int n = 0; n <= PP; ++n) H[n] = 0.0;
  const size_t stride = (size_t)gridDim.y * gridDim.x * (P + 1) * kLanes;
  const double* src = part + ((size_t)bo.obj * gridDim.x + bo.blk) * (size_t)(P + 1) * kLanes;
  for (int z = wave; z < nsplit; z += kWaves) {
    const double* o = src + (size_t)z * stride;
    // all orders of a split in flight at once (a latency-bound sum)
#pragma unroll
    for (int n = 0; n <= PP; ++n) H[n] += o[n * kLanes + lane];
  }
  reduce_and_evaluate<PP, kKind>(H, red, wave, lane, j, jc, dx, dy, dz, d, krep, Qh, Mh,
                                 out_mode, Aout, ldA, Xout, ldX, colmax);
}

__global__ void __launch_bounds__(kLanes * kWaves) k_series_finish(
    const ObjDesc* __restrict__ objs, int type, const double* __restrict__ krep, int Qh, int Mh,
    int out_mode, double2* __restrict__ Aout, int ldA, float2* __restrict__ Xout, int ldX,
    unsigned* __restrict__ colmax, const double* __restrict__ part, int nsplit) {
  __shared__ double red[P + 1][kLanes];
  const BlockObj bo = xcd_block_map();
  const ObjDesc& d = objs[bo.obj];
  if (d.nser <= 0 || d.type != type || d.shc) return;
  if (type == PII_GEOM_TRIANGLE) {
#define PII_FIN_BODY(PP)                                                                     \
  series_finish_body<PP, kKindTri>(d, bo, &red[0][0], krep, Qh, Mh, out_mode, Aout, ldA, \
                                   Xout, ldX, colmax, part, nsplit)
    PII_SERIES_SWITCH(PII_FIN_BODY)
#undef PII_FIN_BODY
  } else if (type == PII_GEOM_DISK) {
    series_finish_body<P, kKindDisk>(d, bo, &red[0][0], krep, Qh, Mh, out_mode, Aout, ldA, Xout,
                                     ldX, colmax, part, nsplit);
  } else {
    series_finish_body<P, kKindGauss>(d, bo, &red[0][0], krep, Qh, Mh, out_mode, Aout, ldA, Xout,
                                      ldX, colmax, part, nsplit);
  }
}

// Spherical-harmonic form of the triangle series (objects with cached
// coefficients d.shc; sh_common.hpp): the rows of A = s R are s times
// orthonormal, p = A^T d / s = (p_x, p_y, x) is a unit vector, and
//   H_n(d) = s^(n+2) sum_{m, l} Q_l^m(x) (a_nlm Re z^m + b_nlm Im z^m),
//   z = p_x + i p_y,
// (corners scale by s, area vectors by s^2).
//
// The coefficients are staged through LDS and the orders n split by parity:
// a workgroup takes 128 directions; waves 0 and 1 sum the even n (harmonics
// of odd degree l), waves 2 and 3 the odd n (even l), each for its own 64
// directions over every order m.  The pairs of order m + 1 are loaded from
// global memory while the waves sum order m from LDS (double-buffered, one
// barrier per order).  The even sums give E(kappa) = sum_t c_2t H_2t kappa^2t
// and the odd ones O(kappa) = sum_t c_2t+1 H_2t+1 kappa^2t+1 at every series
// shell; the two halves meet in LDS, eight shells at a time, for the phase
// and the output.
__device__ constexpr sh::Recurrence kShRec{};
__device__ constexpr sh::Layout kShLayout{};
constexpr sh::Layout kShLay{};
constexpr int kShPairs0 = kShLay.C[sh::kL + 1];   // pairs of the block m = 0
constexpr int kShThreads = kLanes * kWaves;
constexpr int kShDirs = kShThreads / 2;           // directions per workgroup
constexpr int kShStage = (kShPairs0 + kShThreads - 1) / kShThreads;
constexpr int kShH = P / 2 + 1;                   // sums per parity (29 even, 28 odd)
constexpr int kShChunk = 8;                       // shells per evaluation round

// Degree l of order m for the sums of parity par (n = 2 i + par in H[i]; the
// degrees of the other parity only advance the recurrence).
template <int l>
__device__ __forceinline__ void sh_degree_par(double (&H)[kShH], const double2* cb,
                                              const double2* ab, int cm, int m, int par,
                                              double x, double zr, double zi, double seed,
                                              double& qa, double& qb) {
  if (l >= m) {
    const double2 r = ab[l];  // (a_lm, b_lm)
    const double q = l == m ? seed : r.x * fma(x, qa, -r.y * qb);
    qb = qa;
    qa = q;
    if (((l + 1) & 1) == par) {
      const double u = q * zr, v = q * zi;
      constexpr int c0 = kShLay.C[l];
#pragma unroll
      for (int k = 0; k < sh::count(l); ++k) {
        const int n = sh::n0(l) + 2 * k;
        const double2 c = cb[c0 + k - cm];
        H[n >> 1] = fma(c.x, u, fma(c.y, v, H[n >> 1]));
      }
    }
  }
  if constexpr (l < sh::kL) sh_degree_par<l + 1>(H, cb, ab, cm, m, par, x, zr, zi, seed, qa, qb);
}

// Four waves per SIMD: 31 KB of LDS per workgroup (the shells' even and odd
// halves reuse the coefficient buffers once every order is summed) and
// registers held to that.
__global__ void __launch_bounds__(kShThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_series_sh(
    const ObjDesc* __restrict__ objs, const double* __restrict__ krep, int Qh, int Mh,
    int out_mode, double2* __restrict__ Aout, int ldA, float2* __restrict__ Xout, int ldX,
    unsigned* __restrict__ colmax) {
  __shared__ double2 cbuf[2][kShPairs0];
  __shared__ double2 abuf[2][sh::kL + 1];  // recurrence factors (a_lm, b_lm) of the order
  __shared__ double dcs[kShDirs];
  static_assert(sizeof(double) * kShChunk * 2 * kShDirs <= sizeof(cbuf),
                "the shells' halves fit in the coefficient buffers");
  // after the order loop (whose last barrier ends the reads of cbuf)
  double (*eo)[2][kShDirs] = reinterpret_cast<double (*)[2][kShDirs]>(&cbuf[0][0]);
  const BlockObj bo = xcd_block_map();
  const ObjDesc& d = objs[bo.obj];
  if (d.nser <= 0 || d.type != PII_GEOM_TRIANGLE || !d.shc) return;
  const int tid = threadIdx.x;
  const int par = __builtin_amdgcn_readfirstlane(tid >> 7);  // 0: even n, 1: odd n
  const int dl = tid & (kShDirs - 1);
  const int j = bo.blk * kShDirs + dl;
  const int jc = j < Mh ? j : Mh - 1;
  const double dx = krep[3 * (size_t)Qh + jc], dy = krep[4 * (size_t)Qh + jc],
               dz = krep[5 * (size_t)Qh + jc];
  if (par == 0) dcs[dl] = dx * d.ctr[0] + dy * d.ctr[1] + dz * d.ctr[2];
  const auto* a = as_const(d.dofs);
  const double sc = d.shs, inv = 1.0 / sc;
  const double px = (a[3] * dx + a[6] * dy + a[9] * dz) * inv;
  const double py = (a[4] * dx + a[7] * dy + a[10] * dz) * inv;
  const double x = (a[5] * dx + a[8] * dy + a[11] * dz) * inv;
  double H[kShH];
#pragma unroll
  for (int i = 0; i < kShH; ++i) H[i] = 0.0;
  // global (not flat) loads: a flat load in flight would make every LDS read
  // of the sums below wait for all outstanding LDS reads
  const auto* src = as_global(d.shc);
  auto pair = [&](const PII_GLOBAL_AS double* g, int i) {
    return make_double2(g[2 * i], g[2 * i + 1]);
  };
  // staging of one order's pairs: every thread up to kShStage of them
  double2 r0 = {}, r1 = {}, r2 = {}, r3 = {}, rab = {};
  static_assert(kShStage == 4, "staging registers");
  static_assert(sh::kL + 1 <= kShThreads, "one thread per recurrence factor");
  auto fetch = [&](int m) {
    const int last = kShLayout.C[sh::kL + 1] - kShLayout.C[m] - 1;
    const auto* g = src + 2 * (size_t)kShLayout.mstart[m];
    r0 = pair(g, min(tid, last));
    r1 = pair(g, min(tid + kShThreads, last));
    r2 = pair(g, min(tid + 2 * kShThreads, last));
    r3 = pair(g, min(tid + 3 * kShThreads, last));
    const int l = min(tid, sh::kL);
    rab = make_double2(kShRec.a[l][m], kShRec.b[l][m]);
  };
  auto put = [&](int b, int m) {
    const int np = kShLayout.C[sh::kL + 1] - kShLayout.C[m];
    if (tid < np) cbuf[b][tid] = r0;
    if (tid + kShThreads < np) cbuf[b][tid + kShThreads] = r1;
    if (tid + 2 * kShThreads < np) cbuf[b][tid + 2 * kShThreads] = r2;
    if (tid + 3 * kShThreads < np) cbuf[b][tid + 3 * kShThreads] = r3;
    if (tid <= sh::kL) abuf[b][tid] = rab;
  };
  fetch(0);
  put(0, 0);
  __syncthreads();
  double zr = 1.0, zi = 0.0;  // z^m
  for (int m = 0; m <= sh::kL; ++m) {
    const int cur = m & 1;
    if (m < sh::kL) fetch(m + 1);
    double qa = 0.0, qb = 0.0;
    sh_degree_par<0>(H, &cbuf[cur][0], &abuf[cur][0], kShLayout.C[m], m, par, x, zr, zi,
                     kShRec.seed[m], qa, qb);
    if (m < sh::kL) put(cur ^ 1, m + 1);
    __syncthreads();
    const double t = zr * px - zi * py;
    zi = zr * py + zi * px;
    zr = t;
  }
  // series coefficients s^(n+2) c_n H_n, n = 2 i + par
  {
    double f = sc * sc;
    if (par) f *= sc;
    const double s2 = sc * sc;
#pragma unroll
    for (int i = 0; i < kShH; ++i) {
      H[i] *= f * series_coef<kKindTri>(2 * i + (i < kShH - 1 ? par : 0));
      f *= s2;
    }
  }
  const int nser = d.nser;
  float mx = 0.0f;
  for (int s0 = 0; s0 < nser; s0 += kShChunk) {
    // this wave's half (E, or O) at up to kShChunk shells
    for (int t = 0; t < kShChunk && s0 + t < nser; ++t) {
      const double kap = krep[6 * (size_t)Qh + (size_t)(s0 + t) * Mh + jc];
      const double k2 = kap * kap;
      double v;
      if (par == 0) {
        v = H[kShH - 1];
#pragma unroll
        for (int i = kShH - 2; i >= 0; --i) v = fma(v, k2, H[i]);
      } else {
        v = H[kShH - 2];
#pragma unroll
        for (int i = kShH - 3; i >= 0; --i) v = fma(v, k2, H[i]);
        v *= kap;
      }
      eo[t][par][dl] = v;
    }
    __syncthreads();
    for (int it = tid; it < kShChunk * kShDirs; it += kShThreads) {
      const int t = it / kShDirs, jl = it - t * kShDirs;
      const int jj = bo.blk * kShDirs + jl;
      if (s0 + t >= nser || jj >= Mh) continue;
      const int q = (s0 + t) * Mh + jj;
      const double kap = krep[6 * (size_t)Qh + q];
      const double E = eo[t][0][jl], O = eo[t][1][jl];
      double sn, cs;
      sincos(kap * dcs[jl], &sn, &cs);
      double ax = cs * E - sn * O, ay = sn * E + cs * O;
      if (kap < 1e-10) { ax = 0.0; ay = 0.0; }  // A(0) = 0 (triangle_mesh_dof.cpp:28)
      if (out_mode == 0) {
        Aout[(size_t)d.col * ldA + q] = make_double2(ax, ay);
      } else {
        const double sw = sqrt(krep[7 * (size_t)Qh + q]);
        const float2 xv = make_float2((float)(ax * sw), (float)(ay * sw));
        Xout[(size_t)d.col * ldX + q] = xv;
        mx = fmaxf(mx, fmaxf(fabsf(xv.x), fabsf(xv.y)));
      }
    }
    __syncthreads();
  }
  if (out_mode == 1) {
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if ((tid & (kLanes - 1)) == 0) atomicMax(colmax + d.col, __float_as_uint(mx));
  }
}

}  // namespace

int series_tri_slots() {
  static std::atomic<int> cache[16];
  return resident_workgroups(k_series_tri, kLanes * kWaves, cache);
}

int series_surfel_slots() {
  static std::atomic<int> cache[16];
  return resident_workgroups(k_series_surfel<true>, kLanes * kWaves, cache);
}

size_t series_split_doubles(int N, int Mh, int nsplit) {
  return nsplit > 1 ? (size_t)nsplit * N * ((Mh + kLanes - 1) / kLanes) * (P + 1) * kLanes : 0;
}

void launch_series_surfel(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_krep,
                          int Qh, int Mh, int out_mode, double2* d_A, int ldA, float2* d_X,
                          int ldX, unsigned* d_colmax, int nsplit, double* d_part) {
  if (N <= 0 || Mh <= 0) return;
  const bool split = nsplit > 1 && d_part;
  const dim3 grid((Mh + kLanes - 1) / kLanes, N, split ? nsplit : 1);
  const dim3 grid2((Mh + kLanes - 1) / kLanes, N);
  double* part = split ? d_part : nullptr;
  hipLaunchKernelGGL(k_series_surfel<true>, grid, dim3(kLanes * kWaves), 0, s, d_objs, d_krep, Qh,
                     Mh, out_mode, d_A, ldA, d_X, ldX, d_colmax, part);
  if (split)
    hipLaunchKernelGGL(k_series_finish, grid2, dim3(kLanes * kWaves), 0, s, d_objs,
                       (int)PII_GEOM_DISK, d_krep, Qh, Mh, out_mode, d_A, ldA, d_X, ldX, d_colmax,
                       part, nsplit);
  hipLaunchKernelGGL(k_series_surfel<false>, grid, dim3(kLanes * kWaves), 0, s, d_objs, d_krep,
                     Qh, Mh, out_mode, d_A, ldA, d_X, ldX, d_colmax, part);
  if (split)
    hipLaunchKernelGGL(k_series_finish, grid2, dim3(kLanes * kWaves), 0, s, d_objs,
                       (int)PII_GEOM_GAUSSIAN, d_krep, Qh, Mh, out_mode, d_A, ldA, d_X, ldX,
                       d_colmax, part, nsplit);
}

void launch_series_sh(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_krep, int Qh,
                      int Mh, int out_mode, double2* d_A, int ldA, float2* d_X, int ldX,
                      unsigned* d_colmax) {
  if (N <= 0 || Mh <= 0) return;
  hipLaunchKernelGGL(k_series_sh, dim3((Mh + kShDirs - 1) / kShDirs, N), dim3(kShThreads), 0, s,
                     d_objs, d_krep, Qh, Mh, out_mode, d_A, ldA, d_X, ldX, d_colmax);
}

void launch_series_tri_raw(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_dirs,
                           int Qd, int Md, double* d_H) {
  if (N <= 0 || Md <= 0) return;
  hipLaunchKernelGGL(k_series_tri, dim3((Md + kLanes - 1) / kLanes, N, 1), dim3(kLanes * kWaves),
                     0, s, d_objs, d_dirs, Qd, Md, 0, nullptr, 0, nullptr, 0, nullptr, d_H);
}

void launch_series_tri(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_krep, int Qh,
                       int Mh, int out_mode, double2* d_A, int ldA, float2* d_X, int ldX,
                       unsigned* d_colmax, int nsplit, double* d_part) {
  if (N <= 0 || Mh <= 0) return;
  const bool split = nsplit > 1 && d_part;
  double* part = split ? d_part : nullptr;
  hipLaunchKernelGGL(k_series_tri, dim3((Mh + kLanes - 1) / kLanes, N, split ? nsplit : 1),
                     dim3(kLanes * kWaves), 0, s, d_objs, d_krep, Qh, Mh, out_mode, d_A, ldA, d_X,
                     ldX, d_colmax, part);
  if (split)
    hipLaunchKernelGGL(k_series_finish, dim3((Mh + kLanes - 1) / kLanes, N),
                       dim3(kLanes * kWaves), 0, s, d_objs, (int)PII_GEOM_TRIANGLE, d_krep, Qh,
                       Mh, out_mode, d_A, ldA, d_X, ldX, d_colmax, part, nsplit);
}

}  // namespace pii
